"""GPU tests of the continuous-batching server (hastycompute_amd.serving): a request's
image is the same alone and in a mix of requests of other sizes and coil counts that
join and leave its batch; the image matches the FP64 oracle's CG on the explicit SENSE
normal operator; admission respects the HBM budget."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import serving

    return serving


def test_alone_vs_mix_identical(S):
    """Same inputs (prepared once: the setup's NUFFT spreading sums with atomics), same
    image bit for bit whether the request runs alone or batched with others."""
    srv = S.ReconServer()
    R = lambda: S.ReconRequest(rid=100, n=64, ncoils=6, nspokes=700, iters=6, seed=5)  # noqa: E731
    prepared = srv.setup(R())
    r_alone = R()
    r_alone.prepared = prepared
    S.ReconServer().run([r_alone], realtime=False)
    mix = [S.ReconRequest(rid=1, n=64, ncoils=8, nspokes=500, iters=9, seed=1),
           S.ReconRequest(rid=2, n=32, ncoils=4, nspokes=300, iters=4, seed=2),
           S.ReconRequest(rid=3, n=64, ncoils=3, nspokes=900, iters=3, seed=3, arrival=0.5),
           S.ReconRequest(rid=4, n=48, ncoils=5, nspokes=400, iters=7, seed=4, arrival=0.7)]
    r_mix = R()
    r_mix.prepared = prepared
    r_mix.arrival = 0.6
    stats = S.ReconServer().run(mix[:3] + [r_mix] + mix[3:], realtime=False)
    d = float((r_mix.image - r_alone.image).abs().max())
    print(f"alone vs mix: max |diff| {d:.3e}, mix mean batch {stats['mean_batch']:.2f}, "
          f"max active {stats['max_active']}", flush=True)
    assert stats["max_active"] >= 3 and stats["requests"] == 5
    assert torch.equal(r_mix.image, r_alone.image)
    # without prepared inputs (the setup rerun) the image agrees to the float rounding of
    # the setup's atomic sums
    r_fresh = R()
    S.ReconServer().run([r_fresh], realtime=False)
    rel = float((r_fresh.image - r_alone.image).norm() / r_alone.image.norm())
    print(f"fresh setup vs prepared: {rel:.2e}", flush=True)
    assert rel < 1e-5


def test_server_image_matches_oracle_cg(S):
    """The served image equals the oracle's CG (same right-hand side) on the explicit
    SENSE normal operator of the request's trajectory and coil maps."""
    from hastycompute_amd import synthetic

    req = S.ReconRequest(rid=7, n=32, ncoils=4, nspokes=400, iters=4, seed=9)
    srv = S.ReconServer()
    req.prepared = srv.setup(req)
    smaps, _, b = req.prepared
    srv.run([req], realtime=False)
    coords = synthetic.kooshball_torch(req.nspokes, 64, (req.seed * 7919) % 1000003, "cuda").cpu().numpy()
    Sd = smaps.cpu().numpy().astype(np.complex128)
    bd = b.cpu().numpy().astype(np.complex128)[None]
    ref = O.conjugate_gradient(lambda v: O.sense_normal_apply(coords.astype(np.float64), v, Sd, list(range(4))), bd,
                               np.zeros_like(bd), req.iters)
    e = O.rel_l2(req.image.numpy()[None].astype(np.complex128), ref)
    print(f"served image vs oracle CG ({req.iters} iterations): {e:.2e}", flush=True)
    assert e < 1e-4


def test_admission_respects_budget(S):
    reqs = [S.ReconRequest(rid=i, n=64, ncoils=8, nspokes=400, iters=3, seed=i) for i in range(5)]
    base = torch.cuda.memory_allocated()
    budget = base + (1 << 30)
    srv = S.ReconServer(hbm_budget=budget)
    stats = srv.run(reqs, realtime=False)
    print(f"budget run: max active {stats['max_active']}, ticks {stats['ticks']}, "
          f"admissions {[(r, round((u + f - base) / 2**20), a) for r, u, f, a in srv.admissions]}", flush=True)
    assert stats["requests"] == 5
    for rid, used, fp, active in srv.admissions:
        assert active == 0 or used + fp <= budget
    assert stats["max_active"] < 5
