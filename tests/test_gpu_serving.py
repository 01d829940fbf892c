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
    bd = b.cpu().numpy().astype(np.complex128)  # [frames = 1, N, N, N]
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


def test_served_cfg5_sizes_alone_vs_mix(S):
    """Requests at the serving workload's sizes (96^3 x 8, 128^3 x 32, 160^3 x 16 coils,
    2N samples per spoke) served together: each image equals the one it gets alone, bit
    for bit, and the 96^3 one equals the device CG on the same inputs."""
    from hastycompute_amd import native

    reqs = lambda: [S.ReconRequest(rid=1, n=96, ncoils=8, nspokes=1000, iters=4, seed=31),  # noqa: E731
                    S.ReconRequest(rid=2, n=128, ncoils=32, nspokes=1200, iters=3, seed=32, arrival=0.5),
                    S.ReconRequest(rid=3, n=160, ncoils=16, nspokes=1000, iters=4, seed=33, arrival=0.2)]
    srv = S.ReconServer()
    prepared = [srv.setup(r) for r in reqs()]
    alone = []
    for r, p in zip(reqs(), prepared):
        r.prepared = p
        S.ReconServer().run([r], realtime=False)
        alone.append(r.image)
    mix = reqs()
    for r, p in zip(mix, prepared):
        r.prepared = p
    stats = S.ReconServer().run(mix, realtime=False)
    assert stats["requests"] == 3 and stats["max_active"] >= 2
    for r, a in zip(mix, alone):
        print(f"N={r.n} C={r.ncoils}: alone == mix {torch.equal(r.image, a)}", flush=True)
        assert torch.equal(r.image, a)
    # the 96^3 request against the device CG on its own operator and right-hand side
    smaps, kern, b = prepared[0]
    op = native.SenseNormalToeplitz(smaps, kern[0].reshape(192, 192, 192), coil_batch=8)
    x, _ = native.cg_solve(lambda v, o: op.apply(v, out=o), b, torch.zeros_like(b), 4)
    e = float((x[0].cpu() - mix[0].image).norm() / x[0].cpu().norm())
    print(f"96^3 served vs device CG: {e:.2e}", flush=True)
    assert e < 1e-5


def test_llr_request_is_cg_then_prox(S):
    """A 3-frame request with the LLR prox: its image is the frames' CG images followed by
    one OldNormal3DBlocksSVT pass (16^3 blocks), which matches the oracle's block SVT."""
    from hastycompute_amd import native

    R = lambda thr: S.ReconRequest(rid=5, n=48, ncoils=4, nspokes=600, iters=5, seed=41, frames=3,  # noqa: E731
                                   llr_thresh=thr)
    srv = S.ReconServer()
    prepared = srv.setup(R(None))
    plain, llr = R(None), R(0.5)
    plain.prepared = llr.prepared = prepared
    S.ReconServer().run([plain], realtime=False)
    S.ReconServer().run([llr], realtime=False)
    assert tuple(plain.image.shape) == (3, 48, 48, 48) and tuple(llr.image.shape) == (3, 48, 48, 48)
    x = plain.image.reshape(3, 1, 48, 48, 48).cuda().contiguous()
    native.llr_normal_blocks(x, (16, 16, 16), (16, 16, 16), 1, 0.5, True)
    assert torch.equal(x.reshape(3, 48, 48, 48).cpu(), llr.image)
    ref = O.normal_3d_blocks_svt(plain.image.reshape(3, 1, 48, 48, 48).numpy().astype(np.complex128),
                                 [16, 16, 16], [16, 16, 16], 1, 0.5, True)
    e = O.rel_l2(llr.image.numpy().reshape(ref.shape), ref)
    changed = O.rel_l2(llr.image.numpy(), plain.image.numpy())
    print(f"LLR request vs oracle block SVT: {e:.2e} (prox changed the frames by {changed:.2e})", flush=True)
    assert e < 1e-4 and changed > 1e-3


def test_async_admission_counts_requests_in_setup(S):
    """Real-time serving with asynchronous setup: a burst of simultaneous arrivals is
    admitted against the HBM in use plus the footprints of requests still in setup."""
    reqs = [S.ReconRequest(rid=i, n=64, ncoils=8, nspokes=400, iters=3, seed=60 + i) for i in range(6)]
    base = torch.cuda.memory_allocated()
    budget = base + (3 << 29)
    srv = S.ReconServer(hbm_budget=budget)
    stats = srv.run(reqs, realtime=True)
    print(f"burst: admissions {[(r, round((u + f - base) / 2**20), a) for r, u, f, a in srv.admissions]}, "
          f"max active {stats['max_active']}", flush=True)
    assert stats["requests"] == 6
    for rid, used, fp, active in srv.admissions:
        assert active == 0 or used + fp <= budget
    assert stats["max_active"] < 6
