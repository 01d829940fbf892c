"""GPU parity of the TorchScript operator layer (torch.classes.Hasty*, through the
hastycompute_amd.ffi shims that mirror hastypy.ffi) against the FP64 oracle.

Bounds (relative L2): NUFFT 1e-5 at tol 1e-6 (1e-8 in FP64 at tol 1e-10), 1e-4 at the
reference's default tol 1e-5; SENSE operators 1e-4 at the default tol; the batched
Toeplitz normal operator 1e-5; block SVT 1e-4.
"""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd.ffi import hasty_interface, hasty_nufft, hasty_sense, hasty_svt

    return SimpleNamespace(nufft=hasty_nufft, sense=hasty_sense, svt=hasty_svt, itf=hasty_interface)


def rnd(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def g(a, dtype=torch.complex64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def h(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128)


def coords32(rng, ndim, M):
    co = rng.uniform(-np.pi, np.pi, (ndim, M)).astype(np.float32)
    return co, co.astype(np.float64)


@pytest.mark.parametrize("fp64", [False, True])
def test_ffi_nufft_types_vs_direct_sum(F, fp64):
    rng = np.random.default_rng(11 + fp64)
    nmodes = [2, 12, 10, 8]
    M = 600
    if fp64:
        co = rng.uniform(-np.pi, np.pi, (3, M))
        cog, co_used, ct, tol, bound = g(co, torch.float64), co, torch.complex128, 1e-10, 1e-8
    else:
        co, co_used = coords32(rng, 3, M)
        cog, ct, tol, bound = g(co, torch.float32), torch.complex64, 1e-6, 1e-5
    f = rnd(rng, nmodes)
    c = rnd(rng, (2, M))
    n2 = F.nufft.Nufft(cog, nmodes, F.nufft.NufftOptions(2, False, tol))
    n1 = F.nufft.Nufft(cog, nmodes, F.nufft.NufftOptions(1, True, tol))
    k = torch.empty((2, M), dtype=ct, device="cuda")
    im = torch.empty(nmodes, dtype=ct, device="cuda")
    n2.apply(g(f, ct), k)
    n1.apply(g(c, ct), im)
    e2 = O.rel_l2(h(k), O.nufft_apply(co_used, nmodes, 2, f, exact=True))
    e1 = O.rel_l2(h(im), O.nufft_apply(co_used, nmodes, 1, c, exact=True))
    print(f"ffi Nufft fp64={fp64}: type2 {e2:.2e} type1 {e1:.2e}", flush=True)
    assert e2 < bound and e1 < bound
    if not fp64:
        # the reference's default options (tol 1e-5)
        n2d = F.nufft.Nufft(cog, nmodes, F.nufft.NufftOptions.type2())
        n2d.apply(g(f, ct), k)
        assert O.rel_l2(h(k), O.nufft_apply(co_used, nmodes, 2, f, exact=True)) < 1e-4


def test_ffi_nufft_checks(F):
    rng = np.random.default_rng(2)
    co, _ = coords32(rng, 3, 100)
    n2 = F.nufft.Nufft(g(co, torch.float32), [1, 8, 8, 8], F.nufft.NufftOptions.type2())
    with pytest.raises(RuntimeError, match="in.sizes"):
        n2.apply(torch.zeros((1, 8, 8, 4), dtype=torch.complex64, device="cuda"),
                 torch.zeros((1, 100), dtype=torch.complex64, device="cuda"))
    with pytest.raises(RuntimeError, match="complex double"):
        n2.apply(torch.zeros((1, 8, 8, 8), dtype=torch.complex128, device="cuda"),
                 torch.zeros((1, 100), dtype=torch.complex128, device="cuda"))
    with pytest.raises(RuntimeError, match="Type 3"):
        F.nufft.Nufft(g(co, torch.float32), [1, 8, 8, 8], F.nufft.NufftOptions(3))


def test_ffi_nufft_normal_with_function_lambda(F):
    rng = np.random.default_rng(5)
    nmodes = [1, 14, 12, 10]
    M = 900
    co, co_used = coords32(rng, 3, M)
    w = rng.uniform(0.2, 1.5, M)
    script = """
def apply(a: List[Tensor], b: List[Tensor]):
    for i in range(len(a)):
        a[i].mul_(b[0])
"""
    fl = F.itf.FunctionLambda(script, "apply", [g(w)])
    op = F.nufft.NufftNormal(g(co, torch.float32), nmodes, F.nufft.NufftOptions(2, False, 1e-6),
                             F.nufft.NufftOptions(1, True, 1e-6))
    x = rnd(rng, nmodes)
    out = torch.empty(nmodes, dtype=torch.complex64, device="cuda")
    st = torch.empty((1, M), dtype=torch.complex64, device="cuda")
    op.apply(g(x), out, st, fl)
    ref = O.nufft_normal_apply(co_used, nmodes, x, func_between=lambda k: k * w[None], exact=True)
    e = O.rel_l2(h(out), ref)
    op.apply(g(x), out, st, None)
    e0 = O.rel_l2(h(out), O.nufft_normal_apply(co_used, nmodes, x, exact=True))
    print(f"ffi NufftNormal: weighted {e:.2e} plain {e0:.2e}", flush=True)
    assert e < 1e-5 and e0 < 1e-5


@pytest.mark.parametrize("coils", [[0, 1, 2, 3], [2, 0, 3]])
def test_ffi_sense_family(F, coils):
    rng = np.random.default_rng(len(coils))
    img = [16, 14, 12]
    M = 1500
    co, co_used = coords32(rng, 3, M)
    S = rnd(rng, [4] + img) * 0.5
    x = rnd(rng, [1] + img)
    y = rnd(rng, (len(coils), M))
    cog, Sg = g(co, torch.float32), g(S)
    nm = [1] + img

    out = torch.empty((len(coils), M), dtype=torch.complex64, device="cuda")
    F.sense.Sense(cog, nm).apply(g(x), out, Sg, coils)
    e_f = O.rel_l2(h(out), O.sense_apply(co_used, x, S, coils, exact=True))

    outa = torch.empty(nm, dtype=torch.complex64, device="cuda")
    F.sense.SenseAdjoint(cog, nm).apply(g(y), outa, Sg, coils)
    e_a = O.rel_l2(h(outa), O.sense_adjoint_apply(co_used, y, S, coils, exact=True))

    outn = torch.empty(nm, dtype=torch.complex64, device="cuda")
    F.sense.SenseNormal(cog, nm).apply(g(x), outn, Sg, coils)
    e_n = O.rel_l2(h(outn), O.sense_normal_apply(co_used, x, S, coils, exact=True))
    print(f"ffi Sense coils={coils}: fwd {e_f:.2e} adj {e_a:.2e} normal {e_n:.2e}", flush=True)
    assert e_f < 1e-4 and e_a < 1e-4 and e_n < 1e-4


@pytest.mark.parametrize("coils", [[0, 1, 2, 3], [2, 0, 3]])
def test_ffi_sense_nop_vs_oracle(F, coils):
    """op::SenseNOp (mriop.cpp:156-266): normal, forward and backward maps, the list form
    and the adjoint, against the oracle's SENSE operators."""
    from hastycompute_amd.ffi import hasty_nufft, hasty_op

    rng = np.random.default_rng(7 + len(coils))
    img = [16, 14, 12]
    M = 1400
    co, co_used = coords32(rng, 3, M)
    S = rnd(rng, [4] + img) * 0.5
    x = rnd(rng, [1] + img)
    y = rnd(rng, (len(coils), M))
    op = hasty_op.SenseNOp(g(co, torch.float32), [1] + img, g(S), coils,
                           hasty_nufft.NufftOptions.type2(), hasty_nufft.NufftOptions.type2())
    e_n = O.rel_l2(h(op.apply(g(x))), O.sense_normal_apply(co_used, x, S, coils, exact=True))
    fwd = op.apply_forward(g(x))
    assert tuple(fwd.shape) == (len(coils), M)
    # row i is coil coils[i] (sense.ixx:581-620 multiplies smaps[coils[idx]]); the oracle's
    # sense_apply follows CUDASense's list-position maps, so it gets the maps in list order
    e_f = O.rel_l2(h(fwd), O.sense_apply(co_used, x, S[coils], list(range(len(coils))), exact=True))
    bwd = op.apply_backward(g(y))
    assert tuple(bwd.shape) == tuple([1] + img)
    # the backward NUFFT is type 1 with the + sign even when type-2 options are passed
    e_b = O.rel_l2(h(bwd), O.sense_adjoint_apply(co_used, y, S, coils, exact=True))
    x2 = rnd(rng, [1] + img)
    lst = op.apply([g(x), g(x2)])
    e_l = max(O.rel_l2(h(lst[0]), O.sense_normal_apply(co_used, x, S, coils, exact=True)),
              O.rel_l2(h(lst[1]), O.sense_normal_apply(co_used, x2, S, coils, exact=True)))
    e_adj = O.rel_l2(h(op.adjoint().apply(g(x2))), h(lst[1]))
    print(f"ffi SenseNOp coils={coils}: normal {e_n:.2e} fwd {e_f:.2e} bwd {e_b:.2e} list {e_l:.2e} "
          f"adjoint {e_adj:.2e}", flush=True)
    assert e_n < 1e-4 and e_f < 1e-4 and e_b < 1e-4 and e_l < 1e-4 and e_adj < 1e-6


def _frames(rng, nframes, E, img, C, M):
    coords, kd, w = [], [], []
    for f in range(nframes):
        co, co_used = coords32(rng, 3, M + 37 * f)
        coords.append((co, co_used))
        kd.append(rnd(rng, (E, C, M + 37 * f)) * 0.1)
        w.append(rng.uniform(0.3, 1.2, M + 37 * f))
    return coords, kd, w


@pytest.mark.parametrize("img,on_gpu", [((12, 16, 10), False), ((32, 32, 32), False), ((32, 32, 32), True)])
def test_ffi_batched_sense_normal_vs_oracle(F, img, on_gpu):
    """BatchedSenseNormal with kdata and weights: out = A^H W (A x - y) per (frame,
    encoding) through the Toeplitz-embedded operator, against the explicit oracle."""
    rng = np.random.default_rng(sum(img) + on_gpu)
    nframes, E, C, M = 2, 2, 4, 1200
    coils = [[0, 1, 2, 3], [3, 1, 0, 2]]
    frames, kd, w = _frames(rng, nframes, E, img, C, M)
    S = np.ascontiguousarray(O.coil_maps(C, max(img), seed=3)[:, :img[0], :img[1], :img[2]])
    x = rnd(rng, (nframes, E) + img)
    dev = (lambda a, dt=torch.complex64: g(a, dt)) if on_gpu else (lambda a, dt=torch.complex64: torch.as_tensor(a).to(dt))
    op = F.sense.BatchedSenseNormal([dev(c[0], torch.float32) for c in frames], dev(S),
                                    [dev(k) for k in kd], [dev(ww, torch.float32) for ww in w])
    out = dev(np.zeros_like(x))
    op.apply(dev(x), out, coils)
    got = out.cpu().numpy().astype(np.complex128)
    errs = []
    for f in range(nframes):
        for e in range(E):
            ref = O.sense_normal_apply(frames[f][1], x[f, e][None], S, coils[f], kdata=kd[f][e], weights=w[f])
            errs.append(O.rel_l2(got[f, e][None], ref))
    print(f"ffi BatchedSenseNormal img={img} gpu={on_gpu}: max {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-5


def test_ffi_batched_sense_and_adjoint_vs_oracle(F):
    rng = np.random.default_rng(21)
    img = (12, 10, 14)
    nframes, E, C, M = 2, 2, 3, 800
    frames, kd, w = _frames(rng, nframes, E, img, C, M)
    S = rnd(rng, (C,) + img) * 0.4
    coils = [[0, 1, 2], [0, 1, 2]]
    x = rnd(rng, (nframes, E) + img)
    cpu = lambda a, dt=torch.complex64: torch.as_tensor(a).to(dt)  # noqa: E731
    cl = [cpu(c[0], torch.float32) for c in frames]
    fwd = F.sense.BatchedSense(cl, cpu(S), [cpu(k) for k in kd], [cpu(ww, torch.float32) for ww in w])
    outs = [torch.zeros((E, C, kd[f].shape[2]), dtype=torch.complex64) for f in range(nframes)]
    fwd.apply(cpu(x), outs, coils)
    ys = [rnd(rng, kd[f].shape) for f in range(nframes)]
    adj = F.sense.BatchedSenseAdjoint(cl, cpu(S), [cpu(k) for k in kd], [cpu(ww, torch.float32) for ww in w])
    outa = torch.zeros((nframes, E) + img, dtype=torch.complex64)
    adj.apply([cpu(yy) for yy in ys], outa, coils)
    ef, ea = [], []
    for f in range(nframes):
        for e in range(E):
            ref = (O.sense_apply(frames[f][1], x[f, e][None], S, coils[f], exact=True) - kd[f][e]) * w[f][None]
            ef.append(O.rel_l2(outs[f][e].numpy().astype(np.complex128), ref))
            refa = O.sense_adjoint_apply(frames[f][1], (ys[f][e] - kd[f][e]) * w[f][None], S, coils[f], exact=True)
            ea.append(O.rel_l2(outa[f, e].numpy().astype(np.complex128)[None], refa))
    print(f"ffi BatchedSense fwd max {max(ef):.2e} adj max {max(ea):.2e}", flush=True)
    assert max(ef) < 1e-4 and max(ea) < 1e-4


@pytest.mark.parametrize("soft,on_gpu", [(True, False), (False, True)])
def test_ffi_normal3d_blocks_svt(F, soft, on_gpu):
    rng = np.random.default_rng(7 + soft)
    Fr, E, X = 6, 2, 16
    base = rnd(rng, (Fr, 3)) @ rnd(rng, (3, E * X ** 3))
    x = (base.reshape(Fr, E, X, X, X) + 0.05 * rnd(rng, (Fr, E, X, X, X))).astype(np.complex64)
    blk = (4, 4, 4)
    A = O.extract_block(x.astype(np.complex128), (0, 0, 0), blk)
    thresh = 0.05 * np.linalg.svd(A, compute_uv=False)[0] if soft else 2
    ref = O.normal_3d_blocks_svt(x.astype(np.complex128), blk, blk, 2, thresh, soft)
    t = g(x) if on_gpu else torch.as_tensor(x)
    F.svt.Normal3DBlocksSVT(None).apply(t, list(blk), list(blk), 2, float(thresh), soft)
    e = O.rel_l2(t.cpu().numpy().astype(np.complex128), ref)
    print(f"ffi Normal3DBlocksSVT soft={soft} gpu={on_gpu}: {e:.2e}", flush=True)
    assert e < 1e-4


def test_ffi_random3d_blocks_svt(F):
    rng = np.random.default_rng(9)
    x = rnd(rng, (5, 2, 12, 12, 12)).astype(np.complex64)
    t = torch.as_tensor(x.copy())
    F.svt.Random3DBlocksSVT(None).apply(t, 6, [4, 4, 4], 0.0, True)
    assert O.rel_l2(t.numpy().astype(np.complex128), x.astype(np.complex128)) < 1e-5
    # a threshold above every singular value zeroes each drawn block
    F.svt.Random3DBlocksSVT(None).apply(t, 3, [4, 4, 4], 1e6, True)
    zero = (t.abs().amax(dim=(0, 1)) == 0).numpy()
    assert 64 <= zero.sum() <= 3 * 64
    assert t.abs().amax() > 0
    # seeded comparison: the blocks the op drew (overlapping ones included), replayed in
    # the same order through the oracle's svt_soft on the Casorati matrices
    x = (rnd(rng, (6, 2, 14, 14, 14)) + 2.0).astype(np.complex64)
    t = torch.as_tensor(x.copy())
    op = F.svt.Random3DBlocksSVT(None)
    op.apply(t, 12, [5, 5, 5], 3.0, True)
    corners = op.last_corners().numpy()
    assert corners.shape == (12, 3) and (corners >= 0).all() and (corners <= 14 - 5 - 1).all()
    ref = x.astype(np.complex128)
    for c in corners:
        blk = O.extract_block(ref, c, [5, 5, 5])
        O.insert_block(ref, c, [5, 5, 5], O.svt_soft(blk, 3.0))
    e = O.rel_l2(t.numpy(), ref)
    print(f"Random3DBlocksSVT replayed through the oracle: {e:.2e}", flush=True)
    assert e < 1e-4 and O.rel_l2(ref, x) > 1e-3


def test_ffi_cpp_power_iteration_and_cg(F):
    """op::PowerIteration (opalgs.cpp:29-44) and op::ConjugateGradient (opalgs.cpp:50-96),
    plain and on l2_reg_linop(A, lam), driven through TorchScript over op::SenseNOp, against
    the oracle's power iteration and CG on the oracle SENSE normal operator; the CG's
    not-positive-definite error is the reference's."""
    from hastycompute_amd.ffi import hasty_nufft, hasty_op

    rng = np.random.default_rng(21)
    img = [12, 12, 12]
    M = 6000
    co, co_used = coords32(rng, 3, M)
    S = rnd(rng, [3] + img) * 0.4
    opts = hasty_nufft.NufftOptions(2, False, 1e-8), hasty_nufft.NufftOptions(1, True, 1e-8)
    op = hasty_op.SenseNOp(g(co, torch.float32), [1] + img, g(S), [0, 1, 2], *opts)
    Aref = lambda v: O.sense_normal_apply(co_used, v, S, [0, 1, 2], exact=True)
    v0 = rnd(rng, [1] + img)
    v = g(v0)
    lam = float(hasty_op.PowerIteration().run(op, v, 25))
    lam_ref = O.power_iteration(Aref, v0, 25)
    e_v = O.rel_l2(h(v), _power_vec(Aref, v0, 25))
    print(f"C++ PowerIteration: lambda {lam:.6e} vs oracle {lam_ref:.6e}, vector {e_v:.2e}", flush=True)
    assert abs(lam - lam_ref) / lam_ref < 1e-5 and e_v < 1e-4
    b = rnd(rng, [1] + img)
    for l2 in (0.0, 0.2 * lam_ref):
        x = torch.zeros((1, *img), dtype=torch.complex64, device="cuda")
        hasty_op.ConjugateGradient(op, g(b), l2).run(x, 8, 0.0)
        ref = O.conjugate_gradient(lambda w: Aref(w) + l2 * w, b, np.zeros_like(b), 8)
        e = O.rel_l2(h(x), ref)
        print(f"C++ ConjugateGradient l2={l2:.3g}: {e:.2e}", flush=True)
        assert e < 1e-4
    neg = hasty_op.ConjugateGradient(op, g(b), -2.0 * lam_ref)
    with pytest.raises(RuntimeError, match="A was not positive definite"):
        neg.run(torch.zeros((1, *img), dtype=torch.complex64, device="cuda"), 3, 0.0)


def _power_vec(A, v0, iters):
    v = np.array(v0, dtype=np.complex128)
    for _ in range(iters):
        v = A(v)
        v = v / np.linalg.norm(v)
    return v


def test_ffi_cpp_batched_cg(F):
    """op::BatchedConjugateGradient (opalgs.ixx:84-136) over a BatchedSenseNormal of 2
    frames x 2 encodings: each problem equals the oracle CG on its frame's Toeplitz SENSE
    normal operator (uniform points, well conditioned), and equals itself solved alone."""
    from hastycompute_amd.ffi import hasty_op

    rng = np.random.default_rng(22)
    n, C, nf, E = 16, 3, 2, 2
    cos = [rng.uniform(-np.pi, np.pi, (3, 3 * n ** 3)) for _ in range(nf)]
    S = rnd(rng, (C, n, n, n)) * 0.3 + 1.0
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=0, keepdims=True))
    bs = F.sense.BatchedSenseNormal([g(c, torch.float32) for c in cos], g(S))
    b = rnd(rng, (nf, E, n, n, n))
    x = torch.zeros((nf, E, n, n, n), dtype=torch.complex64, device="cuda")
    st = hasty_op.BatchedConjugateGradient(bs, g(b)).run(x, 6, 0.0)
    xs = h(x)
    assert tuple(st.shape) == (nf * E, 8) and bool((st[:, 6] == 6).all())
    for f in range(nf):
        diag = O.toeplitz_diagonal(cos[f].astype(np.float32).astype(np.float64), [1, n, n, n])
        A = lambda v: O.sense_normal_toeplitz(diag, v, S, list(range(C)))
        for e in range(E):
            ref = O.conjugate_gradient(A, b[f, e][None], np.zeros((1, n, n, n)), 6)
            err = O.rel_l2(xs[f, e][None], ref)
            print(f"C++ BatchedConjugateGradient frame {f} encoding {e}: {err:.2e}", flush=True)
            assert err < 1e-5
