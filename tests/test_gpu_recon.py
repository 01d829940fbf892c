"""GPU parity of the reconstruction layer (hastycompute_amd.base: MaxEig, GradientDescent
with the LLR prox, FivePointLLR) against the FP64 oracle's restatement of the
reference's Python loop (python/python/hastypy/base/recon.py, solvers.py, opalg.py).

Bounds (relative L2): power-iteration eigenvalue 1e-5; FivePointLLR image after a few
FISTA iterations 1e-4 (the LLR bound; the gradient itself is within 1e-5).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd.base import opalg, proximal, recon, sense, solvers

    return dict(opalg=opalg, proximal=proximal, recon=recon, sense=sense, solvers=solvers)


def rnd(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def make_scan(rng, nframes, n, C, M):
    """Per frame-encoding radial-like trajectories, kdata [1, C, M_p] and weights."""
    coords, kdata, weights = [], [], []
    for p in range(nframes * 5):
        m = M + 31 * p
        co = rng.uniform(-np.pi, np.pi, (3, m)).astype(np.float32)
        coords.append(co)
        kdata.append((0.05 * rnd(rng, (1, C, m))).astype(np.complex64))
        weights.append(rng.uniform(0.5, 1.0, m).astype(np.float32))
    S = np.ascontiguousarray(O.coil_maps(C, n, seed=5)).astype(np.complex64)
    return coords, kdata, weights, S


def test_batched_max_eig_vs_oracle_power_iteration(B):
    rng = np.random.default_rng(3)
    n, C = 16, 3
    coords, _, weights, S = make_scan(rng, 1, n, C, 700)
    coords, weights = coords[:3], weights[:3]
    x0 = (rng.uniform(0, 1, (3, 1, n, n, n)) + 1j * rng.uniform(0, 1, (3, 1, n, n, n))).astype(np.complex64)
    lam = B["recon"].batched_max_eig([torch.as_tensor(c).cuda() for c in coords], torch.as_tensor(S).cuda(),
                                     [torch.as_tensor(w).cuda() for w in weights], max_iter=6,
                                     x0=torch.as_tensor(x0)).cpu().numpy()
    for p in range(3):
        co = coords[p].astype(np.float64)
        ref = O.power_iteration(lambda v: O.sense_normal_apply(co, v, S.astype(np.complex128), list(range(C)),
                                                               weights=weights[p].astype(np.float64)),
                                x0[p].astype(np.complex128), 6)
        e = abs(lam[p] - ref) / ref
        print(f"max_eig p={p}: gpu {lam[p]:.6e} oracle {ref:.6e} rel {e:.2e}", flush=True)
        assert e < 1e-5


def test_conjugate_gradient_python_layer(B):
    """hastycompute_amd.base.solvers.ConjugateGradient on the batched normal operator
    equals the oracle CG on the explicit operator."""
    rng = np.random.default_rng(4)
    n, C = 16, 3
    coords, _, _, S = make_scan(rng, 1, n, C, 2500)
    co = coords[0]
    opalg, sense, solvers = B["opalg"], B["sense"], B["solvers"]
    A = sense.BatchedSenseNormal([torch.as_tensor(co).cuda()], torch.as_tensor(S).cuda())
    b = rnd(rng, (1, 1, n, n, n)).astype(np.complex64)
    x = opalg.Vector(torch.zeros((1, 1, n, n, n), dtype=torch.complex64, device="cuda"))
    cg = solvers.ConjugateGradient(A, opalg.Vector(torch.as_tensor(b).cuda()), x, max_iter=6)
    xg = cg.run().get_tensor().cpu().numpy()
    Sd = S.astype(np.complex128)
    ref = O.conjugate_gradient(lambda v: O.sense_normal_apply(co.astype(np.float64), v, Sd, list(range(C))),
                               b[0].astype(np.complex128), np.zeros((1, n, n, n)), 6)
    e = O.rel_l2(xg[0], ref)
    print(f"python-layer CG 6 iters: {e:.2e}", flush=True)
    assert e < 1e-5


def test_five_point_llr_vs_oracle_loop(B):
    rng = np.random.default_rng(11)
    nframes, n, C = 4, 32, 3
    coords, kdata, weights, S = make_scan(rng, nframes, n, C, 1500)
    recon, proximal = B["recon"], B["proximal"]
    x0 = (0.1 * rnd(rng, (nframes, 5, n, n, n))).astype(np.complex64)
    opts = proximal.SVTOptions(block_shapes=[8, 8, 8], block_strides=[8, 8, 8], block_iter=2, thresh=0.02,
                               soft=True)
    fp = recon.FivePointLLR(torch.as_tensor(S).cuda(), [torch.as_tensor(c).cuda() for c in coords],
                            [torch.as_tensor(k).cuda() for k in kdata], opts,
                            weights_vec=[torch.as_tensor(w).cuda() for w in weights], seed=1)
    iters = 3
    out = fp.run(torch.as_tensor(x0).cuda(), iter=iters).cpu().numpy()
    ref = O.five_point_llr([c.astype(np.float64) for c in coords], S.astype(np.complex128),
                           [k.astype(np.complex128) for k in kdata], [w.astype(np.float64) for w in weights],
                           x0.astype(np.complex128), fp.stepsize, iters, [8, 8, 8], [8, 8, 8], 2, 0.02, True)
    e = O.rel_l2(out, ref)
    print(f"FivePointLLR {iters} FISTA iterations (step {fp.stepsize:.3e}): {e:.2e}", flush=True)
    assert e < 1e-4
    # the LLR prox did something: without it the result differs
    ref0 = O.five_point_llr([c.astype(np.float64) for c in coords], S.astype(np.complex128),
                            [k.astype(np.complex128) for k in kdata], [w.astype(np.float64) for w in weights],
                            x0.astype(np.complex128), fp.stepsize, iters, [8, 8, 8], [8, 8, 8], 2, 0.0, True)
    assert O.rel_l2(ref0, ref) > 10 * e


def test_pdhg_python_layer(B):
    """base.solvers.PrimalDualHybridGradient on min_x 0.5|A x - y|^2 + lam |x|_1 with the
    batched SENSE forward / adjoint as K / K^H equals the oracle's PDHG (reference order)
    on the explicit operators, for the reference's and the classic extrapolation."""
    rng = np.random.default_rng(6)
    n, C = 12, 2
    coords, kdata, _, S = make_scan(rng, 1, n, C, 1800)
    co, y = coords[0], kdata[0]
    opalg, sense, solvers, proximal = B["opalg"], B["sense"], B["solvers"], B["proximal"]
    cot, St = torch.as_tensor(co).cuda(), torch.as_tensor(S).cuda()
    K = sense.BatchedSense([cot], St)
    KH = sense.BatchedSenseAdjoint([cot], St)
    yt = torch.as_tensor(y).cuda()
    tau, sigma, lam = 0.4, 0.4, 0.02

    class DualL2(proximal.ProximalOperator):
        # prox of sigma f*, f(z) = 0.5 |z - y|^2: (v - sigma y) / (1 + sigma)
        def _apply(self, alpha, inp):
            t = inp.get_tensorlist()[0]
            return opalg.Vector([(t - sigma * yt) / (1 + sigma)])

    class Soft(proximal.ProximalOperator):
        def _apply(self, alpha, inp):
            t = inp.get_tensor()
            return opalg.Vector(torch.polar(torch.clamp(t.abs() - lam * tau, min=0.0), t.angle()))

    Sd, cod, yd = S.astype(np.complex128), co.astype(np.float64), y[0].astype(np.complex128)
    Kd = lambda v: O.sense_apply(cod, v[None], Sd, list(range(C)))  # noqa: E731
    KHd = lambda w: O.sense_adjoint_apply(cod, w, Sd, list(range(C)))[0]  # noqa: E731
    fc = lambda v, s: (v - sigma * yd) / (1 + sigma)  # noqa: E731
    g = lambda v, t: np.exp(1j * np.angle(v)) * np.maximum(np.abs(v) - lam * tau, 0.0)  # noqa: E731
    for classic in (False, True):
        x = opalg.Vector(torch.zeros((1, 1, n, n, n), dtype=torch.complex64, device="cuda"))
        yv = opalg.Vector([torch.zeros((1, C, co.shape[1]), dtype=torch.complex64, device="cuda")])
        pd = solvers.PrimalDualHybridGradient(DualL2(K.oshape, None, False), Soft(K.ishape, None, False), K, KH, x,
                                              yv, opalg.Vector(torch.tensor(tau)), opalg.Vector(torch.tensor(sigma)),
                                              0.0, max_iter=8, classic=classic)
        xg = pd.run().get_tensor().cpu().numpy()[0, 0]
        ref, _ = O.pdhg(Kd, KHd, fc, g, np.zeros((n, n, n)), np.zeros((C, co.shape[1])), tau, sigma, 0.0, 8,
                        classic=classic)
        e = O.rel_l2(xg, ref)
        print(f"PDHG classic={classic}: {e:.2e}", flush=True)
        assert e < 1e-4


@pytest.mark.parametrize("shape,axis", [((2, 1, 12, 10, 9), (True, True, True)), ((3, 16, 20), (False, True)),
                                        ((1, 1, 7, 8, 5), (True, False, True))])
def test_dct_and_l1dct_prox_vs_oracle(B, shape, axis):
    import scipy.fft as sfft
    from hastycompute_amd import native

    rng = np.random.default_rng(sum(shape))
    x = rnd(rng, shape)
    nd = len(axis)
    axes = tuple(len(shape) - nd + i for i, a in enumerate(axis) if a)
    xt = torch.as_tensor(x.astype(np.complex64)).cuda()
    native.dctn_(xt, axis)
    torch.cuda.synchronize()
    assert O.rel_l2(xt.cpu().numpy(), sfft.dctn(x, axes=axes)) < 2e-6
    native.dctn_(xt, axis, inverse=True)
    torch.cuda.synchronize()
    assert O.rel_l2(xt.cpu().numpy(), x) < 2e-6
    lam = 0.3 * float(np.median(np.abs(sfft.dctn(x, axes=axes))))
    prox = B["proximal"].ProximalL1DCT(shape, axis, lam, inplace=False)
    got = prox(B["opalg"].Vector(torch.as_tensor(x.astype(np.complex64)).cuda())).get_tensor().cpu().numpy()
    e = O.rel_l2(got, O.l1dct_prox(x, lam, axis))
    print(f"l1dct prox {shape} {axis}: {e:.2e}", flush=True)
    assert e < 2e-6


def test_five_point_full_vs_oracle(B):
    """FivePointFULL (recon.py:47-83): frames stacked per encoding, GD with the L1-DCT prox
    against the oracle's loop with the same max eigenvalue."""
    rng = np.random.default_rng(8)
    n, C, nframes = 12, 2, 2
    coords, kdata, weights, S = make_scan(rng, nframes, n, C, 600)
    recon = B["recon"]
    cvec = [torch.as_tensor(c).cuda() for c in coords]
    kvec = [torch.as_tensor(k).cuda() for k in kdata]
    wvec = [torch.as_tensor(w).cuda() for w in weights]
    sc, sk = recon.stack_frame_datas_fivepoint(cvec, kvec)
    sw = [torch.cat([wvec[f * 5 + e] for f in range(nframes)]) for e in range(5)]
    max_eig, lamda, iters = 3.0, 2e-3, 4
    fp = recon.FivePointFULL(torch.as_tensor(S).cuda(), sc, sk, sw, lamda=lamda, max_eig=max_eig)
    x0 = (0.1 * rnd(rng, (5, 1, n, n, n))).astype(np.complex64)
    xg = fp.run(torch.as_tensor(x0).cuda(), iter=iters).cpu().numpy()
    ref = O.five_point_full([c.cpu().numpy().astype(np.float64) for c in sc], S.astype(np.complex128),
                            [k.cpu().numpy().astype(np.complex128) for k in sk],
                            [w.cpu().numpy().astype(np.float64) for w in sw], x0, max_eig, lamda, iters)
    e = O.rel_l2(xg, ref)
    print(f"FivePointFULL {iters} iterations: {e:.2e}", flush=True)
    assert e < 1e-4


def test_create_spokes_vs_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import simulate

    xa, za = [0.3, 2.0, 1.2], [1.1, 5.0, 0.2]
    sp = simulate.create_spokes(3, 40, noise=0.0, xangles=xa, zangles=za).cpu().numpy()
    for i in range(3):
        ref = O.create_spoke(40, xa[i], za[i])
        assert np.abs(sp[:, i * 40:(i + 1) * 40] - ref).max() < 1e-5
    cr = simulate.create_spokes(3, 40, noise=0.0, crop_factor=1.4, xangles=xa, zangles=za)
    for i in range(3):
        ref = O.create_spoke(40, xa[i], za[i], crop_factor=1.4)
        assert cr[i].shape == ref.shape and np.abs(cr[i].cpu().numpy() - ref).max() < 1e-5
    # jitter: the samples leave the line by N(0, noise) per axis
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    jit = simulate.create_spokes(400, 64, noise=0.01, xangles=[0.0] * 400, zangles=[0.0] * 400, generator=g)
    std = float(jit[0].std())
    assert 0.009 < std < 0.011
    rnd_angles = simulate.create_spokes(200, 16, generator=g)
    assert rnd_angles.shape == (3, 3200) and float(rnd_angles.abs().max()) < 1.1 * np.pi
