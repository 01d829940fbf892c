"""GPU parity of the circulant preconditioner (CirculantPreconditionerOp,
lib/mri/precond.cpp:27-117) and of the preconditioned device CG against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def N():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import native

    return native


def g(a, dtype=torch.complex64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def h(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128)


def maps(rng, C, shape):
    m = (rng.standard_normal((C,) + shape) + 1j * rng.standard_normal((C,) + shape)) * 0.3 + 1.0
    return m / np.sqrt(np.sum(np.abs(m) ** 2, axis=0, keepdims=True))


@pytest.mark.parametrize("shape,weighted", [((8, 16, 12), True), ((10, 12, 16), False), ((5, 12), True)])
def test_precond_diagonal_matches_oracle(N, shape, weighted):
    rng = np.random.default_rng(sum(shape))
    C, M = 3, 4000
    S = maps(rng, C, shape)
    co = rng.uniform(-np.pi, np.pi, (len(shape), M))
    w = rng.uniform(0.3, 2.0, M) if weighted else None
    ref = O.circulant_precond_diagonal(S, co, w, lam=0.05)
    for centered in (True, False):
        d = h(N.circulant_precond_diagonal(g(S), g(co, torch.float32),
                                           g(w, torch.float32) if w is not None else None, lamda=0.05,
                                           centered=centered))
        exp = ref if centered else np.fft.ifftshift(ref)
        e = O.rel_l2(d, exp)
        print(f"precond diagonal {shape} weighted={weighted} centered={centered}: {e:.2e}", flush=True)
        assert e < 1e-5


@pytest.mark.parametrize("shape", [(8, 16, 12), (6, 10), (9, 15)])
def test_precond_apply_matches_oracle(N, shape):
    rng = np.random.default_rng(len(shape) * 7 + shape[0])
    P = 3
    d = rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)
    x = rng.standard_normal((P,) + shape) + 1j * rng.standard_normal((P,) + shape)
    idx = [1, 0, 1]
    for centered in (True, False):
        for adjoint in (False, True):
            out = h(N.circulant_precond_apply(g(x), g(d), diag_index=idx, centered=centered, adjoint=adjoint))
            for p in range(P):
                ref = O.circulant_precond_apply(d[idx[p]], x[p], centered, adjoint)
                e = O.rel_l2(out[p], ref)
                assert e < 1e-5, (shape, centered, adjoint, p, e)
    # in place, one diagonal for all problems
    xt = g(x)
    N.circulant_precond_apply(xt, g(d[0]), out=xt)
    got = h(xt)
    for p in range(P):
        assert O.rel_l2(got[p], O.circulant_precond_apply(d[0], x[p])) < 1e-5


def test_preconditioned_device_cg_matches_oracle(N):
    """Device PCG on the Toeplitz SENSE normal operator with the circulant preconditioner
    against the oracle's ConjugateGradient with P (same diagonal, FP64)."""
    rng = np.random.default_rng(5)
    shape = (16, 32, 32)
    C, P = 4, 2
    S = maps(rng, C, shape)
    cos = [O.kooshball(300, 32, offset=300 * i).astype(np.float32).astype(np.float64) for i in range(P)]
    K = torch.stack([N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)[1] for co in cos])
    diags = [O.toeplitz_diagonal(co, [1, *shape]) for co in cos]
    pd = [O.circulant_precond_diagonal(S, co, lam=1e-3) for co in cos]
    dev_pd = torch.stack([N.circulant_precond_diagonal(g(S), g(co, torch.float32), lamda=1e-3) for co in cos])
    op = N.SenseNormalToeplitz(g(S), K)
    b = rng.standard_normal((P,) + shape) + 1j * rng.standard_normal((P,) + shape)
    iters = 8

    def A(v, out):
        op.apply(v, out=out, kernel_index=list(range(P)))

    def Pc(v, out):
        N.circulant_precond_apply(v, dev_pd, out=out, diag_index=list(range(P)))

    x, st = N.cg_solve(A, g(b), torch.zeros_like(g(b)), iters, apply_P=Pc)
    x = h(x)
    for p in range(P):
        ref = O.conjugate_gradient(lambda v: O.sense_normal_toeplitz(diags[p], v[None], S, range(C))[0], b[p],
                                   np.zeros(shape, complex), iters,
                                   P=lambda v: O.circulant_precond_apply(pd[p], v))
        e = O.rel_l2(x[p], ref)
        print(f"PCG problem {p}: rel diff vs oracle {e:.2e}, iterations {st[p, 6].item():.0f}", flush=True)
        assert e < 2e-4


def test_preconditioner_speeds_up_cg(N):
    """On an undersampled kooshball SENSE problem with Tikhonov regularisation,
    (A^H A + mu I) x = b for a random (white) x, the circulant preconditioner built with
    lambda = mu (the eigenvalues of the nearest circulant of A^H A + mu I) lowers the error
    to the solution reached in a fixed number of device-CG iterations. Without mu the
    unsampled corners of k-space give near-zero eigenvalues and 1/p amplifies them, which
    is why the reference exposes lambda. (For a smooth phantom plain CG is already good:
    its first iterations fit the strongly sampled low frequencies.)"""
    from hastycompute_amd import synthetic

    shape = (64, 64, 64)
    C = 8
    Sg = synthetic.coil_maps_torch(C, 64, "cuda").to(torch.complex64).contiguous()
    co = O.kooshball(600, 128).astype(np.float32)
    K = N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)[1][None]
    op = N.SenseNormalToeplitz(Sg, K)
    rng = np.random.default_rng(3)
    xt = g(rng.standard_normal((1,) + shape) + 1j * rng.standard_normal((1,) + shape))
    d0 = N.circulant_precond_diagonal(Sg, g(co, torch.float32), lamda=0.0)
    mu = 1e-2 * torch.mean(torch.real(1.0 / d0)).item()
    pdiag = N.circulant_precond_diagonal(Sg, g(co, torch.float32), lamda=mu)

    def A(v, out):
        op.apply(v, out=out)
        out.add_(v, alpha=mu)

    b = torch.empty_like(xt)
    A(xt, b)

    def Pc(v, out):
        N.circulant_precond_apply(v, pdiag, out=out)

    err = {}
    for name, pc in [("cg", None), ("pcg", Pc)]:
        x, _ = N.cg_solve(A, b, torch.zeros_like(b), 10, apply_P=pc)
        err[name] = (torch.linalg.vector_norm(x - xt) / torch.linalg.vector_norm(xt)).item()
    print(f"mu = {mu:.3e}; error after 10 iterations: CG {err['cg']:.3e}, PCG {err['pcg']:.3e}", flush=True)
    assert err["pcg"] < 0.75 * err["cg"]


def test_precond_linop_matches_oracle(N):
    """base.precond.PrecondLinop (reconstruct_util.py:436-464): fftn(ifftn(x) * P_f) per
    frame with P_f the frame's diagonal in FFT order; and ConjugateGradient with it."""
    from hastycompute_amd.base import precond as BP
    from hastycompute_amd.base.opalg import Vector

    rng = np.random.default_rng(21)
    shape = (8, 12, 10)
    C, F = 2, 3
    S = maps(rng, C, shape)
    cos = [rng.uniform(-np.pi, np.pi, (3, 1500)) for _ in range(F)]
    op = BP.PrecondLinop(g(S), [g(c, torch.float32) for c in cos])
    x = rng.standard_normal((F, 1) + shape) + 1j * rng.standard_normal((F, 1) + shape)
    out = h(op(Vector(g(x))).get_tensor())
    for f in range(F):
        Pf = np.fft.ifftshift(O.circulant_precond_diagonal(S, cos[f]))
        ref = np.fft.fftn(np.fft.ifftn(x[f, 0]) * Pf)
        assert O.rel_l2(out[f, 0], ref) < 1e-5
