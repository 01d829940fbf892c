"""Pins for the FP64 oracle (CPU only). Each check is chosen so that a plausible slip in
the oracle (a transposed coordinate row, a wrong sign or mode offset, a missing scale)
fails it; none of them re-types the oracle's own formula."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")
rng = np.random.default_rng(1234)


def cx(d):
    return (np.asarray(d["re"]) + 1j * np.asarray(d["im"])).reshape(d["shape"])


def rnd(shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def cartesian(N):
    g = [-np.pi + 2 * np.pi * np.arange(n) / n for n in N]
    G = np.meshgrid(*g, indexing="ij")
    return np.stack([x.ravel() for x in G])


@pytest.mark.parametrize("N", [(7,), (6, 5), (4, 6, 5), (3, 2, 5)])
def test_direct_sum_is_shifted_dft_on_cartesian_grid(N):
    # x_m = -pi + 2 pi m/N makes type 2 a DFT of f with phase (-1)^k (numpy.fft pins it).
    co = cartesian(N)
    f = rnd(N)
    fs = f.copy()
    for d, n in enumerate(N):
        sh = [1] * len(N)
        sh[d] = n
        fs = fs * np.exp(1j * np.pi * (np.arange(n) - n // 2)).reshape(sh)
    # place frequency k at index k mod n, then forward DFT
    ref = np.fft.fftn(np.roll(fs, [-(n // 2) for n in N], axis=tuple(range(len(N)))))
    c = O.nudft_type2(co, f, -1).reshape(N)
    assert np.abs(c - ref).max() <= 1e-11 * np.abs(ref).max()
    # type 1 with +1 is the adjoint: on the full grid it is N^d times the inverse
    back = O.nudft_type1(co, c.ravel(), N, +1)
    assert np.abs(back - np.prod(N) * f).max() <= 1e-10 * np.abs(f).max() * np.prod(N)


def test_direct_sums_adjoint_and_row_pairing():
    co = rng.uniform(-np.pi, np.pi, (3, 40))
    N = (3, 4, 5)
    f, c = rnd(N), rnd(40)
    lhs = np.vdot(O.nudft_type2(co, f, -1), c)
    rhs = np.vdot(f, O.nudft_type1(co, c, N, +1))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # a single mode at index n gives exp(-i (n - N//2) . x): row d pairs with dim d
    f = np.zeros(N, dtype=complex)
    f[2, 0, 4] = 1.0
    k = np.array([2 - 1, 0 - 2, 4 - 2])
    assert np.allclose(O.nudft_type2(co, f, -1), np.exp(-1j * (k @ co)), atol=1e-13)


@pytest.mark.parametrize("tol", [1e-6, 1e-9, 1e-12])
def test_es_nufft_matches_direct_sum(tol):
    co = rng.uniform(-3 * np.pi, 3 * np.pi, (3, 300))
    N = (10, 9, 8)
    f, c = rnd(N), rnd(300)
    e2 = O.rel_l2(O.nufft_type2(co, f, -1, tol), O.nudft_type2(co, f, -1))
    e1 = O.rel_l2(O.nufft_type1(co, c, N, +1, tol), O.nudft_type1(co, c, N, +1))
    assert e2 < 10 * tol and e1 < 10 * tol
    # opposite sign too
    assert O.rel_l2(O.nufft_type2(co, f, +1, tol), O.nudft_type2(co, f, +1)) < 10 * tol


def test_es_nufft_2d_1d():
    co = rng.uniform(-np.pi, np.pi, (2, 200))
    f = rnd((12, 7))
    assert O.rel_l2(O.nufft_type2(co, f, -1, 1e-10), O.nudft_type2(co, f, -1)) < 1e-9
    co1 = rng.uniform(-np.pi, np.pi, (1, 100))
    c = rnd(100)
    assert O.rel_l2(O.nufft_type1(co1, c, (16,), 1, 1e-10), O.nudft_type1(co1, c, (16,), 1)) < 1e-9


def test_golden_deterministic_3d_closed_form():
    fx = json.load(open(GOLDEN))["deterministic_3d"]
    co = np.asarray(fx["coords"], dtype=np.float64)
    # f[0,0,0,1] = 2 -> k = (-2,-2,-1); f[0,2,1,3] = 3 -> k = (0,-1,1); type 2 with sign -1, /sqrt(64)
    closed = (2 * np.exp(-1j * (np.array([-2, -2, -1]) @ co)) + 3 * np.exp(-1j * (np.array([0, -1, 1]) @ co))) / 8
    assert np.allclose(cx(fx["type2"])[0], closed, atol=1e-12)
    got = O.nufft_apply(co, fx["nmodes"], 2, cx(fx["input"]), exact=True)
    assert np.allclose(got, cx(fx["type2"]), atol=1e-12)
    back = O.nufft_apply(co, fx["nmodes"], 1, got, exact=True)
    assert np.allclose(back, cx(fx["type1_of_type2"]), atol=1e-12)


def test_golden_deterministic_1d_closed_form():
    fx = json.load(open(GOLDEN))["deterministic_1d"]
    co = np.asarray(fx["coords"], dtype=np.float64)
    # sum over k=-4..3 of exp(+i k x) at x = -pi + 2 pi m / 8: 8 (-1)^m... only m = 4 (x = 0) is nonzero
    out = cx(fx["type2"])[0]
    expect = np.zeros(8, dtype=complex)
    expect[4] = 8 / math.sqrt(8)
    assert np.allclose(out, expect, atol=1e-5)  # float32 coordinates
    got = O.nufft_apply(co, fx["nmodes"], 2, cx(fx["input"]), positive=True, exact=True)
    assert np.allclose(got, cx(fx["type2"]), atol=1e-12)


def test_golden_compare_normal_methods_toeplitz_scale():
    fx = json.load(open(GOLDEN))["compare_normal_methods"]
    normal, toep = cx(fx["normal"]), cx(fx["toeplitz"])
    # the reference's Toeplitz operator is 2^-d times its explicit normal operator
    assert np.allclose(toep, normal / 8, atol=1e-12 * np.abs(normal).max())
    co = np.asarray(fx["coords"], dtype=np.float64)
    assert np.allclose(O.nufft_normal_apply(co, fx["nmodes"], np.ones(fx["nmodes"], complex), exact=True), normal)


def test_toeplitz_equals_explicit_normal_random():
    co = rng.uniform(-np.pi, np.pi, (3, 150))
    nm = [1, 6, 5, 4]
    x = rnd(nm)
    diag = O.toeplitz_diagonal(co, nm, exact=True)
    t = O.toeplitz_apply(diag, x)
    e = O.nufft_normal_apply(co, nm, x, exact=True)
    assert O.rel_l2(t * 8, e) < 1e-12
    # ES-built diagonal agrees with the exact one
    diag2 = O.toeplitz_diagonal(co, nm, exact=False)
    assert O.rel_l2(diag2, diag) < 1e-10


def test_sense_normal_hermitian_psd_and_cartesian_identity():
    N = (4, 5, 6)
    co = rng.uniform(-np.pi, np.pi, (3, 200))
    S = O.coil_maps(3, 6)[:, :4, :5, :6]
    S = S / np.sqrt(np.sum(np.abs(S) ** 2, axis=0, keepdims=True))
    x, y = rnd((1,) + N), rnd((1,) + N)
    Ax = O.sense_normal_apply(co, x, S, [0, 1, 2], exact=True)
    Ay = O.sense_normal_apply(co, y, S, [0, 1, 2], exact=True)
    assert abs(np.vdot(y, Ax) - np.vdot(Ay, x)) < 1e-10 * abs(np.vdot(y, Ax))
    assert np.vdot(x, Ax).real > 0
    # Toeplitz form at SenseNormal scale equals the explicit form
    diag = O.toeplitz_diagonal(co, [1, *N], exact=True)
    assert O.rel_l2(O.sense_normal_toeplitz(diag, x, S, [0, 1, 2]), Ax) < 1e-12
    # full Cartesian sampling: F^H F = N^d I and sum |S_c|^2 = 1 -> identity
    coc = cartesian(N)
    assert O.rel_l2(O.sense_normal_apply(coc, x, S, [0, 1, 2], exact=True), x) < 1e-11


def test_sense_gradient_form():
    N = (4, 4, 4)
    co = rng.uniform(-np.pi, np.pi, (3, 100))
    S = O.coil_maps(2, 4)
    x = rnd((1,) + N)
    y = rnd((2, 100))
    w = rng.uniform(0.5, 1.5, 100)
    g = O.sense_normal_apply(co, x, S, [0, 1], kdata=y, weights=w, exact=True)
    # A^H W (A x - y) written out with the forward and adjoint operators
    Ax = O.sense_apply(co, x, S, [0, 1], exact=True)
    expect = O.sense_adjoint_apply(co, (Ax - y) * w[None], S, [0, 1], exact=True)
    assert O.rel_l2(g, expect) < 1e-12


def test_cg_matches_dense_solve_and_detects_indefinite():
    n = 12
    Q = rnd((n, n))
    A = Q.conj().T @ Q + 0.5 * np.eye(n)
    b = rnd(n)
    x = O.conjugate_gradient(lambda v: A @ v, b, np.zeros(n, complex), iters=3 * n)
    assert O.rel_l2(x, np.linalg.solve(A, b)) < 1e-8
    with pytest.raises(O.NotPositiveDefinite):
        O.conjugate_gradient(lambda v: -(A @ v), b, np.zeros(n, complex), iters=3)
    # tol stops before any update when the initial residual is already below it
    x0 = np.linalg.solve(A, b)
    assert np.array_equal(O.conjugate_gradient(lambda v: A @ v, b, x0, iters=5, tol=1e-3), x0)


def test_svt_closed_forms():
    R, F, r = 300, 20, 5
    U, _ = np.linalg.qr(rnd((R, r)))
    V, _ = np.linalg.qr(rnd((F, r)))
    s = np.array([10.0, 7.0, 5.0, 3.0, 1.0])
    A = (U * s) @ V.conj().T
    lam = 2.0
    out = O.svt_soft(A, lam)
    assert O.rel_l2(out, (U * np.maximum(s - lam, 0)) @ V.conj().T) < 1e-12
    assert O.rel_l2(O.svt_hard(A, 2), (U[:, :2] * s[:2]) @ V[:, :2].conj().T) < 1e-12


def test_llr_shift_sequence_is_libstdcxx():
    for it, shape in [(1, (16, 16, 16)), (6, (16, 16, 16)), (20, (8, 4, 12)), (4, (1, 2, 3))]:
        assert np.array_equal(O.llr_shifts(it, shape), O.llr_shifts_libstdcxx(it, shape))


def test_llr_block_grid_skip_rule():
    it = O.normal_blocks((128, 128, 128), (16, 16, 16), (16, 16, 16), 1)
    assert len(it[0]) == 343 and max(b[0] for b in it[0]) == 96
    assert len(O.normal_blocks((32, 32, 32), (16, 16, 16), (16, 16, 16), 1)[0]) == 1
    with pytest.raises(RuntimeError):
        O.normal_blocks((32, 32, 32), (16, 16, 16), (16, 16, 16), 0)


def test_extract_insert_roundtrip_and_layout():
    x = rnd((4, 2, 6, 6, 6))
    A = O.extract_block(x, (1, 2, 0), (2, 3, 4))
    assert A.shape == (2 * 2 * 3 * 4, 4)
    # row (e, x, y, z) = (1, 1, 2, 3), column f = 3
    assert A[((1 * 2 + 1) * 3 + 2) * 4 + 3, 3] == x[3, 1, 2, 4, 3]
    y = x.copy()
    O.insert_block(y, (1, 2, 0), (2, 3, 4), A)
    assert np.array_equal(x, y)


def test_normal_blocks_svt_sequential_semantics():
    x = rnd((6, 1, 8, 8, 8))
    out = O.normal_3d_blocks_svt(x, (4, 4, 4), (4, 4, 4), 1, 0.0, True)
    # thresh 0 keeps everything; blocks touching the far edge are skipped either way
    assert O.rel_l2(out, x) < 1e-12
    out = O.normal_3d_blocks_svt(x, (4, 4, 4), (4, 4, 4), 1, 1e6, True)
    assert np.abs(out[:, :, :4, :4, :4]).max() == 0.0
    assert np.array_equal(out[:, :, 4:, :, :], x[:, :, 4:, :, :])


def test_power_iteration_converges_to_largest_eigenvalue():
    """MaxEig: on a Hermitian PSD matrix with a spectral gap the estimate approaches the
    largest eigenvalue (numpy eigvalsh) from below; a slip (returning |x|, normalising
    before applying A) fails."""
    Q, _ = np.linalg.qr(rnd((12, 12)))
    lam = np.array([5.0, 2.0] + [1.0] * 10)
    A = (Q * lam) @ Q.conj().T
    x0 = rnd(12)
    est = O.power_iteration(lambda v: A @ v, x0, 40)
    assert abs(est - 5.0) < 1e-9
    assert O.power_iteration(lambda v: A @ v, x0, 3) <= 5.0 + 1e-12


def test_gradient_descent_lasso_orthonormal_closed_form():
    """GD / FISTA with a soft-threshold prox on 0.5|Q x - b|^2 + lam |x|_1, Q unitary,
    step 1: the minimiser is soft(Q^H b, lam) and both variants reach it and stay."""
    Q, _ = np.linalg.qr(rnd((10, 10)))
    b = rnd(10)
    lam = 0.7

    def soft(v, t=lam):
        return np.exp(1j * np.angle(v)) * np.maximum(np.abs(v) - t, 0.0)

    grad = lambda v: Q.conj().T @ (Q @ v - b)  # noqa: E731
    xs = soft(Q.conj().T @ b)
    for acc in (False, True):
        x = O.gradient_descent(grad, rnd(10), soft, 1.0, 5, accelerate=acc)
        assert np.abs(x - xs).max() < 1e-12
    # without a prox, FISTA on an ill-conditioned least-squares problem converges to lstsq
    r2 = np.random.default_rng(7)
    A = (r2.standard_normal((30, 8)) + 1j * r2.standard_normal((30, 8))) @ np.diag(np.linspace(1.0, 0.5, 8))
    bb = r2.standard_normal(30) + 1j * r2.standard_normal(30)
    L = np.linalg.eigvalsh(A.conj().T @ A).max()
    x = O.gradient_descent(lambda v: A.conj().T @ (A @ v - bb), np.zeros(8), None, 1.0 / L, 2000, accelerate=True)
    assert np.abs(x - np.linalg.lstsq(A, bb, rcond=None)[0]).max() < 1e-8


def _dense(op, shape):
    """Matrix of a linear map on arrays of `shape` (column j = op(e_j))."""
    n = int(np.prod(shape))
    cols = []
    for j in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1.0
        cols.append(np.asarray(op(e.reshape(shape))).ravel())
    return np.stack(cols, axis=1)


def test_circulant_precond_is_frobenius_optimal_circulant():
    """1/diag are the eigenvalues of the circulant matrix closest in Frobenius norm to
    T = sum_c S_c^H N1 diag(sqrt(w)) N2 S_c, i.e. the diagonal of U T U^H with U the
    centred unitary DFT (brute force on small 2-D problems): a wrong scale, a dropped
    sqrt, the other half of the doubled grid ([1::2]) or an uncentred order fails it;
    lambda is added to those eigenvalues. Even sizes only: for odd N the reference's
    [::2] reads doubled-grid frequency 2i - N, half a bin off the N grid (kept as is)."""
    r2 = np.random.default_rng(11)
    for img in [(6, 8), (4, 10)]:
        C, M = 2, 50
        S = r2.standard_normal((C,) + img) + 1j * r2.standard_normal((C,) + img)
        co = r2.uniform(-np.pi, np.pi, (2, M))
        w = r2.uniform(0.2, 1.5, M)
        nm = [1] + list(img)

        def T(x):
            out = np.zeros(img, dtype=np.complex128)
            for c in range(C):
                k = O.nufft_apply(co, nm, 2, (x * S[c])[None], exact=True)
                out += np.conj(S[c]) * O.nufft_apply(co, nm, 1, k * np.sqrt(w)[None], exact=True)[0]
            return out

        Tm = _dense(T, img)
        U = _dense(lambda v: O.fftc(v), img)
        eig = np.diag(U @ Tm @ U.conj().T)
        d = O.circulant_precond_diagonal(S, co, w, exact=True)
        assert np.abs(1.0 / d.ravel() - eig).max() < 1e-10 * np.abs(eig).max()
        d2 = O.circulant_precond_diagonal(S, co, w, lam=0.3, exact=True)
        assert np.abs(1.0 / d2.ravel() - (eig + 0.3)).max() < 1e-10 * np.abs(eig).max()
        # the ES NUFFT PSF gives the same diagonal
        d3 = O.circulant_precond_diagonal(S, co, w)
        assert O.rel_l2(d3, d) < 1e-9


def test_circulant_precond_apply_is_circular_convolution():
    """The apply is a circular convolution with ifftn(d) (ifftn(ifftshift(d)) for a
    centred d: the data are not shifted); the adjointified apply with fftn(conj d) / N.
    Checked against direct circular sums on odd and even sizes."""
    r2 = np.random.default_rng(12)
    for shape in [(4, 6), (5, 3, 4)]:
        d = r2.standard_normal(shape) + 1j * r2.standard_normal(shape)
        x = r2.standard_normal(shape) + 1j * r2.standard_normal(shape)
        n = int(np.prod(shape))
        idx = np.array(np.unravel_index(np.arange(n), shape))

        def circ(h):
            out = np.zeros(n, dtype=np.complex128)
            xf = x.ravel()
            for k in range(n):
                j = tuple(((idx[:, k][:, None] - idx) % np.array(shape)[:, None]))
                out[k] = np.sum(h[j] * xf)
            return out.reshape(shape)

        for centered in (False, True):
            dn = np.fft.ifftshift(d) if centered else d
            h = np.fft.ifftn(dn)
            assert O.rel_l2(O.circulant_precond_apply(d, x, centered), circ(h)) < 1e-12
            g = np.fft.fftn(np.conj(dn)) / n
            assert O.rel_l2(O.circulant_precond_apply(d, x, centered, adjoint=True), circ(g)) < 1e-12


def test_preconditioned_cg_matches_dense_solve():
    """CG with a preconditioner (the _P branch): on an ill-conditioned SPD system with an
    SPD preconditioner it reaches numpy.linalg.solve, and in fewer iterations than plain CG
    when P approximates the inverse."""
    n = 30
    Q, _ = np.linalg.qr(rnd((n, n)))
    lam = np.logspace(0, 4, n)
    A = (Q * lam) @ Q.conj().T
    Pm = (Q * (1.0 / (lam * np.linspace(0.8, 1.2, n)))) @ Q.conj().T
    b = rnd(n)
    xs = np.linalg.solve(A, b)
    xp = O.conjugate_gradient(lambda v: A @ v, b, np.zeros(n, complex), iters=3 * n, P=lambda v: Pm @ v)
    assert O.rel_l2(xp, xs) < 1e-8
    xp5 = O.conjugate_gradient(lambda v: A @ v, b, np.zeros(n, complex), iters=8, P=lambda v: Pm @ v)
    xc5 = O.conjugate_gradient(lambda v: A @ v, b, np.zeros(n, complex), iters=8)
    assert O.rel_l2(xp5, xs) < 0.1 * O.rel_l2(xc5, xs)


def test_pdhg_fixed_point_and_classic_closed_form():
    """PDHG on min_x 0.5|x - b|^2 + lam |x|_1 with K = I (proxfc = projection onto the
    lam-ball, proxg(v, t) = (v + t b) / (1 + t)). The classic extrapolation reaches
    soft(b, lam), with and without acceleration. The reference's order (x_bar from the
    previous x_bar) does not converge here, but keeps the primal-dual solution
    (soft(b, lam), b - soft(b, lam)) fixed, and its first step equals the classic one:
    a sign error in the dual step or a dropped step size breaks both."""
    b = rnd(40)
    lam = 0.6
    soft = np.exp(1j * np.angle(b)) * np.maximum(np.abs(b) - lam, 0.0)

    def clip(v, s):
        m = np.abs(v)
        return np.where(m > lam, v * (lam / np.maximum(m, 1e-300)), v)

    I = lambda v: v  # noqa: E731
    pg = lambda v, t: (v + t * b) / (1 + t)  # noqa: E731
    for acc in (0.0, 0.5):
        x, _ = O.pdhg(I, I, clip, pg, np.zeros(40), np.zeros(40), 0.5, 0.5, acc, 300 if acc == 0 else 20000,
                      classic=True)
        assert np.abs(x - soft).max() < (1e-8 if acc == 0 else 1e-6), acc
        xr, yr = O.pdhg(I, I, clip, pg, soft, b - soft, 0.5, 0.5, acc, 25)
        assert np.abs(xr - soft).max() < 1e-12 and np.abs(yr - (b - soft)).max() < 1e-12
        x0 = rnd(40)
        a1 = O.pdhg(I, I, clip, pg, x0, np.zeros(40), 0.5, 0.5, acc, 1)
        c1 = O.pdhg(I, I, clip, pg, x0, np.zeros(40), 0.5, 0.5, acc, 1, classic=True)
        assert np.array_equal(a1[0], c1[0]) and np.array_equal(a1[1], c1[1])


def test_dct_matches_cosine_sums_and_l1dct_prox():
    """The oracle's L1-DCT prox rests on scipy's dctn / idctn: checked against the cosine
    sums y_k = 2 sum x_n cos(pi k (2n+1) / 2N) and their inverse on an odd/even 3-D size,
    one and all axes; lam = 0 gives back x, a huge lam gives 0."""
    import scipy.fft as sfft
    x = rnd((5, 6, 3))
    for axes in [(0, 1, 2), (1,)]:
        ref = x.copy()
        for ax in axes:
            N = x.shape[ax]
            n = np.arange(N)
            Cm = 2.0 * np.cos(np.pi * np.outer(n, 2 * n + 1) / (2 * N))
            ref = np.moveaxis(np.tensordot(Cm, np.moveaxis(ref, ax, 0), axes=(1, 0)), 0, ax)
        assert O.rel_l2(sfft.dctn(x, axes=axes), ref) < 1e-13
        assert O.rel_l2(sfft.idctn(ref, axes=axes), x) < 1e-13
    assert O.rel_l2(O.l1dct_prox(x, 0.0, (True, True, True)), x) < 1e-13
    assert np.abs(O.l1dct_prox(x, 1e6, (True, True, True))).max() == 0.0
    # a leading batch axis is not transformed
    xb = rnd((2, 4, 5))
    pb = O.l1dct_prox(xb, 0.5, (True, True))
    assert O.rel_l2(pb[1], O.l1dct_prox(xb[1], 0.5, (True, True))) < 1e-13


def test_create_spoke_closed_form():
    """Without jitter a PC-VIPR spoke is t (sin x sin z, -sin x cos z, cos x) for
    t = pi linspace(-1/3, 1, S): a swapped rotation order or a sign error in either
    rotation fails; cropping keeps exactly the samples inside the cube."""
    S = 50
    for xa, za in [(0.3, 1.1), (2.0, 5.0), (0.0, 0.7)]:
        t = np.pi * np.linspace(-1.0 / 3.0, 1.0, S)
        d = np.array([np.sin(xa) * np.sin(za), -np.sin(xa) * np.cos(za), np.cos(xa)])
        sp = O.create_spoke(S, xa, za)
        assert np.abs(sp - d[:, None] * t[None, :]).max() < 1e-12
        c = O.create_spoke(S, xa, za, crop_factor=1.5)
        full = 1.5 * d[:, None] * t[None, :]
        kept = full[:, np.all(np.abs(full) < np.pi, axis=0)]
        assert c.shape == kept.shape and np.abs(c - kept).max() < 1e-12
