"""FP64 CPU oracle for the batched SENSE hot path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker
(or the timed CPU baseline), never as the product. The product path lives in
``hastycompute_amd`` and fails loudly when its HIP library is missing.

Every function restates the reference's algorithm with the reference's conventions
(cited per function) in float64, from definitions:

- NUFFT type-1/type-2: direct sums (``nudft_type1/2``, C, OpenMP) and an ES-kernel
  NUFFT (``nufft_type1/2``: C spreading/interpolation + ``scipy.fft`` + Gauss-Legendre
  deconvolution) that is pinned against the direct sums in the tests.
- Toeplitz PSF/diagonal and apply, SENSE forward/adjoint/normal, CG, SVT, the LLR
  block grid and its libstdc++ shift sequence.

How each function is pinned (what a plausible mistake would fail) is listed in the
tests (tests/test_oracle.py) next to the check.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
import threading
from typing import Callable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None
_lib_lock = threading.Lock()


def build() -> str:
    """Compile the oracle's C/C++ restatement (oracle/Makefile) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            lib = ctypes.CDLL(_LIB_PATH)
            dp = ctypes.POINTER(ctypes.c_double)
            i64p = ctypes.POINTER(ctypes.c_int64)
            for name in ("ora_nudft_type1", "ora_nudft_type2"):
                fn = getattr(lib, name)
                fn.argtypes = [ctypes.c_int, i64p, ctypes.c_int64, dp, dp, dp, ctypes.c_int, ctypes.c_int]
                fn.restype = None
            lib.ora_spread.argtypes = [ctypes.c_int, i64p, ctypes.c_int64, dp, dp, dp, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int]
            lib.ora_interp.argtypes = lib.ora_spread.argtypes
            lib.ora_llr_shifts.argtypes = [ctypes.c_int64, i64p, ctypes.POINTER(ctypes.c_int32)]
            _lib = lib
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i64(vals) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(vals, dtype=np.int64))


def default_threads() -> int:
    return max(1, len(os.sched_getaffinity(0)))


# ---------------------------------------------------------------------------------------
# NUFFT: definitions (direct sums)
# ---------------------------------------------------------------------------------------

def nudft_type2(coords: np.ndarray, f: np.ndarray, isign: int = -1, nthreads: int = 0) -> np.ndarray:
    """c_j = sum_k f_k exp(isign i k.x_j), k_d = n_d - floor(N_d/2); coords [ndim, M].

    Direct sum (the definition FINUFFT approximates, called at lib/fft/nufft.cpp:747-764).
    """
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    ndim, M = coords.shape
    f = np.ascontiguousarray(f, dtype=np.complex128)
    assert f.ndim == ndim
    out = np.empty(M, dtype=np.complex128)
    _load().ora_nudft_type2(ndim, _i64(f.shape).ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), M,
                            _dptr(coords), _dptr(f.view(np.float64)), _dptr(out.view(np.float64)),
                            int(isign), int(nthreads))
    return out


def nudft_type1(coords: np.ndarray, c: np.ndarray, nmodes: Sequence[int], isign: int = 1,
                nthreads: int = 0) -> np.ndarray:
    """f_k = sum_j c_j exp(isign i k.x_j) on the modes nmodes (image dims only)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    ndim, M = coords.shape
    assert len(nmodes) == ndim
    c = np.ascontiguousarray(c, dtype=np.complex128).reshape(M)
    out = np.empty(tuple(nmodes), dtype=np.complex128)
    _load().ora_nudft_type1(ndim, _i64(nmodes).ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), M,
                            _dptr(coords), _dptr(c.view(np.float64)), _dptr(out.view(np.float64)),
                            int(isign), int(nthreads))
    return out


# ---------------------------------------------------------------------------------------
# NUFFT: ES-kernel method (fast FP64 oracle, pinned against the direct sums)
# ---------------------------------------------------------------------------------------

def es_params(tol: float, upsampfac: float = 2.0):
    """Kernel width ns and shape beta for upsampfac 2 (FINUFFT paper, Sec. 4 and its
    defaults): ns = ceil(-log10(tol/10)) clamped to [2, 16]; beta = c(ns)*ns with
    c = 2.20, 2.26, 2.38 for ns = 2, 3, 4 and 2.30 otherwise."""
    assert upsampfac == 2.0
    ns = int(math.ceil(-math.log10(tol / 10.0)))
    ns = max(2, min(16, ns))
    c = {2: 2.20, 3: 2.26, 4: 2.38}.get(ns, 2.30)
    return ns, c * ns


def next235even(n: int) -> int:
    """Smallest even integer >= n whose only prime factors are 2, 3, 5."""
    if n <= 2:
        return 2
    if n % 2:
        n += 1
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 2


def fine_size(n: int, ns: int, upsampfac: float = 2.0) -> int:
    return next235even(max(int(upsampfac * n), 2 * ns))


def es_fourier(nf: int, ns: int, beta: float, n: int) -> np.ndarray:
    """Phi(k) = integral of the ES kernel times cos(2 pi k z / nf) over its support,
    for the modes k = m - n//2, m = 0..n-1 (Gauss-Legendre, 200 nodes on [0, ns/2])."""
    xg, wg = np.polynomial.legendre.leggauss(200)
    z = (xg + 1.0) * (ns / 4.0)
    w = wg * (ns / 4.0)
    u = 2.0 * z / ns
    phi = np.exp(beta * (np.sqrt(np.maximum(1.0 - u * u, 0.0)) - 1.0))
    k = np.arange(n) - n // 2
    return 2.0 * (np.cos(2.0 * np.pi * np.outer(k, z) / nf) @ (w * phi))


def _spread(coords, c, nf, ns, beta, nthreads):
    ndim, M = coords.shape
    grid = np.empty(tuple(nf), dtype=np.complex128)
    _load().ora_spread(ndim, _i64(nf).ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), M, _dptr(coords),
                       _dptr(c.view(np.float64)), _dptr(grid.view(np.float64)), ns, beta, nthreads)
    return grid


def _interp(coords, grid, ns, beta, nthreads):
    ndim, M = coords.shape
    out = np.empty(M, dtype=np.complex128)
    _load().ora_interp(ndim, _i64(grid.shape).ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), M,
                       _dptr(coords), _dptr(np.ascontiguousarray(grid).view(np.float64)),
                       _dptr(out.view(np.float64)), ns, beta, nthreads)
    return out


def _mode_index(n: int, nf: int) -> np.ndarray:
    return (np.arange(n) - n // 2) % nf


def nufft_type1(coords: np.ndarray, c: np.ndarray, nmodes: Sequence[int], isign: int = 1,
                tol: float = 1e-12, nthreads: int = 0) -> np.ndarray:
    """ES-kernel type-1 (spread, FFT, deconvolve); same definition as nudft_type1."""
    import scipy.fft as sfft
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    ndim, M = coords.shape
    ns, beta = es_params(tol)
    nf = [fine_size(n, ns) for n in nmodes]
    nthreads = nthreads or default_threads()
    grid = _spread(coords, np.ascontiguousarray(c, dtype=np.complex128).reshape(M), nf, ns, beta, nthreads)
    # exp(+i k l h) sums are an unnormalised inverse DFT; exp(-i ...) the forward one.
    if isign > 0:
        G = sfft.ifftn(grid, workers=nthreads) * np.prod(nf)
    else:
        G = sfft.fftn(grid, workers=nthreads)
    idx = np.ix_(*[_mode_index(n, m) for n, m in zip(nmodes, nf)])
    out = G[idx]
    for d in range(ndim):
        phi = es_fourier(nf[d], ns, beta, nmodes[d])
        shp = [1] * ndim
        shp[d] = nmodes[d]
        out = out / phi.reshape(shp)
    return out


def nufft_type2(coords: np.ndarray, f: np.ndarray, isign: int = -1, tol: float = 1e-12,
                nthreads: int = 0) -> np.ndarray:
    """ES-kernel type-2 (deconvolve, FFT, interpolate); same definition as nudft_type2."""
    import scipy.fft as sfft
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    ndim, M = coords.shape
    nmodes = f.shape
    ns, beta = es_params(tol)
    nf = [fine_size(n, ns) for n in nmodes]
    nthreads = nthreads or default_threads()
    fc = np.array(f, dtype=np.complex128)
    for d in range(ndim):
        phi = es_fourier(nf[d], ns, beta, nmodes[d])
        shp = [1] * ndim
        shp[d] = nmodes[d]
        fc = fc / phi.reshape(shp)
    grid = np.zeros(tuple(nf), dtype=np.complex128)
    grid[np.ix_(*[_mode_index(n, m) for n, m in zip(nmodes, nf)])] = fc
    if isign < 0:
        G = sfft.fftn(grid, workers=nthreads)
    else:
        G = sfft.ifftn(grid, workers=nthreads) * np.prod(nf)
    return _interp(coords, np.ascontiguousarray(G), ns, beta, nthreads)


# ---------------------------------------------------------------------------------------
# Reference-semantics wrappers (lib/fft/nufft.cpp)
# ---------------------------------------------------------------------------------------

def nufft_apply(coords: np.ndarray, nmodes: Sequence[int], nufft_type: int, x: np.ndarray,
                positive: Optional[bool] = None, exact: bool = False, tol: float = 1e-12) -> np.ndarray:
    """hasty::fft::(CUDA)Nufft::apply (nufft.cpp:104-123, 487-506): the (cu)finufft
    transform over ntransf = nmodes[0] vectors, then division by sqrt(prod nmodes[1:]).
    Default signs follow NufftOptions::type1()/type2() (nufft.ixx:40-42): type-1
    positive (+1), type-2 negative (-1). x is [ntransf, M] (type 1) or nmodes (type 2)."""
    ntransf = int(nmodes[0])
    img = list(nmodes[1:])
    nvox = float(np.prod(img))
    if positive is None:
        positive = nufft_type == 1
    isign = 1 if positive else -1
    outs = []
    for t in range(ntransf):
        if nufft_type == 1:
            fn = nudft_type1 if exact else (lambda co, cc, nm, isign: nufft_type1(co, cc, nm, isign, tol))
            outs.append(fn(coords, x[t], img, isign))
        elif nufft_type == 2:
            fn = nudft_type2 if exact else (lambda co, ff, isign: nufft_type2(co, ff, isign, tol))
            outs.append(fn(coords, x[t], isign))
        else:
            raise RuntimeError("Only Nufft Type 1 and Type 2 is supported")
    return np.stack(outs) / math.sqrt(nvox)


def nufft_normal_apply(coords: np.ndarray, nmodes: Sequence[int], x: np.ndarray,
                       func_between: Optional[Callable[[np.ndarray], np.ndarray]] = None,
                       exact: bool = False) -> np.ndarray:
    """hasty::fft::CUDANufftNormal::apply (nufft.cpp:862-869): type-2, optional
    func_between on the [ntransf, M] storage, then type-1, each scaled as nufft_apply."""
    k = nufft_apply(coords, nmodes, 2, x, exact=exact)
    if func_between is not None:
        k = func_between(k)
    return nufft_apply(coords, nmodes, 1, k, exact=exact)


def toeplitz_psf(coords: np.ndarray, nmodes: Sequence[int], weights: Optional[np.ndarray] = None,
                 exact: bool = False) -> np.ndarray:
    """The PSF the reference builds in build_diagonal_base (nufft.cpp:932-950): a delta at
    mode index nmodes_ns[d]/2 (frequency 0) on the doubled grid, pushed through the
    explicit normal operator with the doubled grid's scaling. In exact arithmetic this is
    psf[n] = (2N)^-d sum_j w_j exp(+i (n - N).x_j), returned as [ntransf, 2N, ...]."""
    img = [2 * int(n) for n in nmodes[1:]]
    M = coords.shape[1]
    c = np.ones(M, dtype=np.complex128) if weights is None else np.asarray(weights, dtype=np.complex128)
    if exact:
        p = nudft_type1(coords, c, img, 1)
    else:
        p = nufft_type1(coords, c, img, 1)
    p = p / float(np.prod(img))
    return np.broadcast_to(p, (int(nmodes[0]),) + tuple(img)).copy()


def toeplitz_diagonal(coords: np.ndarray, nmodes: Sequence[int], weights: Optional[np.ndarray] = None,
                      exact: bool = False) -> np.ndarray:
    """CUDANormalNufftToeplitz::build_diagonal (nufft.cpp:932-1053): psf, ifftshift over
    the image dims, unnormalised fftn."""
    p = toeplitz_psf(coords, nmodes, weights, exact)
    axes = tuple(range(1, p.ndim))
    return np.fft.fftn(np.fft.ifftshift(p, axes=axes), axes=axes)


def toeplitz_apply(diag: np.ndarray, x: np.ndarray, workers: Optional[int] = None) -> np.ndarray:
    """CUDANormalNufftToeplitz::apply (nufft.cpp:1163-1170): fftn(x, s=2N) (zero pad at
    the end of each dim), multiply by diag, ifftn ('backward', 1/prod(2N)), crop [:N]."""
    axes = tuple(range(1, x.ndim))
    s = [2 * n for n in x.shape[1:]]
    if workers:
        import scipy.fft as sfft
        y = sfft.ifftn(sfft.fftn(x, s=s, axes=axes, workers=workers) * diag, axes=axes, workers=workers)
    else:
        y = np.fft.ifftn(np.fft.fftn(x, s=s, axes=axes) * diag, axes=axes)
    return y[(slice(None),) + tuple(slice(0, n) for n in x.shape[1:])]


# ---------------------------------------------------------------------------------------
# SENSE (lib/mri/sense.ixx)
# ---------------------------------------------------------------------------------------

def sense_apply(coords, x, smaps, coils, exact=False):
    """sense_apply with ntransf=1 and no manipulators (sense.ixx:348-423): out[i] =
    N2(x * smaps[i]) for the i-th listed coil. The reference multiplies smaps[idx] (the
    list position, sense.ixx:395), not smaps[coils[idx]]; with the default coil list
    0..C-1 both agree, and this restatement uses the list position as the reference does."""
    nm = [1] + list(smaps.shape[1:])
    return np.concatenate([nufft_apply(coords, nm, 2, x[0:1] * smaps[i][None], exact=exact)
                           for i, _ in enumerate(coils)])


def sense_adjoint_apply(coords, y, smaps, coils, exact=False):
    """sense_adjoint_apply with an accumulating [1, ...] out (sense.ixx:426-497):
    out = sum_i conj(smaps[coils[i]]) * N1(y[i])."""
    nm = [1] + list(smaps.shape[1:])
    out = np.zeros(nm, dtype=np.complex128)
    for i, c in enumerate(coils):
        out += np.conj(smaps[c])[None] * nufft_apply(coords, nm, 1, y[i:i + 1], exact=exact)
    return out


def sense_normal_apply(coords, x, smaps, coils, kdata=None, weights=None, exact=False):
    """sense_normal_apply (sense.ixx:500-578) with the batched midmanip of
    BatchedSenseNormal (batch_sense.cpp:452-486): for the i-th listed coil,
    k = N2(x * S_c); k = (k - kdata[i]) * weights (each part only if given);
    out += conj(S_c) * N1(k)."""
    nm = [1] + list(smaps.shape[1:])
    out = np.zeros(nm, dtype=np.complex128)
    for i, c in enumerate(coils):
        k = nufft_apply(coords, nm, 2, x[0:1] * smaps[c][None], exact=exact)
        if kdata is not None:
            k = k - kdata[i][None]
        if weights is not None:
            k = k * weights.reshape(1, -1)
        out += np.conj(smaps[c])[None] * nufft_apply(coords, nm, 1, k, exact=exact)
    return out


def sense_normal_toeplitz(diag, x, smaps, coils, workers: Optional[int] = None):
    """The Toeplitz-embedded SENSE normal operator at the scale of sense_normal_apply:
    sum_c conj(S_c) * 2^d * toeplitz_apply(diag, S_c x). The 2^d undoes the doubled-grid
    scaling of the reference's PSF (nufft.cpp:932-971 vs 862-880)."""
    d = x.ndim - 1
    out = np.zeros_like(x, dtype=np.complex128)
    for c in coils:
        out += np.conj(smaps[c])[None] * (2 ** d) * toeplitz_apply(diag, x * smaps[c][None], workers)
    return out


# ---------------------------------------------------------------------------------------
# Circulant preconditioner (lib/mri/precond.cpp:27-117)
# ---------------------------------------------------------------------------------------

def fftc(x: np.ndarray, norm: Optional[str] = "ortho") -> np.ndarray:
    """hasty::fft::fftc (lib/fft/fft.cpp:7-18): fftshift(fftn(ifftshift(x))) over all
    axes, norm "ortho" by default (fft.ixx:12-13)."""
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x), norm=norm))


def ifftc(x: np.ndarray, norm: Optional[str] = "ortho") -> np.ndarray:
    """hasty::fft::ifftc (fft.cpp:20-32)."""
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x), norm=norm))


def circulant_precond_diagonal(smaps: np.ndarray, coords: np.ndarray, weights: Optional[np.ndarray] = None,
                               lam: float = 0.0, exact: bool = False, tol: float = 1e-12) -> np.ndarray:
    """CirculantPreconditionerOp::build_diagonal (precond.cpp:65-117, "based on sigpy":
    sigpy.mri.circulant_precond): psf = NUFFTAdjoint of sqrt(w) (ones without weights) onto
    the doubled modes (nufft_apply scaling 1/sqrt(prod 2N)); per coil
    xcorr = ifftc(|fftc(resize(conj S_c, 2N))|^2) * psf, p_c = fftc(xcorr)[::2, ...] * scale
    with scale = prod(2N)^1.5 / prod(N)^2; p = sum_c p_c + lambda; p == 0 -> 1; 1 / p.
    resize puts the map at the low corner (torch_util.cpp:240-253). Centred order.
    The reference allocates p with the sample count as its shape and resizes the map with
    its coil axis included (precond.cpp:94, 100-102), which raise at run time, and its
    second multiply by sqrt(w) (precond.cpp:108-109) cannot broadcast against the image;
    this follows the intended image-shaped sum, as sigpy does."""
    img = list(smaps.shape[1:])
    ndim = len(img)
    img2 = [2 * n for n in img]
    scale = float(np.prod(img2)) ** 1.5 / float(np.prod(img)) ** 2
    ones = np.ones(coords.shape[1], dtype=np.complex128)
    if weights is not None:
        ones = ones * np.sqrt(np.asarray(weights, dtype=np.float64))
    psf = nufft_apply(coords, [1] + img2, 1, ones[None], exact=exact, tol=tol)[0]
    idx = (slice(None, None, 2),) * ndim
    corner = tuple(slice(0, n) for n in img)
    p_inv = np.zeros(img, dtype=np.complex128)
    for S in smaps:
        padded = np.zeros(img2, dtype=np.complex128)
        padded[corner] = np.conj(S)
        xcorr = ifftc(np.abs(fftc(padded)) ** 2)
        xcorr = xcorr * psf
        p_inv += fftc(xcorr)[idx] * scale
    p_inv += lam
    p_inv[p_inv == 0] = 1.0
    return 1.0 / p_inv


def circulant_precond_apply(diag: np.ndarray, x: np.ndarray, centered: bool = True,
                            adjoint: bool = False) -> np.ndarray:
    """CirculantPreconditionerOp::apply (precond.cpp:27-50) without a norm: ifftc(d fftc(x))
    (centred) or ifftn(d fftn(x)); adjointify: fftc(conj(d) ifftc(x)) / fftn(conj(d) ifftn(x)).
    x is one image [N...]."""
    if centered:
        if adjoint:
            return fftc(np.conj(diag) * ifftc(x, None), None)
        return ifftc(diag * fftc(x, None), None)
    if adjoint:
        return np.fft.fftn(np.conj(diag) * np.fft.ifftn(x))
    return np.fft.ifftn(diag * np.fft.fftn(x))


# ---------------------------------------------------------------------------------------
# Conjugate gradient (lib/op/opalgs.cpp:50-96)
# ---------------------------------------------------------------------------------------

class NotPositiveDefinite(RuntimeError):
    pass


def conjugate_gradient(A: Callable[[np.ndarray], np.ndarray], b: np.ndarray, x: np.ndarray,
                       iters: int, tol: float = 0.0, P: Optional[Callable] = None) -> np.ndarray:
    """ConjugateGradient::run: r = b - A x; z = P r; p = z; rzold = Re<r,z>; per iteration:
    stop if sqrt(rzold) < tol; Ap = A p; pAp = Re<p,Ap> (<= 0 raises 'A was not positive
    definite'); alpha = rzold/pAp; x += alpha p; r -= alpha Ap; z = P r; rznew = Re<r,z>;
    p = z + (rznew/rzold) p; rzold = rznew."""
    x = np.array(x, dtype=np.complex128)
    if iters < 1:
        return x
    r = b - A(x)
    z = P(r) if P is not None else r
    p = z.copy()
    rzold = float(np.real(np.vdot(r, z)))
    resid = math.sqrt(rzold)
    for _ in range(iters):
        if resid < tol:
            return x
        Ap = A(p)
        pAp = float(np.real(np.vdot(p, Ap)))
        if pAp <= 0.0:
            raise NotPositiveDefinite("A was not positive definite")
        alpha = rzold / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        z = P(r) if P is not None else r
        rznew = float(np.real(np.vdot(r, z)))
        p = p * (rznew / rzold) + z
        rzold = rznew
        resid = math.sqrt(rzold)
    return x


def power_iteration(A: Callable[[np.ndarray], np.ndarray], x0: np.ndarray, iters: int) -> float:
    """MaxEig (python/python/hastypy/base/opalg.py:647-674; op::PowerIteration,
    lib/op/opalgs.cpp:29-44): per iteration y = A x, lambda = |y|_2, x = y / lambda."""
    x = np.array(x0, dtype=np.complex128)
    lam = float("inf")
    for _ in range(iters):
        y = A(x)
        lam = float(np.linalg.norm(y))
        x = y / lam
    return lam


def gradient_descent(gradf: Callable[[np.ndarray], np.ndarray], x0: np.ndarray,
                     prox: Optional[Callable[[np.ndarray], np.ndarray]], alpha: float, iters: int,
                     accelerate: bool = True) -> np.ndarray:
    """GradientDescent.run (python/python/hastypy/base/solvers.py:85-162): per iteration
    x_old = x; (accelerated: x = z); x = x - alpha grad f(x); x = prox(x); accelerated:
    t' = (1 + sqrt(1 + 4 t^2)) / 2, z = x + ((t - 1) / t') (x - x_old). The prox takes
    no step size: ProximalOperator.__call__ drops it (proximal.py:39-40)."""
    x = np.array(x0, dtype=np.complex128)
    z = x.copy()
    t = 1.0
    for _ in range(iters):
        x_old = x.copy()
        if accelerate:
            x = z
        x = x - alpha * gradf(x)
        if prox is not None:
            x = prox(x)
        if accelerate:
            t_old = t
            t = (1.0 + math.sqrt(1.0 + 4.0 * t_old * t_old)) / 2.0
            z = x + ((t_old - 1.0) / t) * (x - x_old)
    return x


def pdhg(K: Callable, KH: Callable, proxfc: Callable, proxg: Callable, x0: np.ndarray, y0: np.ndarray,
         tau: float, sigma: float, acceleration: float, iters: int, classic: bool = False):
    """PrimalDualHybridGradient._update (python/python/hastypy/base/solvers.py:191-218) in
    the reference's order: y += sigma K(x_bar); y = proxfc(y, sigma); x -= tau K^H(y);
    x = proxg(x, tau); theta = 1/sqrt(1 + 2 acceleration tau_min), tau *= theta and
    sigma /= theta unless theta == 1; x_bar = x + theta (x - x_bar_previous) (the
    previous x_bar: solvers.py:210-215, the stored x_old is unused). classic=True
    extrapolates from the previous x (Chambolle-Pock). Returns (x, y)."""
    x = np.array(x0, dtype=np.complex128)
    y = np.array(y0, dtype=np.complex128)
    xbar = x.copy()
    tau_mul = float(tau)
    for _ in range(iters):
        y = proxfc(y + sigma * K(xbar), sigma)
        xold = x
        x = proxg(x - tau * KH(y), tau)
        theta = 1.0 / math.sqrt(1.0 + 2.0 * acceleration * tau_mul)
        if not math.isclose(theta, 1.0):
            tau_mul *= theta
            tau *= theta
            sigma /= theta
        xbar = x + theta * (x - (xold if classic else xbar))
    return x, y


def five_point_llr(coords_list, smaps, kdata_list, weights_list, x0, stepsize: float, iters: int,
                   block_strides, block_shape, block_iter: int, thresh: float, soft: bool,
                   accelerate: bool = True) -> np.ndarray:
    """FivePointLLR.run with solver 'GD' (python/python/hastypy/base/recon.py:87-205):
    x0 [F, 5, N...]; the gradient of problem p = 5 f + e is sense_normal_apply with
    kdata/weights (BatchedSenseNormal on the [F*5, 1, N...] view); the prox is
    normal_3d_blocks_svt on the [F, 5, N...] view."""
    F = x0.shape[0]
    sp = x0.shape[2:]
    coils = list(range(smaps.shape[0]))

    def grad(xg):
        out = np.empty_like(xg)
        for p in range(F * 5):
            w = None if weights_list is None else weights_list[p]
            kd = None if kdata_list is None else kdata_list[p][0]
            out[p] = sense_normal_apply(coords_list[p], xg[p], smaps, coils, kdata=kd, weights=w)
        return out

    def prox(xg):
        return normal_3d_blocks_svt(xg.reshape((F, 5) + sp), block_strides, block_shape, block_iter, thresh,
                                    soft).reshape((F * 5, 1) + sp)

    x = gradient_descent(grad, np.asarray(x0, dtype=np.complex128).reshape((F * 5, 1) + sp), prox, stepsize,
                         iters, accelerate)
    return x.reshape((F, 5) + sp)


def create_spoke(samp_per_spoke: int, xangle: float, zangle: float, crop_factor: float = 1.0,
                 jitter: Optional[np.ndarray] = None) -> np.ndarray:
    """create_spoke, method 'PCVIPR' (python/python/hastypy/util/image_creation.py:13-57),
    with the angles given and the Gaussian jitter passed in ([3, S], zero if None): the
    line k_z = pi linspace(-1/3, 1, S) plus jitter, rotated by Rz(zangle) Rx(xangle), then
    scaled by crop_factor and cut to |k_d| < pi when crop_factor != 1."""
    S = samp_per_spoke
    spoke = np.zeros((3, S))
    if jitter is not None:
        spoke += jitter
    spoke[2] += np.pi * np.linspace(-1.0 / 3.0, 1.0, S)
    cx, sx, cz, sz = math.cos(xangle), math.sin(xangle), math.cos(zangle), math.sin(zangle)
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cx, -sx], [0.0, sx, cx]])
    rz = np.array([[cz, -sz, 0.0], [sz, cz, 0.0], [0.0, 0.0, 1.0]])
    spoke = rz @ rx @ spoke
    if crop_factor != 1.0:
        spoke = spoke * crop_factor
        spoke = spoke[:, np.all(np.abs(spoke) < np.pi, axis=0)]
    return spoke


def l1dct_prox(x: np.ndarray, lam: float, axis: Sequence[bool]) -> np.ndarray:
    """ProximalL1DCT.tensor_apply (python/python/hastypy/base/proximal.py:139-163): over
    every trailing volume of len(axis) dims, dctn over the axes marked True (scipy's
    unnormalised DCT-II, norm=None, what cupyx.scipy.fft.dctn computes), then
    exp(i angle c) max(|c| - lam, 0), then idctn."""
    import scipy.fft as sfft
    x = np.array(x, dtype=np.complex128)
    nd = len(axis)
    axes = tuple(x.ndim - nd + i for i, a in enumerate(axis) if a)
    c = sfft.dctn(x, axes=axes)
    c = np.exp(1j * np.angle(c)) * np.maximum(np.abs(c) - lam, 0.0)
    return sfft.idctn(c, axes=axes)


def five_point_full(coords_list, smaps, kdata_list, weights_list, x0, max_eig: float, lamda: float, iters: int,
                    accelerate: bool = True) -> np.ndarray:
    """FivePointFULL.run with solver 'GD' (python/python/hastypy/base/recon.py:47-83) on the
    5 stacked encodings: x0 [5, 1, N...], step 1 / max_eig, gradient sense_normal_apply with
    kdata/weights, prox l1dct_prox with threshold lamda * max_eig (lamda / step) when
    lamda != 0."""
    coils = list(range(smaps.shape[0]))

    def grad(xg):
        out = np.empty_like(xg)
        for p in range(xg.shape[0]):
            w = None if weights_list is None else weights_list[p]
            kd = None if kdata_list is None else kdata_list[p][0]
            out[p] = sense_normal_apply(coords_list[p], xg[p], smaps, coils, kdata=kd, weights=w)
        return out

    th = lamda * max_eig
    prox = (lambda v: l1dct_prox(v, th, (True, True, True))) if lamda != 0.0 else None
    return gradient_descent(grad, np.asarray(x0, dtype=np.complex128), prox, 1.0 / max_eig, iters, accelerate)


# ---------------------------------------------------------------------------------------
# SVT / LLR (lib/op/svt.cpp, lib/op/svt.ixx)
# ---------------------------------------------------------------------------------------

def svt_soft(A: np.ndarray, lam: float) -> np.ndarray:
    """svt_soft (svt.cpp:117-144): U diag(max(|S|-lam, 0) sign S) Vh from the thin SVD."""
    U, S, Vh = np.linalg.svd(A, full_matrices=False)
    ldiff = np.abs(S) - lam
    S2 = np.sign(S) * np.abs(0.5 * (np.abs(ldiff) + ldiff))
    return (U * S2[None, :]) @ Vh


def svt_hard(A: np.ndarray, rank: int) -> np.ndarray:
    """svt_hard (svt.cpp:62-88): keep the first `rank` singular values, zero the rest."""
    U, S, Vh = np.linalg.svd(A, full_matrices=False)
    S = S.copy()
    S[rank:] = 0.0
    return (U * S[None, :]) @ Vh


_MINSTD_M = 2147483647
_MINSTD_A = 16807


def llr_shifts(block_iter: int, block_shape: Sequence[int]) -> np.ndarray:
    """The reference's LLR shift sequence (svt.cpp:378-391) restated: libstdc++'s
    std::default_random_engine is minstd_rand0 (x <- 16807 x mod 2^31-1, seed 1, range
    [1, 2^31-2]); std::uniform_int_distribution<int>(0, b) draws by downscaling:
    scaling = urngrange // (b+1), reject r - 1 >= (b+1)*scaling, return (r-1)//scaling.
    Returns int32 [3, block_iter], shifts[d][0] = 0."""
    state = 1
    urngrange = (_MINSTD_M - 1) - 1 - 1  # max() - min() = (2^31-2) - 1
    out = np.zeros((3, block_iter), dtype=np.int32)
    for d in range(3):
        if block_iter > 1:
            uerange = int(block_shape[d]) + 1
            scaling = urngrange // uerange
            past = uerange * scaling
            for b in range(1, block_iter):
                while True:
                    state = (_MINSTD_A * state) % _MINSTD_M
                    r = state - 1
                    if r < past:
                        break
                out[d, b] = r // scaling
    return out


def llr_shifts_libstdcxx(block_iter: int, block_shape: Sequence[int]) -> np.ndarray:
    """The same sequence drawn by libstdc++ itself (oracle/llr_shifts.cpp)."""
    out = np.zeros((3, block_iter), dtype=np.int32)
    _load().ora_llr_shifts(int(block_iter), _i64(block_shape).ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out


def normal_blocks(shape_xyz: Sequence[int], block_strides: Sequence[int], block_shape: Sequence[int],
                  block_iter: int, shifts: Optional[np.ndarray] = None):
    """Block corners in the order OldNormal3DBlocksSVT::apply enqueues them
    (svt.cpp:393-436): per iteration, nx, ny, nz loops over S/stride blocks starting at
    n*stride + shift, skipping any block whose end is >= the size (so the last row is
    never processed, svt.cpp:404). Returns a list (per iteration) of [(sx,sy,sz), ...]."""
    if block_iter < 1:
        raise RuntimeError("Can't perform Normal3DBlocksSVT with block_iter < 1")
    if shifts is None:
        shifts = llr_shifts(block_iter, block_shape)
    Sx, Sy, Sz = shape_xyz
    bx, by, bz = Sx // block_strides[0], Sy // block_strides[1], Sz // block_strides[2]
    iters = []
    for it in range(block_iter):
        blocks = []
        for nx in range(bx):
            sx = nx * block_strides[0] + int(shifts[0, it])
            if sx + block_shape[0] >= Sx:
                continue
            for ny in range(by):
                sy = ny * block_strides[1] + int(shifts[1, it])
                if sy + block_shape[1] >= Sy:
                    continue
                for nz in range(bz):
                    sz = nz * block_strides[2] + int(shifts[2, it])
                    if sz + block_shape[2] >= Sz:
                        continue
                    blocks.append((sx, sy, sz))
        iters.append(blocks)
    return iters


def extract_block(x: np.ndarray, corner, block_shape) -> np.ndarray:
    """extract_block(Block<3>, perms=None, flatten_split=0, transpose=true)
    (svt.cpp:37-60): x[..., block] flattened to [F, E*bx*by*bz] and transposed, giving
    the Casorati matrix with rows (e, x, y, z) and columns f."""
    sx, sy, sz = corner
    b = x[:, :, sx:sx + block_shape[0], sy:sy + block_shape[1], sz:sz + block_shape[2]]
    return b.reshape(b.shape[0], -1).T


def insert_block(x: np.ndarray, corner, block_shape, casorati: np.ndarray) -> None:
    """insert_block, the inverse of extract_block (svt.cpp:208-243)."""
    sx, sy, sz = corner
    F, E = x.shape[0], x.shape[1]
    x[:, :, sx:sx + block_shape[0], sy:sy + block_shape[1], sz:sz + block_shape[2]] = \
        casorati.T.reshape(F, E, *block_shape)


def normal_3d_blocks_svt(x: np.ndarray, block_strides, block_shape, block_iter: int, thresh: float,
                         soft: bool) -> np.ndarray:
    """OldNormal3DBlocksSVT::apply (svt.cpp:354-449; the TorchScript target at
    python/cpp/py_svt.cpp:47-50): blocks applied sequentially in enqueue order; each
    block is replaced by svt_soft(block, thresh) or svt_hard(block, int(thresh+0.5))."""
    x = np.array(x, dtype=np.complex128)
    for blocks in normal_blocks(x.shape[2:], block_strides, block_shape, block_iter):
        for corner in blocks:
            A = extract_block(x, corner, block_shape)
            A2 = svt_soft(A, thresh) if soft else svt_hard(A, int(thresh + 0.5))
            insert_block(x, corner, block_shape, A2)
    return x


# ---------------------------------------------------------------------------------------
# Synthetic inputs (deterministic, seeded) shared by tests, smoke and bench
# ---------------------------------------------------------------------------------------

GOLDEN_MEANS = (0.4656, 0.6823)


def kooshball(nspokes: int, nsamp: int, offset: int = 0, dtype=np.float64) -> np.ndarray:
    """3D radial kooshball, golden-means directions: spoke n has u_z = frac(n phi1),
    theta = 2 pi frac(n phi2), samples k in [-pi, pi) along the direction. Returns
    [3, nspokes*nsamp] (row 0 pairs with image dim 0)."""
    phi1, phi2 = 0.465571231876768, 0.682327803828019
    n = np.arange(offset, offset + nspokes, dtype=np.float64)
    uz = np.mod(n * phi1, 1.0) * 2.0 - 1.0
    th = 2.0 * np.pi * np.mod(n * phi2, 1.0)
    r = np.sqrt(np.maximum(1.0 - uz * uz, 0.0))
    dirs = np.stack([uz, r * np.cos(th), r * np.sin(th)])  # [3, S]
    t = -np.pi + 2.0 * np.pi * np.arange(nsamp) / nsamp
    return (dirs[:, :, None] * t[None, None, :]).reshape(3, -1).astype(dtype)


def coil_maps(ncoils: int, n: int, seed: int = 0) -> np.ndarray:
    """Smooth Gaussian-magnitude x linear-phase coil maps normalised so sum_c |S_c|^2 = 1."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n) - n / 2) / n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    maps = []
    for c in range(ncoils):
        ctr = rng.uniform(-0.5, 0.5, 3)
        w = rng.uniform(0.3, 0.6)
        mag = np.exp(-((Z - ctr[0]) ** 2 + (Y - ctr[1]) ** 2 + (X - ctr[2]) ** 2) / (2 * w * w))
        kv = rng.uniform(-3, 3, 3)
        maps.append(mag * np.exp(1j * (kv[0] * Z + kv[1] * Y + kv[2] * X)))
    S = np.stack(maps)
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=0, keepdims=True))
    return S


def ellipsoid_phantom(n: int, seed: int = 0, nell: int = 6) -> np.ndarray:
    """Seeded random-ellipsoid phantom (complex, [1, n, n, n])."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n) - n / 2) / (n / 2)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    img = np.zeros((n, n, n), dtype=np.complex128)
    for _ in range(nell):
        c = rng.uniform(-0.4, 0.4, 3)
        a = rng.uniform(0.1, 0.5, 3)
        v = rng.uniform(0.2, 1.0) * np.exp(1j * rng.uniform(0, 2 * np.pi))
        mask = ((Z - c[0]) / a[0]) ** 2 + ((Y - c[1]) / a[1]) ** 2 + ((X - c[2]) / a[2]) ** 2 <= 1.0
        img[mask] += v
    return img[None]


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
