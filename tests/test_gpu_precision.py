"""Precision study on the plain single-volume shapes (64^3, 8 coils, kooshball 2000 spokes
x 64 samples, NUFFT tolerance 1e-6): per step, the relative L2 error against the FP64
oracle of the build's path (FP32 data, FP64 Toeplitz-kernel build, FP64 CG scalars) and
of an all-FP32 variant (FP32 PSF NUFFT for the kernel, FP32 CG scalars on the host).

Bounds: NUFFT type-1/type-2 1e-5; normal operator 1e-5 (the FP64-built kernel must also
beat the FP32-built one); CG image after 15 iterations: reported, and the FP64-scalar
path must not be worse than the all-FP32 variant (the kooshball leaves the cube's
corners unsampled, so CG amplifies any FP32 rounding there; the error of both variants
is printed for the record).
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import native

    n, C = 64, 8
    co = O.kooshball(2000, 64)
    co32 = co.astype(np.float32)
    cod = co32.astype(np.float64)
    S = O.coil_maps(C, n, seed=1)
    x_true = O.ellipsoid_phantom(n, seed=2)
    return dict(native=native, n=n, C=C, co32=co32, cod=cod, S=S, x_true=x_true)


def rel(a, b):
    return O.rel_l2(np.asarray(a, dtype=np.complex128), np.asarray(b, dtype=np.complex128))


def test_precision_study(setup):
    N = setup["native"]
    n, C, co32, cod, S, x_true = (setup[k] for k in ("n", "C", "co32", "cod", "S", "x_true"))
    dev = torch.device("cuda:0")
    cog = torch.as_tensor(co32).to(dev)
    nvox = n ** 3
    rows = []

    # NUFFT type-2 / type-1 (scaled as CUDANufft::apply)
    f = (x_true * S[0][None]).astype(np.complex64)
    k_gpu = N.nufft(cog, torch.as_tensor(f).to(dev), [n] * 3, 2, -1, 1e-6, 1 / math.sqrt(nvox)).cpu().numpy()
    k_ref = O.nufft_type2(cod, f[0].astype(np.complex128), -1, 1e-12) / math.sqrt(nvox)
    e2 = rel(k_gpu[0], k_ref)
    rng = np.random.default_rng(3)
    c = (rng.standard_normal(cod.shape[1]) + 1j * rng.standard_normal(cod.shape[1])).astype(np.complex64)
    im_gpu = N.nufft(cog, torch.as_tensor(c[None]).to(dev), [n] * 3, 1, 1, 1e-6, 1 / math.sqrt(nvox)).cpu().numpy()
    im_ref = O.nufft_type1(cod, c.astype(np.complex128), [n] * 3, 1, 1e-12) / math.sqrt(nvox)
    e1 = rel(im_gpu[0], im_ref)
    rows += [("NUFFT type-2", e2, None), ("NUFFT type-1", e1, None)]
    print(f"precision: NUFFT done ({e2:.2e}, {e1:.2e})", flush=True)

    # Toeplitz kernels: FP64 build (the path) and an FP32 PSF (all-FP32 variant)
    _, kern64 = N.toeplitz_build(cog, [n] * 3, tol=1e-8, diag=False)
    L = 2 * n
    ones = torch.ones((1, cog.shape[1]), dtype=torch.complex64, device=dev)
    psf32 = N.nufft(cog, ones, [L] * 3, 1, 1, 1e-6, 1.0)[0] / float(L ** 3)
    psf32[0, :, :] = 0
    psf32[:, 0, :] = 0
    psf32[:, :, 0] = 0
    # the operator's real kernel: Re DFT(ifftshift(psf)) / N^d, psf carrying (2N)^-d
    kern32 = (torch.fft.fftn(torch.fft.ifftshift(psf32)).real / float(nvox)).contiguous()
    diag = O.toeplitz_diagonal(cod, [1, n, n, n])
    Sg = torch.as_tensor(S).to(torch.complex64).to(dev)
    xg = torch.as_tensor(x_true).to(torch.complex64).to(dev)
    ref_normal = O.sense_normal_toeplitz(diag, x_true, S, list(range(C)), workers=O.default_threads())
    op64 = N.SenseNormalToeplitz(Sg, kern64[None].contiguous())
    op32 = N.SenseNormalToeplitz(Sg, kern32[None].contiguous())
    en64 = rel(op64.apply(xg).cpu().numpy(), ref_normal)
    en32 = rel(op32.apply(xg).cpu().numpy(), ref_normal)
    rows += [("normal op", en64, en32)]
    print(f"precision: normal op done ({en64:.2e}, {en32:.2e})", flush=True)

    # CG 15 iterations on b = A^H (A x_true + noise)
    noise = 1e-3 * (rng.standard_normal((C, cod.shape[1])) + 1j * rng.standard_normal((C, cod.shape[1])))
    y = np.stack([O.nufft_type2(cod, (x_true[0] * S[ci]), -1, 1e-12) / math.sqrt(nvox) for ci in range(C)]) + noise
    b = sum(np.conj(S[ci])[None] * O.nufft_type1(cod, y[ci], [n] * 3, 1, 1e-12)[None] / math.sqrt(nvox)
            for ci in range(C))
    ref_cg = O.conjugate_gradient(lambda v: O.sense_normal_toeplitz(diag, v, S, list(range(C)),
                                                                    workers=O.default_threads()),
                                  b, np.zeros_like(b), 15)
    bg = torch.as_tensor(b).to(torch.complex64).to(dev)
    x64, _ = N.cg_solve(lambda v, o: op64.apply(v, out=o), bg, torch.zeros_like(bg), 15, check_pd=False)
    # all-FP32 variant: FP32 kernel, FP32 scalars read back every iteration
    x32 = torch.zeros_like(bg)
    r = bg - op32.apply(x32)
    p = r.clone()
    rz = torch.vdot(r.flatten(), r.flatten()).real
    for _ in range(15):
        Ap = op32.apply(p)
        alpha = rz / torch.vdot(p.flatten(), Ap.flatten()).real
        x32 += alpha * p
        r -= alpha * Ap
        rz_new = torch.vdot(r.flatten(), r.flatten()).real
        p = r + (rz_new / rz) * p
        rz = rz_new
    ec64 = rel(x64.cpu().numpy(), ref_cg)
    ec32 = rel(x32.cpu().numpy(), ref_cg)
    rows += [("CG image (15 it)", ec64, ec32)]

    print("\nprecision vs FP64 oracle (relative L2): step | this path | all-FP32 variant", flush=True)
    for name, a, b_ in rows:
        print(f"  {name:18s} {a:.3e}  {'' if b_ is None else f'{b_:.3e}'}", flush=True)
    assert e1 < 1e-5 and e2 < 1e-5
    assert en64 < 1e-5 and en64 <= en32
    assert ec64 <= ec32 * 1.05
