"""GPU parity of the benchmarked 4D-flow configuration of the fused Toeplitz SENSE normal
operator: 128^3 images, 16 coils, oversampled grid 256^3, several problems with their own
frame kernels in each z-convolution launch (the reference runs one CUDASenseNormal per
frame and encoding, lib/mri/sense.ixx:533-577 inside lib/mri/batch_sense.cpp:383-450).

Each problem is compared with the FP64 oracle's Toeplitz SENSE normal operator on the
GPU-built complex diagonal, so the test pins the fused FFT/convolution pipeline at
L = 256 for every launch position; a second test checks that a problem's 15-iteration
batched CG result is bit-identical alone and inside the 100-problem batch."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_IMG = 128


@pytest.fixture(scope="module")
def N():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import native

    return native


@pytest.fixture(scope="module")
def frames(N):
    """Three kooshball frames of 800, 1150 and 1500 spokes x 128 samples: float32 kernels
    (fused operator input) and the complex diagonals the oracle applies."""
    kern, diags = [], []
    off = 0
    for sp in (800, 1150, 1500):
        co = torch.as_tensor(O.kooshball(sp, 128, offset=off).astype(np.float32)).cuda()
        off += sp
        d, k = N.toeplitz_build(co, [N_IMG] * 3, tol=1e-8)
        kern.append(k)
        torch.cuda.synchronize()
        diags.append(d.cpu().numpy()[None])
        del co, d
    return torch.stack(kern), diags


def _opt(name, value):
    from hastycompute_amd import _capi as C

    prev = C.get_option(name)
    C.set_option(name, value)
    return prev


def _oracle(diag, x, S, coils):
    return O.sense_normal_toeplitz(diag, x.astype(np.complex128), S, coils, workers=O.default_threads())


@pytest.mark.parametrize("case", ["bench", "strided_uneven"])
def test_headline_launch_vs_oracle(N, frames, case):
    from hastycompute_amd import synthetic

    K, diags = frames
    n, C = N_IMG, 16
    if case == "bench":
        # the benchmark's launch shapes: 4 problems per chunk and per z-conv launch,
        # 16 coils in one coil batch; problem 4 is a second chunk (p0 = 4)
        P, pbatch, zbatch, cbatch, stride = 5, 4, 16, 16, 0
        kidx = [0, 1, 2, 1, 0]
        ncoil = None
    else:
        # 6 problems in one chunk split over two z-conv launches (q0 = 0, 4), two coil
        # batches with accumulation, per-problem coil counts and coil-map windows
        P, pbatch, zbatch, cbatch, stride = 6, 6, 4, 8, 4
        kidx = [2, 0, 1, 2, 1, 0]
        ncoil = [16, 9, 16, 12, 5, 16]
    rows = C + stride * (P - 1)
    S = synthetic.coil_maps_torch(rows, n, "cuda", seed=11 + stride)
    g = torch.Generator(device="cuda")
    g.manual_seed(7 + P)
    x = torch.randn((P, n, n, n), dtype=torch.complex64, device="cuda", generator=g)
    sub = torch.randn((P, n, n, n), dtype=torch.complex64, device="cuda", generator=g)
    coils = list(range(C))[::-1] if case == "bench" else list(range(C))
    prev = _opt("sense_zconv_batch", zbatch)
    try:
        op = N.SenseNormalToeplitz(S, K, coils=coils, coil_batch=cbatch, problem_batch=pbatch)
        assert op.layout != 0
        out = op.apply(x, kernel_index=kidx, problem_ncoil=ncoil, coil_problem_stride=stride)
        outs = op.apply(x, kernel_index=kidx, sub=sub, problem_ncoil=ncoil, coil_problem_stride=stride)
        torch.cuda.synchronize()
    finally:
        _opt("sense_zconv_batch", prev)
    Sh = S.cpu().numpy()
    for p in range(P):
        nc = C if ncoil is None else ncoil[p]
        pc = [c + p * stride for c in coils[:nc]]
        ref = _oracle(diags[kidx[p]], x[p:p + 1].cpu().numpy(), Sh, pc)
        e = O.rel_l2(out[p:p + 1].cpu().numpy(), ref)
        es = O.rel_l2(outs[p:p + 1].cpu().numpy(), ref - sub[p:p + 1].cpu().numpy().astype(np.complex128))
        print(f"{case}: problem {p} kernel {kidx[p]} coils {nc}: rel L2 {e:.2e} (with sub {es:.2e})", flush=True)
        assert e < 1e-5 and es < 1e-5


def test_batched_cg_problem_alone_equals_in_batch(N):
    """15 device-CG iterations over the 100 problems of a 5-encoding x 20-frame scan at
    128^3 x 16 coils (4 problems per launch): problems 0, 37 and 99 solved alone give
    the same bits as inside the batch."""
    from hastycompute_amd import synthetic

    n, C, F, E = N_IMG, 16, 20, 5
    rng = np.random.default_rng(2024)
    spokes = rng.integers(800, 1501, F)
    kern, off = [], 0
    for f in range(F):
        co = synthetic.kooshball_torch(int(spokes[f]), 128, off, "cuda")
        off += int(spokes[f])
        _, k = N.toeplitz_build(co, [n] * 3, tol=1e-8, diag=False)
        kern.append(k)
    K = torch.stack(kern)
    del kern
    S = synthetic.coil_maps_torch(C, n, "cuda", seed=3)
    op = N.SenseNormalToeplitz(S, K, coil_batch=16, problem_batch=4)
    P = F * E
    kidx = [p // E for p in range(P)]
    g = torch.Generator(device="cuda")
    g.manual_seed(99)
    b = torch.randn((P, n, n, n), dtype=torch.complex64, device="cuda", generator=g)
    xb, st = N.cg_solve(lambda v, o: op.apply(v, out=o, kernel_index=kidx), b, torch.zeros_like(b), 15)
    assert bool((st[:, 6] == 15).all())
    for p in (0, 37, 99):
        xa, _ = N.cg_solve(lambda v, o: op.apply(v, out=o, kernel_index=[kidx[p]]), b[p:p + 1].clone(),
                           torch.zeros_like(b[p:p + 1]), 15)
        same = torch.equal(xa[0], xb[p])
        print(f"problem {p}: alone == in batch: {same}", flush=True)
        assert same
