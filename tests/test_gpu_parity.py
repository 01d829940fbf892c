"""GPU parity: the HIP path through the C ABI against the FP64 oracle on seeded inputs.

Tolerances (relative L2 vs FP64): NUFFT 1e-5 at tol 1e-6 kernels, Toeplitz/SENSE normal
operator 1e-5, CG image 1e-5 (against the oracle's CG on the same operator), LLR 1e-4.
"""
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")


@pytest.fixture(scope="module")
def N():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import native

    return native


def cx(d):
    return (np.asarray(d["re"]) + 1j * np.asarray(d["im"])).reshape(d["shape"])


def g(a, dtype=torch.complex64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def h(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128)


def rnd(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# (64, 5, 7): the dim-0 pass is a strided 128-point register FFT over 35 lines, so its
# last bundle of 16 holds 3
@pytest.mark.parametrize("dims", [(10, 9, 8), (12, 7), (16,), (6, 5, 4), (64, 5, 7)])
@pytest.mark.parametrize("fp64", [False, True])
def test_nufft_small_vs_direct_sum(N, dims, fp64):
    rng = np.random.default_rng(len(dims) * 10 + fp64)
    M = 400
    co = rng.uniform(-np.pi, np.pi, (len(dims), M))
    f = rnd(rng, dims)
    c = rnd(rng, M)
    tol = 1e-10 if fp64 else 1e-6
    rt, ct = (torch.float64, torch.complex128) if fp64 else (torch.float32, torch.complex64)
    cog = g(co, rt)
    co_used = co.astype(np.float32).astype(np.float64) if not fp64 else co
    out2 = h(N.nufft(cog, g(f[None], ct), dims, 2, -1, tol))[0]
    out1 = h(N.nufft(cog, g(c[None], ct), dims, 1, +1, tol))[0]
    e2 = O.rel_l2(out2, O.nudft_type2(co_used, f, -1))
    e1 = O.rel_l2(out1, O.nudft_type1(co_used, c, dims, +1))
    bound = 1e-8 if fp64 else 1e-5
    print(f"dims={dims} fp64={fp64} type2 {e2:.2e} type1 {e1:.2e}", flush=True)
    assert e2 < bound and e1 < bound
    # the other sign
    out2p = h(N.nufft(cog, g(f[None], ct), dims, 2, +1, tol))[0]
    assert O.rel_l2(out2p, O.nudft_type2(co_used, f, +1)) < bound


def test_golden_deterministic_3d(N):
    fx = json.load(open(GOLDEN))["deterministic_3d"]
    co = g(np.asarray(fx["coords"], dtype=np.float32), torch.float32)
    nm = fx["nmodes"]
    k = N.nufft(co, g(cx(fx["input"])), nm[1:], 2, -1, 1e-6, scale=1 / math.sqrt(64))
    assert O.rel_l2(h(k), cx(fx["type2"])) < 1e-5
    back = N.nufft(co, k, nm[1:], 1, +1, 1e-6, scale=1 / math.sqrt(64))
    assert O.rel_l2(h(back), cx(fx["type1_of_type2"])) < 1e-5


def test_golden_deterministic_1d(N):
    fx = json.load(open(GOLDEN))["deterministic_1d"]
    co = g(np.asarray(fx["coords"], dtype=np.float32), torch.float32)
    out = N.nufft(co, g(cx(fx["input"])), [8], 2, +1, 1e-6, scale=1 / math.sqrt(8))
    ref = cx(fx["type2"])
    assert np.abs(h(out) - ref).max() < 1e-5 * np.abs(ref).max()


def test_golden_compare_normal_methods(N):
    fx = json.load(open(GOLDEN))["compare_normal_methods"]
    co = g(np.asarray(fx["coords"], dtype=np.float32), torch.float32)
    nm = fx["nmodes"]
    diag, kern = N.toeplitz_build(co, nm[1:])
    assert O.rel_l2(h(diag)[None], cx(fx["toeplitz_diagonal"])) < 1e-5
    ones = torch.ones(nm, dtype=torch.complex64, device="cuda")
    out3 = N.toeplitz_apply(diag[None].contiguous(), ones)
    assert O.rel_l2(h(out3), cx(fx["toeplitz"])) < 1e-5
    # SENSE normal with one unit coil = the explicit normal (2^d times the Toeplitz result)
    S = torch.ones((1, *nm[1:]), dtype=torch.complex64, device="cuda")
    op = N.SenseNormalToeplitz(S, kern)
    out = op.apply(ones)
    assert O.rel_l2(h(out), cx(fx["normal"])) < 1e-5


def test_nufft_cfg1_sizes_vs_oracle(N):
    rng = np.random.default_rng(7)
    co = O.kooshball(2000, 64)
    x = O.ellipsoid_phantom(64, seed=3)[0]
    cog = g(co, torch.float32)
    co32 = co.astype(np.float32).astype(np.float64)
    k = h(N.nufft(cog, g(x[None]), [64, 64, 64], 2, -1, 1e-6))[0]
    kref = O.nufft_type2(co32, x, -1, 1e-12)
    e2 = O.rel_l2(k, kref)
    y = rnd(rng, co.shape[1])
    img = h(N.nufft(cog, g(y[None]), [64, 64, 64], 1, +1, 1e-6))[0]
    e1 = O.rel_l2(img, O.nufft_type1(co32, y, (64, 64, 64), +1, 1e-12))
    print(f"cfg1 NUFFT type2 {e2:.2e} type1 {e1:.2e}", flush=True)
    assert e2 < 1e-5 and e1 < 1e-5


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("traj", ["kooshball", "uniform", "pileup"])
@pytest.mark.parametrize("tol,bound", [(1e-6, 1e-5), (1e-3, 1e-2)])
def test_nufft_type1_spread_variants(N, variant, traj, tol, bound):
    # Every 3-D complex64 spreader against the FP64 direct sum. "pileup" puts 1500 points
    # on one position plus 500 in its tile, so the tile splits into several chunks whose
    # points all update the same cells; the kooshball has every spoke's k = 0 sample at
    # the origin.
    from hastycompute_amd import _capi

    rng = np.random.default_rng(11)
    n = 24
    if traj == "kooshball":
        co = O.kooshball(300, 24)
    elif traj == "uniform":
        co = rng.uniform(-np.pi, np.pi, (3, 6000))
    else:
        p0 = rng.uniform(-np.pi, np.pi, (3, 1))
        co = np.concatenate([np.repeat(p0, 1500, axis=1), p0 + rng.uniform(-0.3, 0.3, (3, 500)),
                             rng.uniform(-np.pi, np.pi, (3, 2000))], axis=1)
    co32 = co.astype(np.float32).astype(np.float64)
    y = rnd(rng, co.shape[1])
    print(f"spread variant {variant} {traj} tol {tol}", flush=True)
    prev = _capi.get_option("nufft_spread_variant")
    _capi.set_option("nufft_spread_variant", variant)
    try:
        img = h(N.nufft(g(co, torch.float32), g(y[None]), [n, n, n], 1, +1, tol))[0]
    finally:
        _capi.set_option("nufft_spread_variant", prev)
    e = O.rel_l2(img, O.nufft_type1(co32, y, (n, n, n), +1, 1e-12))
    print(f"  rel err {e:.2e}", flush=True)
    assert e < bound


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("fp64,tol,bound", [(True, 1e-10, 1e-8), (False, 1e-8, 1e-5)])
def test_nufft_type1_spread_variants_fp64_and_wide(N, variant, fp64, tol, bound):
    # The spreaders with complex128 data (ns = 11: 121 window rows, two per lane in the
    # window-row spreaders, 363 or 242 owned cells per wave and point in the plane-dealt
    # ones) and complex64 with a kernel wider than 8 (tol 1e-8: ns = 9),
    # on a kooshball with every spoke's k = 0 sample at the origin, against the direct sum.
    from hastycompute_amd import _capi

    n = 20
    co = O.kooshball(250, 20)
    rng = np.random.default_rng(12)
    y = rnd(rng, co.shape[1])
    rt, ct = (torch.float64, torch.complex128) if fp64 else (torch.float32, torch.complex64)
    co_used = co if fp64 else co.astype(np.float32).astype(np.float64)
    prev = _capi.get_option("nufft_spread_variant")
    _capi.set_option("nufft_spread_variant", variant)
    try:
        plan = N.NufftPlan(g(co, rt), [n, n, n], 1, +1, 1, tol)
        ns = plan.info()[1]
        img = torch.empty((1, n, n, n), dtype=ct, device="cuda")
        plan.execute(g(y[None], ct), img)
        img = h(img)[0]
        plan.close()
    finally:
        _capi.set_option("nufft_spread_variant", prev)
    e = O.rel_l2(img, O.nudft_type1(co_used, y, (n, n, n), +1))
    print(f"spread variant {variant} {'fp64' if fp64 else 'fp32'} tol {tol} (ns {ns}): {e:.2e}", flush=True)
    assert ns >= 9 and e < bound


@pytest.mark.parametrize("n,traj,tol,bound,fp64", [(24, "kooshball", 1e-6, 1e-5, False),
                                                   (24, "pileup", 1e-6, 1e-5, False),
                                                   (32, "uniform", 1e-6, 1e-5, False),
                                                   (40, "kooshball", 1e-8, 1e-6, False),
                                                   (24, "pileup", 1e-3, 1e-2, False),
                                                   (32, "kooshball", 1e-12, 1e-5, False),
                                                   (24, "kooshball", 1e-10, 1e-8, True),
                                                   (24, "pileup", 1e-10, 1e-8, True),
                                                   (32, "uniform", 1e-8, 1e-7, True)])
def test_nufft_type1_gather_spreader(N, n, traj, tol, bound, fp64):
    # The output-driven spreader (variant 6): 16^3-cell cores, fine grids 48^3 (three
    # cores per dim: every neighbour of a core at the edge is a periodic image), 64^3 and
    # 80^3; the kooshball and the pileup (1500 points on one position) make cores with
    # more (point, core) pairs than one part holds, which are zeroed first and spread in
    # atomic parts; ns from 4 to 16; complex64 (128-point parts) and complex128
    # (32-point parts, the FP64 Toeplitz-kernel build's spreader). Against the direct sum,
    # with two transforms per plan (the core lists are built once), and against the tile
    # spreader (variant 3) to rounding.
    from hastycompute_amd import _capi

    rng = np.random.default_rng(n + int(-np.log10(tol)))
    if traj == "kooshball":
        co = O.kooshball(400, 2 * n)
    elif traj == "uniform":
        co = rng.uniform(-np.pi, np.pi, (3, 8000))
    else:
        p0 = rng.uniform(-np.pi, np.pi, (3, 1))
        co = np.concatenate([np.repeat(p0, 1500, axis=1), p0 + rng.uniform(-0.3, 0.3, (3, 3000)),
                             rng.uniform(-np.pi, np.pi, (3, 2000))], axis=1)
    rt, ct = (torch.float64, torch.complex128) if fp64 else (torch.float32, torch.complex64)
    co_used = co if fp64 else co.astype(np.float32).astype(np.float64)
    M = co.shape[1]
    y = rnd(rng, (2, M))
    prev = _capi.get_option("nufft_spread_variant")
    outs = {}
    try:
        for v in (6, 3):
            _capi.set_option("nufft_spread_variant", v)
            plan = N.NufftPlan(g(co, rt), [n, n, n], 1, +1, 2, tol)
            nf, ns, _ = plan.info()
            img = torch.empty((2, n, n, n), dtype=ct, device="cuda")
            plan.execute(g(y, ct), img)
            outs[v] = h(img)
            plan.close()
    finally:
        _capi.set_option("nufft_spread_variant", prev)
    assert all(f % 16 == 0 and f >= 48 for f in nf)
    for t in range(2):
        e = O.rel_l2(outs[6][t], O.nudft_type1(co_used, y[t], (n, n, n), +1))
        e3 = O.rel_l2(outs[6][t], outs[3][t])
        print(f"gather n {n} {traj} tol {tol} (ns {ns}, fp64 {fp64}) transform {t}: {e:.2e}, vs tiles {e3:.1e}",
              flush=True)
        assert e < bound and e3 < (1e-12 if fp64 else 1e-5)


@pytest.mark.parametrize("n", [24, 20])
def test_nufft_no_points(N, n):
    # An empty point set (M = 0, empty tensors with null data pointers): type 1 gives a
    # zero image, type 2 writes nothing; at n = 24 (fine grid 48^3) through the
    # output-driven spreader, whose plain-stored cores hold the zeros, at n = 20 through
    # the tile spreader and the zero pass. A plan reused after points are set again still
    # matches the direct sum.
    co = torch.empty((3, 0), dtype=torch.float32, device="cuda")
    c = torch.empty((1, 0), dtype=torch.complex64, device="cuda")
    p1 = N.NufftPlan(co, [n, n, n], 1, +1, 1, 1e-6)
    f = torch.full((1, n, n, n), 7.0 + 0j, dtype=torch.complex64, device="cuda")
    p1.execute(c, f)
    torch.cuda.synchronize()
    assert torch.count_nonzero(f).item() == 0
    p1.close()
    p2 = N.NufftPlan(co, [n, n, n], 2, -1, 1, 1e-6)
    p2.execute(c, torch.ones((1, n, n, n), dtype=torch.complex64, device="cuda"))
    torch.cuda.synchronize()
    p2.close()
    # the same plan shape with points
    rng = np.random.default_rng(n)
    pts = rng.uniform(-np.pi, np.pi, (3, 500))
    y = rnd(rng, 500)
    img = h(N.nufft(g(pts, torch.float32), g(y[None]), [n, n, n], 1, +1, 1e-6))[0]
    assert O.rel_l2(img, O.nudft_type1(pts.astype(np.float32).astype(np.float64), y, (n, n, n), +1)) < 1e-5


@pytest.mark.parametrize("n", [64, 96, 128, 160])
def test_nufft_register_fft(N, n):
    # n^3 modes with fine grids 128^3 .. 320^3: every axis pass runs complex64 lines of
    # 16 R (R = 8, 12, 16, 20) through the register 16 x R four-step kernel (option
    # fft_r16 = 1): contiguous and strided lines, zero-padded and centred inputs (type 2),
    # cropped and centred outputs (type 1), deconvolution scales. Type 2 and type 1
    # against the oracle's FP64 NUFFT (tol 1e-12) and against the LDS Stockham kernel
    # (fft_r16 = 0) to rounding.
    from hastycompute_amd import _capi

    rng = np.random.default_rng(n)
    co = O.kooshball(150, 2 * n)
    co32 = co.astype(np.float32).astype(np.float64)
    M = co.shape[1]
    f = rnd(rng, (n, n, n))
    y = rnd(rng, M)
    prev = _capi.get_option("fft_r16")
    outs = {}
    try:
        for r16 in (1, 0):
            _capi.set_option("fft_r16", r16)
            outs[r16] = (h(N.nufft(g(co, torch.float32), g(f[None]), [n, n, n], 2, -1, 1e-6))[0],
                         h(N.nufft(g(co, torch.float32), g(y[None]), [n, n, n], 1, +1, 1e-6))[0])
    finally:
        _capi.set_option("fft_r16", prev)
    e2 = O.rel_l2(outs[1][0], O.nufft_type2(co32, f, -1, 1e-12))
    e1 = O.rel_l2(outs[1][1], O.nufft_type1(co32, y, (n, n, n), +1, 1e-12))
    d2 = O.rel_l2(outs[1][0], outs[0][0])
    d1 = O.rel_l2(outs[1][1], outs[0][1])
    print(f"{n}^3 register FFT: type 2 {e2:.2e}, type 1 {e1:.2e}; vs Stockham {d2:.1e}, {d1:.1e}", flush=True)
    assert e2 < 1e-5 and e1 < 1e-5 and d2 < 1e-6 and d1 < 1e-6


@pytest.mark.parametrize("locality", [0, 1, 2, 3])
@pytest.mark.parametrize("traj", ["kooshball", "pileup"])
def test_nufft_point_orders(N, locality, traj):
    # Every point ordering (bins by super-tile or in plain order, workgroups XCD-aware or
    # not): type 2 and type 1 against the direct sums.
    from hastycompute_amd import _capi

    rng = np.random.default_rng(13)
    n = 28
    if traj == "kooshball":
        co = O.kooshball(400, 28)
    else:
        p0 = rng.uniform(-np.pi, np.pi, (3, 1))
        co = np.concatenate([np.repeat(p0, 1500, axis=1), rng.uniform(-np.pi, np.pi, (3, 3000))], axis=1)
    co32 = co.astype(np.float32).astype(np.float64)
    f = rnd(rng, (n, n, n))
    y = rnd(rng, co.shape[1])
    prev = _capi.get_option("nufft_locality")
    _capi.set_option("nufft_locality", locality)
    try:
        out2 = h(N.nufft(g(co, torch.float32), g(f[None]), [n, n, n], 2, -1, 1e-6))[0]
        out1 = h(N.nufft(g(co, torch.float32), g(y[None]), [n, n, n], 1, +1, 1e-6))[0]
    finally:
        _capi.set_option("nufft_locality", prev)
    e2 = O.rel_l2(out2, O.nudft_type2(co32, f, -1))
    e1 = O.rel_l2(out1, O.nudft_type1(co32, y, (n, n, n), +1))
    print(f"locality {locality} {traj}: type 2 {e2:.2e}, type 1 {e1:.2e}", flush=True)
    assert e2 < 1e-5 and e1 < 1e-5


@pytest.mark.parametrize("nm", [[1, 8, 6, 10], [2, 6, 6, 6], [1, 12, 9], [3, 20]])
def test_toeplitz_diag_and_apply_modes(N, nm):
    rng = np.random.default_rng(sum(nm))
    d = len(nm) - 1
    co = rng.uniform(-np.pi, np.pi, (d, 300))
    co32 = co.astype(np.float32).astype(np.float64)
    cog = g(co, torch.float32)
    diag, _ = N.toeplitz_build(cog, nm[1:], kernel=False)
    dref = O.toeplitz_diagonal(co32, nm, exact=True)
    assert O.rel_l2(h(diag)[None], dref[:1]) < 1e-5
    dg = diag.expand(nm[0], *diag.shape).contiguous()
    x = rnd(rng, nm)
    mul = rnd(rng, nm)
    ref = O.toeplitz_apply(dref, x)
    out = N.toeplitz_apply(dg, g(x), mode=0)
    assert O.rel_l2(h(out), ref) < 1e-5
    base = rnd(rng, nm)
    out = g(base)
    N.toeplitz_apply(dg, g(x), mode=1, out=out)
    assert O.rel_l2(h(out), base + ref) < 1e-5
    out = g(base)
    N.toeplitz_apply(dg, g(x), mode=2, out=out, mul=g(mul))
    assert O.rel_l2(h(out), base + ref * mul) < 1e-5
    xi = g(x)
    N.toeplitz_apply(dg, xi, mode=3)
    assert O.rel_l2(h(xi), ref) < 1e-5


@pytest.mark.parametrize("n,C,M", [(16, 4, 6000), (12, 3, 2500)])
def test_sense_normal_toeplitz_vs_explicit(N, n, C, M):
    rng = np.random.default_rng(n + C)
    co = O.kooshball(M // (2 * n), 2 * n)
    co32 = co.astype(np.float32).astype(np.float64)
    S = O.coil_maps(C, n, seed=1)
    x = rnd(rng, (1, n, n, n))
    w = rng.uniform(0.5, 1.5, co.shape[1])
    y = rnd(rng, (C, co.shape[1]))
    coils = list(range(C))[::-1]
    ref = O.sense_normal_apply(co32, x, S, coils)
    cog = g(co, torch.float32)
    _, kern = N.toeplitz_build(cog, [n, n, n])
    op = N.SenseNormalToeplitz(g(S), kern, coils=coils, coil_batch=3)
    out = op.apply(g(x))
    e = O.rel_l2(h(out), ref)
    print(f"sense normal n={n} C={C}: {e:.2e}", flush=True)
    assert e < 1e-5
    # gradient form A^H W (A x - y) = T_w x - A^H W y
    wt = torch.as_tensor(w.astype(np.float32)).cuda()
    _, kw = N.toeplitz_build(cog, [n, n, n], weights=wt)
    refg = O.sense_normal_apply(co32, x, S, coils, kdata=y[coils], weights=w.astype(np.float32))
    b = O.sense_adjoint_apply(co32, y[coils] * w.astype(np.float32)[None], S, coils)
    opw = N.SenseNormalToeplitz(g(S), kw, coils=coils)
    outg = opw.apply(g(x), sub=g(b))
    assert O.rel_l2(h(outg), refg) < 1e-5


def test_sense_normal_cfg1_batched_problems(N):
    # config-1 sizes (64^3, 8 coils, 2000 spokes x 64) and two problems with distinct kernels
    rng = np.random.default_rng(11)
    n, C = 64, 8
    S = O.coil_maps(C, n, seed=2)
    cos = [O.kooshball(2000, 64, offset=0), O.kooshball(1500, 64, offset=2000)]
    kernels = []
    for co in cos:
        _, k = N.toeplitz_build(g(co, torch.float32), [n, n, n])
        kernels.append(k)
    K = torch.stack(kernels)
    x = rnd(rng, (2, n, n, n))
    op = N.SenseNormalToeplitz(g(S), K, coil_batch=4)
    out = h(op.apply(g(x), kernel_index=[1, 0]))
    for p, kidx in enumerate([1, 0]):
        co32 = cos[kidx].astype(np.float32).astype(np.float64)
        diag = O.toeplitz_diagonal(co32, [1, n, n, n])
        ref = O.sense_normal_toeplitz(diag, x[p:p + 1], S, list(range(C)))
        e = O.rel_l2(out[p:p + 1], ref)
        print(f"cfg1 sense normal problem {p}: {e:.2e}", flush=True)
        assert e < 1e-5


@pytest.mark.parametrize("traj,l2", [("uniform", 0.0), ("kooshball", 0.0), ("kooshball", 0.3)])
def test_device_cg_vs_oracle_cg(N, traj, l2):
    # Uniform-density points give a well-conditioned system, where FP32 and FP64 CG
    # iterates agree to ~1e-7, so the image is held to 1e-5. On a kooshball the later CG
    # iterates are ill-conditioned in their inputs: relative perturbations of the
    # operator's outputs at the size of the device operator's own error (~1e-7) move the
    # FP64 CG's 10th iterate by 2-6 %. There the device CG is held to that envelope: the
    # largest distance (and phantom error) the oracle's FP64 CG reaches over three seeds
    # of Gaussian output noise at 3x the device operator's measured error. With the
    # reference's l2_reg_linop (python/python/hastypy/base/opalg.py:626-627) at
    # 0.3 lambda_max the system is well posed (at this size 0.1 lambda_max still leaves
    # the FP64 CG 5e-5 sensitive to 1e-7 output noise; 0.3: 5e-7) and the image is held
    # to 1e-5 again.
    rng = np.random.default_rng(5)
    n, C, P, iters = 16, 4, 2, 10
    S = O.coil_maps(C, n, seed=4)
    if traj == "uniform":
        cos = [rng.uniform(-np.pi, np.pi, (3, 3 * n ** 3)) for _ in range(P)]
    else:
        cos = [O.kooshball(600, 32, offset=0), O.kooshball(500, 32, offset=700)]
    kerns, diags = [], []
    for co in cos:
        d, k = N.toeplitz_build(g(co, torch.float32), [n, n, n])
        kerns.append(k)
        diags.append(O.toeplitz_diagonal(co.astype(np.float32).astype(np.float64), [1, n, n, n]))
    op = N.SenseNormalToeplitz(g(S), torch.stack(kerns))
    xt = np.concatenate([O.ellipsoid_phantom(n, seed=10 + p) for p in range(P)])
    Arefs = [lambda v, p=p: O.sense_normal_toeplitz(diags[p], v, S, list(range(C))) for p in range(P)]
    if traj == "uniform":
        b = rnd(rng, (P, n, n, n))
    else:
        b = np.concatenate([Arefs[p](xt[p:p + 1]) for p in range(P)])
    x0 = np.zeros_like(b)
    lam = 0.0
    if l2:
        v = rnd(np.random.default_rng(1), (1, n, n, n))
        for _ in range(40):
            w = Arefs[0](v)
            lmax = np.linalg.norm(w)
            v = w / lmax
        lam = l2 * lmax

    def A(v, out):
        op.apply(v, out=out, kernel_index=[0, 1])

    x, state = N.cg_solve(A, g(b), g(x0), iters, l2=lam)
    xs = h(x)
    Ab = h(op.apply(g(b), kernel_index=[0, 1]))
    for p in range(P):
        Aref = (lambda v, p=p: Arefs[p](v) + lam * v)
        xr = O.conjugate_gradient(Aref, b[p:p + 1], x0[p:p + 1], iters)
        e = O.rel_l2(xs[p:p + 1], xr)
        print(f"CG {traj} l2={l2} problem {p}: image vs FP64 CG {e:.2e}", flush=True)
        if traj == "uniform" or l2:
            assert e < 1e-5
        else:
            e_op = O.rel_l2(Ab[p:p + 1], Arefs[p](b[p:p + 1]))
            env_d, env_ph = 0.0, 0.0
            for seed in range(3):
                r = np.random.default_rng(seed)

                def Anoisy(v, r=r, Aref=Aref):
                    y = Aref(v)
                    return y + 3 * e_op * np.linalg.norm(y) / math.sqrt(y.size) * (
                        r.standard_normal(y.shape) + 1j * r.standard_normal(y.shape)) / math.sqrt(2)

                xn = O.conjugate_gradient(Anoisy, b[p:p + 1], x0[p:p + 1], iters)
                env_d = max(env_d, O.rel_l2(xn, xr))
                env_ph = max(env_ph, O.rel_l2(xn, xt[p:p + 1]))
            eg, er = O.rel_l2(xs[p:p + 1], xt[p:p + 1]), O.rel_l2(xr, xt[p:p + 1])
            print(f"   device operator error {e_op:.2e}; FP64 CG under 3x that output noise: up to {env_d:.2e} "
                  f"away, phantom error up to {env_ph:.4f}; device phantom error {eg:.4f}, FP64 CG {er:.4f}",
                  flush=True)
            assert e <= max(1.25 * env_d, 1e-5)
            assert eg <= 1.02 * max(env_ph, er)
    st = state.cpu().numpy()
    assert np.all(st[:, 6] == iters) and np.all(st[:, 4] == 0)


def test_device_cg_not_positive_definite_flag(N):
    # A = -I: pAp < 0 on the first step; the host raises the reference's message
    b = torch.ones((1, 4, 4, 4), dtype=torch.complex64, device="cuda")

    def A(v, out):
        out.copy_(-v)

    with pytest.raises(RuntimeError, match="positive definite"):
        N.cg_solve(A, b, torch.zeros_like(b), 3)


def _lowrank(rng, F, E, X, rank, noise):
    t = rnd(rng, (F, rank))
    sp = rnd(rng, (rank, E * X * X * X))
    return ((t @ sp).reshape(F, E, X, X, X) + noise * rnd(rng, (F, E, X, X, X))) / math.sqrt(rank)


@pytest.mark.parametrize("F,E,X,blk,it,soft,rel", [(20, 1, 32, 8, 3, True, 0.05), (12, 2, 24, 8, 2, True, 0.05),
                                                  (20, 1, 32, 8, 2, False, 0.0), (100, 1, 32, 16, 2, True, 0.05),
                                                  (100, 1, 48, 16, 2, True, 0.01)])
def test_llr_normal_blocks_vs_oracle(N, F, E, X, blk, it, soft, rel):
    # rel = 0.01: the LLR-regularised 4D-flow configuration's threshold, 0.01 sigma_max,
    # on its 16^3 blocks of 100 frames (4096 x 100 Casorati matrices), where most small
    # singular values survive the threshold and the kept eigenvectors are many
    rng = np.random.default_rng(F + X)
    x = _lowrank(rng, F, E, X, 5, 0.05).astype(np.complex64)
    A = O.extract_block(x.astype(np.complex128), (0, 0, 0), (blk,) * 3)
    smax = np.linalg.svd(A, compute_uv=False)[0]
    thresh = rel * smax if soft else 3
    ref = O.normal_3d_blocks_svt(x.astype(np.complex128), (blk,) * 3, (blk,) * 3, it, thresh, soft)
    xg = g(x)
    N.llr_normal_blocks(xg, (blk,) * 3, (blk,) * 3, it, thresh, soft)
    e = O.rel_l2(h(xg), ref)
    print(f"LLR F={F} E={E} X={X} blk={blk} iters={it} soft={soft}: {e:.2e}", flush=True)
    assert e < 1e-4


@pytest.mark.parametrize("F,soft", [(100, True), (20, False)])
def test_llr_repeatable(N, F, soft):
    """The LLR prox is deterministic: partial Grams are summed in a fixed order, the
    eigensolver and A W have no atomics, so two runs on the same input agree bit for bit."""
    rng = np.random.default_rng(11 + F)
    x = _lowrank(rng, F, 1, 64, 5, 0.05).astype(np.complex64)
    A = O.extract_block(x.astype(np.complex128), (0, 0, 0), (16,) * 3)
    thresh = 0.05 * np.linalg.svd(A, compute_uv=False)[0] if soft else 3
    x1, x2 = g(x), g(x)
    N.llr_normal_blocks(x1, (16,) * 3, (16,) * 3, 2, thresh, soft)
    N.llr_normal_blocks(x2, (16,) * 3, (16,) * 3, 2, thresh, soft)
    torch.cuda.synchronize()
    assert not torch.equal(x1, g(x))
    assert torch.equal(x1, x2)


def test_svt_matrix_vs_numpy(N):
    rng = np.random.default_rng(3)
    for R, F in [(500, 20), (4096, 100), (64, 100)]:
        A = (rnd(rng, (R, 6)) @ rnd(rng, (6, F)) + 0.1 * rnd(rng, (R, F))).astype(np.complex64)
        s = np.linalg.svd(A.astype(np.complex128), compute_uv=False)
        lam = 0.3 * s[2]
        Ag = g(A)
        N.svt_matrix(Ag, lam, True)
        e = O.rel_l2(h(Ag), O.svt_soft(A.astype(np.complex128), lam))
        Ah = g(A)
        N.svt_matrix(Ah, 4, False)
        e2 = O.rel_l2(h(Ah), O.svt_hard(A.astype(np.complex128), 4))
        print(f"svt R={R} F={F}: soft {e:.2e} hard {e2:.2e}", flush=True)
        assert e < 1e-4 and e2 < 1e-4


@pytest.mark.parametrize("F", [1, 2, 3])
def test_svt_matrix_smallest(N, F):
    """One to three columns: n = 2 or 4 in the eigensolver (a single pair, or the
    padding column), soft and hard thresholds."""
    rng = np.random.default_rng(40 + F)
    A = rnd(rng, (50, F)).astype(np.complex64)
    s = np.linalg.svd(A.astype(np.complex128), compute_uv=False)
    for lam in (0.5 * s[-1], 0.5 * (s[0] + s[-1]), 2.0 * s[0]):
        Ag = g(A)
        N.svt_matrix(Ag, lam, True)
        e = O.rel_l2(h(Ag), O.svt_soft(A.astype(np.complex128), lam)) if lam < s[0] else float(np.abs(h(Ag)).max())
        print(f"svt F={F} lam={lam:.3g}: {e:.2e}", flush=True)
        assert e < 1e-4
    for rank in range(F + 1):
        Ah = g(A)
        N.svt_matrix(Ah, rank, False)
        ref = O.svt_hard(A.astype(np.complex128), rank)
        e2 = O.rel_l2(h(Ah), ref) if rank > 0 else float(np.abs(h(Ah)).max())
        assert e2 < 1e-4, (rank, e2)


@pytest.mark.parametrize("R,F", [(300, 33), (400, 128), (250, 100), (40, 7)])
def test_svt_matrix_many_kept(N, R, F):
    """Thresholds that keep all, about half, or one singular value: the eigenvectors are
    rebuilt in chunks of F/2, so keeping more than F/2 takes several; odd F pads G."""
    rng = np.random.default_rng(R + F)
    # prescribed spectrum, halved at the two hard-threshold ranks so the kept subspace is
    # well separated (a Gram matrix in FP32 resolves gaps down to ~ F eps |G|)
    k = np.arange(F)
    sv = (1.0 + 0.5 * (1.0 - k / F)) * np.where(k >= F - 1, 0.5, 1.0) * np.where(k >= F // 2 + 1, 0.5, 1.0)
    sv[0] *= 3.0
    Q1 = np.linalg.qr(rnd(rng, (R, F)))[0]
    Q2 = np.linalg.qr(rnd(rng, (F, F)))[0]
    A = ((Q1 * sv[None, :]) @ Q2.conj().T).astype(np.complex64)
    s = np.linalg.svd(A.astype(np.complex128), compute_uv=False)
    for lam in (0.5 * s[-1], s[F // 2], 0.5 * (s[0] + s[1])):
        Ag = g(A)
        N.svt_matrix(Ag, lam, True)
        e = O.rel_l2(h(Ag), O.svt_soft(A.astype(np.complex128), lam))
        print(f"svt R={R} F={F} kept={int((s > lam).sum())}: soft {e:.2e}", flush=True)
        assert e < 1e-4
    for rank in (F - 1, F // 2 + 1):
        Ah = g(A)
        N.svt_matrix(Ah, rank, False)
        e2 = O.rel_l2(h(Ah), O.svt_hard(A.astype(np.complex128), rank))
        print(f"svt R={R} F={F} hard rank={rank}: {e2:.2e}", flush=True)
        assert e2 < 1e-4
