"""The launch-shape autotuner (hasty_autotune_sense_normal, hasty_nufft_autotune,
hasty_autotune_llr): choices inside their ranges, cached per shape, and results with the
tuned shapes equal to the default shapes' within floating-point reordering."""
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


@pytest.fixture()
def restore_options():
    from hastycompute_amd import _capi as C

    names = ("sense_xy_variant", "sense_zconv_batch", "sense_zconv_variant", "nufft_spread_variant",
             "nufft_spread_tile", "nufft_spread_chunk", "llr_eig_threads", "fft_strided_lines")
    saved = {k: C.get_option(k) for k in names}
    yield
    for k, v in saved.items():
        C.set_option(k, v)


@pytest.mark.parametrize("shape,C", [((32, 64, 64), 4), ((128, 32, 32), 3)])
def test_autotune_sense_normal(N, restore_options, shape, C):
    from hastycompute_amd import _capi as Cap
    from hastycompute_amd import autotune

    rng = np.random.default_rng(1)
    P = 6
    co = torch.as_tensor(rng.uniform(-np.pi, np.pi, (3, 3000)).astype(np.float32)).cuda()
    _, k = N.toeplitz_build(co, list(shape), diag=False)
    S = torch.randn((C,) + shape, dtype=torch.complex64, device="cuda")
    x = torch.randn((P,) + shape, dtype=torch.complex64, device="cuda")
    ref = N.SenseNormalToeplitz(S, k, coil_batch=C, problem_batch=1).apply(x)
    t = autotune.sense_normal(shape, C, P)
    print(f"autotune {shape} x {C} coils: {t}", flush=True)
    assert t["sense_xy_variant"] in (0, 1, 2, 3) and 1 <= t["sense_zconv_batch"] <= 16
    assert 1 <= t["problem_batch"] <= min(P, 8) and 1 <= t["coil_batch"] <= C
    assert t["ms_per_problem"] > 0
    assert Cap.get_option("sense_xy_variant") == t["sense_xy_variant"]
    assert autotune.sense_normal(shape, C, P) == t  # cached
    op = N.SenseNormalToeplitz(S, k, coil_batch=t["coil_batch"], problem_batch=t["problem_batch"])
    out = op.apply(x)
    torch.cuda.synchronize()
    e = O.rel_l2(out.cpu().numpy(), ref.cpu().numpy())
    assert e < 1e-5, e


def test_autotune_nufft_type1(N, restore_options):
    from hastycompute_amd import autotune

    n = 32
    co = torch.as_tensor(O.kooshball(300, 64).astype(np.float32)).cuda()
    M = co.shape[1]
    c = torch.randn((1, M), dtype=torch.complex64, device="cuda")
    plan = N.NufftPlan(co, [n, n, n], 1, 1, 1, 1e-6)
    f0 = torch.empty((1, n, n, n), dtype=torch.complex64, device="cuda")
    plan.execute(c, f0)
    f = torch.empty_like(f0)
    t = autotune.nufft(plan, c, f)
    print(f"autotune nufft: {t}", flush=True)
    assert t["nufft_spread_variant"] in range(7) and t["execute_ms"] > 0
    assert t["nufft_spread_tile"] in (16, 20) and t["nufft_spread_chunk"] in (64, 128, 256)
    plan.execute(c, f)
    torch.cuda.synchronize()
    assert O.rel_l2(f.cpu().numpy(), f0.cpu().numpy()) < 1e-5
    plan.close()


def test_autotune_llr_row_split(N, restore_options):
    from hastycompute_amd import autotune

    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    dims = (20, 1, 48, 48, 48)
    x = torch.randn(dims, dtype=torch.complex64, device="cuda", generator=g)
    before = N.llr_normal_blocks(x.clone(), (16, 16, 16), (16, 16, 16), 1, 2.0, True)
    t = autotune.llr(dims, (16, 16, 16), (16, 16, 16))
    print(f"autotune llr: {t}", flush=True)
    assert 1 <= t["llr_row_split"] <= 8 and t["llr_eig_threads"] in (256, 512, 768)
    after = N.llr_normal_blocks(x.clone(), (16, 16, 16), (16, 16, 16), 1, 2.0, True)
    torch.cuda.synchronize()
    assert O.rel_l2(after.cpu().numpy(), before.cpu().numpy()) < 1e-4


@pytest.mark.parametrize("tile,chunk", [(16, 64), (24, 1024), (13, 256)])
def test_nufft_spread_tile_options_vs_direct_sum(N, restore_options, tile, chunk):
    # plans made under the nufft_spread_tile / nufft_spread_chunk options (tiles from
    # 13^3 to 24^3 cells with halo, 64..1024 points per workgroup) spread exactly what the
    # direct sum gives, on a radial trajectory (a crowded centre) at ns = 7
    from hastycompute_amd import _capi as C

    n = 24
    co = O.kooshball(200, 48)
    M = co.shape[1]
    rng = np.random.default_rng(tile + chunk)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    C.set_option("nufft_spread_variant", 3)  # a tile spreader (the output-driven one has fixed cores)
    C.set_option("nufft_spread_tile", tile)
    C.set_option("nufft_spread_chunk", chunk)
    plan = N.NufftPlan(torch.as_tensor(co.astype(np.float32)).cuda(), [n, n, n], 1, 1, 1, 1e-6)
    f = torch.empty((1, n, n, n), dtype=torch.complex64, device="cuda")
    plan.execute(torch.as_tensor(c[None]).cuda(), f)
    torch.cuda.synchronize()
    plan.close()
    ref = O.nudft_type1(co.astype(np.float32).astype(np.float64), c.astype(np.complex128), [n, n, n], 1)
    e = O.rel_l2(f.cpu().numpy()[0].astype(np.complex128), ref)
    print(f"tile {tile} chunk {chunk}: {e:.2e}", flush=True)
    assert e < 1e-5


def test_tile_option_does_not_bind_the_output_driven_spreader(N, restore_options):
    # a 28-cell tile (28^3 complex64 = 172 KB, more than a CU's LDS) is refused by the tile
    # spreaders on the host, before any launch, but does not concern the output-driven
    # spreader, which has no tile
    from hastycompute_amd import _capi as C

    n = 24
    co = O.kooshball(100, 48)
    M = co.shape[1]
    rng = np.random.default_rng(28)
    y = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    C.set_option("nufft_spread_tile", 28)
    img = N.nufft(torch.as_tensor(co.astype(np.float32)).cuda(), torch.as_tensor(y[None]).to(torch.complex64).cuda(),
                  [n, n, n], 1, +1, 1e-6)
    ref = O.nudft_type1(co.astype(np.float32).astype(np.float64), y, (n, n, n), +1)
    assert O.rel_l2(img.cpu().numpy()[0].astype(np.complex128), ref) < 1e-5
    C.set_option("nufft_spread_variant", 3)
    with pytest.raises(Exception, match="exceeds LDS"):
        N.nufft(torch.as_tensor(co.astype(np.float32)).cuda(), torch.as_tensor(y[None]).to(torch.complex64).cuda(),
                [n, n, n], 1, +1, 1e-6)
    torch.cuda.synchronize()


@pytest.mark.parametrize("fp64", [False, True])
def test_fft_strided_lines_same_bits(N, restore_options, fp64):
    # the lines per workgroup of the strided axis FFT passes (1, 8, auto, 64,
    # which the LDS limit halves) only regroup independent lines: type-2 NUFFTs (FFT
    # passes, then a gather) give the same bits for every value; type 1 and the FP64
    # Toeplitz kernel build, whose tile flushes add with global atomics in no fixed
    # order, agree to rounding; type 1 matches the direct sum
    from hastycompute_amd import _capi as C

    n = 24
    rdt, cdt = (torch.float64, torch.complex128) if fp64 else (torch.float32, torch.complex64)
    co = O.kooshball(120, 48)
    M = co.shape[1]
    rng = np.random.default_rng(7)
    c = torch.as_tensor((rng.standard_normal((1, M)) + 1j * rng.standard_normal((1, M)))).to(cdt).cuda()
    img = torch.as_tensor((rng.standard_normal((1, n, n, n)) + 1j * rng.standard_normal((1, n, n, n)))).to(cdt).cuda()
    cot = torch.as_tensor(co).to(rdt).cuda()
    tol = 1e-10 if fp64 else 1e-6
    outs = []
    for lines in (1, 8, 0, 64):
        C.set_option("fft_strided_lines", lines)
        p1 = N.NufftPlan(cot, [n, n, n], 1, 1, 1, tol)
        f = torch.empty((1, n, n, n), dtype=cdt, device="cuda")
        p1.execute(c, f)
        p2 = N.NufftPlan(cot, [n, n, n], 2, -1, 1, tol)
        d = torch.empty((1, M), dtype=cdt, device="cuda")
        p2.execute(d, img)
        torch.cuda.synchronize()
        p1.close()
        p2.close()
        _, k = N.toeplitz_build(cot.float(), [n, n, n], diag=False)
        outs.append((f.clone(), d.clone(), k.clone()))
    rt = 1e-13 if fp64 else 1e-6
    for lines, o in zip((8, 0, 64), outs[1:]):
        assert torch.equal(outs[0][1], o[1])
        e1 = O.rel_l2(o[0].cpu().numpy(), outs[0][0].cpu().numpy())
        ek = O.rel_l2(o[2].cpu().numpy(), outs[0][2].cpu().numpy())
        print(f"fp64={fp64} lines {lines}: type 1 {e1:.1e}, kernel {ek:.1e}", flush=True)
        assert e1 < rt and ek < 1e-6
    ref = O.nudft_type1(co.astype(np.float64) if fp64 else co.astype(np.float32).astype(np.float64),
                        c.cpu().numpy()[0].astype(np.complex128), [n, n, n], 1)
    e = O.rel_l2(outs[0][0].cpu().numpy()[0].astype(np.complex128), ref)
    print(f"fp64={fp64}: type 1 vs direct sum {e:.2e}", flush=True)
    assert e < (1e-9 if fp64 else 1e-5)


@pytest.mark.parametrize("threads", [256, 512, 768])
def test_llr_eig_threads_vs_oracle(N, restore_options, threads):
    # the LLR prox with each eigensolver workgroup size (3, 5 or 9 blocks per thread at
    # 100 frames) against the oracle's LAPACK SVT of the same blocks
    from hastycompute_amd import _capi as C

    rng = np.random.default_rng(threads)
    F, n = 100, 32
    U = rng.standard_normal((5, n ** 3)) + 1j * rng.standard_normal((5, n ** 3))
    V = rng.standard_normal((F, 5)) + 1j * rng.standard_normal((F, 5))
    x = ((V @ U).reshape(F, 1, n, n, n) + 0.05 * (rng.standard_normal((F, 1, n, n, n)) +
                                                 1j * rng.standard_normal((F, 1, n, n, n)))).astype(np.complex64)
    C.set_option("llr_eig_threads", threads)
    blk = (16, 16, 16)
    sig = np.linalg.svd(x[:, 0, :16, :16, :16].reshape(F, -1).T.astype(np.complex128), compute_uv=False)[0]
    lam = 0.01 * sig
    out = N.llr_normal_blocks(torch.as_tensor(x).cuda(), blk, blk, 1, float(lam), True).cpu().numpy()
    ref = O.normal_3d_blocks_svt(x.astype(np.complex128), list(blk), list(blk), 1, lam, True)
    e = O.rel_l2(out.astype(np.complex128), ref)
    print(f"llr eig threads {threads}: {e:.2e}", flush=True)
    assert e < 1e-4
