"""GPU parity of the fused three-kernel Toeplitz SENSE normal operator (layouts 1 and 2)
against the FP64 oracle and against the generic five-pass path (layout 0)."""
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


def rnd(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def maps(rng, C, shape):
    m = rnd(rng, (C,) + shape) * 0.3 + 1.0
    return m / np.sqrt(np.sum(np.abs(m) ** 2, axis=0, keepdims=True))


@pytest.mark.parametrize("N0,Nn,C,cb,P", [(32, 32, 3, 2, 2), (40, 64, 4, 4, 3), (24, 96, 2, 1, 1), (20, 32, 5, 3, 4),
                                         (8, 128, 3, 2, 3),
                                         # N = 160: coset kernels one coset per step, the
                                         # inverse in two row halves (N0 = 20: a short last
                                         # group of 8 planes); 2 N0 = 320 and 300 z lines
                                         (8, 160, 3, 2, 3), (20, 160, 2, 1, 2), (160, 32, 2, 2, 2),
                                         (150, 32, 2, 1, 1)])
def test_fused_vs_oracle(N, N0, Nn, C, cb, P):
    rng = np.random.default_rng(N0 + Nn + C)
    shape = (N0, Nn, Nn)
    S = maps(rng, C, shape)
    cos = [rng.uniform(-np.pi, np.pi, (3, 3000)) for _ in range(2)]
    kern = []
    diags = []
    for co in cos:
        _, k = N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)
        kern.append(k)
        diags.append(O.toeplitz_diagonal(co.astype(np.float32).astype(np.float64), [1, *shape]))
    K = torch.stack(kern)
    x = rnd(rng, (P,) + shape)
    sub = rnd(rng, (P,) + shape)
    kidx = [p % 2 for p in range(P)][::-1]
    coils = list(range(C))[::-1]
    op = N.SenseNormalToeplitz(g(S), K, coils=coils, coil_batch=cb, problem_batch=2)
    assert op.layout != 0
    out = h(op.apply(g(x), kernel_index=kidx))
    outs = h(op.apply(g(x), kernel_index=kidx, sub=g(sub)))
    for p in range(P):
        ref = O.sense_normal_toeplitz(diags[kidx[p]], x[p:p + 1], S, coils)
        e = O.rel_l2(out[p:p + 1], ref)
        es = O.rel_l2(outs[p:p + 1], ref - sub[p:p + 1])
        print(f"fused N0={N0} N={Nn} C={C} cb={cb} p={p}: {e:.2e} (sub {es:.2e})", flush=True)
        assert e < 1e-5 and es < 1e-5


def test_fused_matches_generic_and_oracle_at_128(N):
    # the benchmark size: 128^3, two coils, kooshball kernel; the oracle gets the GPU-built
    # complex diagonal, so this pins the fused FFT pipeline itself at L = 256
    rng = np.random.default_rng(128)
    n, C = 128, 2
    co = O.kooshball(1000, 128)
    cog = g(co, torch.float32)
    diag, kern = N.toeplitz_build(cog, [n, n, n], tol=1e-8)
    S = g(maps(rng, C, (n, n, n)))
    x = g(rnd(rng, (1, n, n, n)))
    fused = N.SenseNormalToeplitz(S, kern, coil_batch=2)
    generic = N.SenseNormalToeplitz(S, kern, coil_batch=2, fused=False)
    assert fused.layout != 0 and generic.layout == 0
    of = h(fused.apply(x))
    og = h(generic.apply(x))
    e_fg = O.rel_l2(of, og)
    ref = O.sense_normal_toeplitz(h(diag)[None], h(x), h(S), [0, 1], workers=O.default_threads())
    e_f = O.rel_l2(of, ref)
    print(f"128^3: fused vs generic {e_fg:.2e}, fused vs oracle FFT pipeline {e_f:.2e}", flush=True)
    assert e_fg < 1e-5 and e_f < 1e-5


def test_fused_repeatable(N):
    # same inputs, same launch sequence -> identical bits (no atomics in the fused path)
    rng = np.random.default_rng(3)
    shape = (32, 64, 64)
    S = g(maps(rng, 3, shape))
    co = rng.uniform(-np.pi, np.pi, (3, 2000))
    _, k = N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)
    op = N.SenseNormalToeplitz(S, k, coil_batch=3)
    x = g(rnd(rng, (2,) + shape))
    a = op.apply(x).clone()
    b = op.apply(x).clone()
    assert torch.equal(a, b)


def test_fused_160_every_variant_runs_the_coset_kernels(N):
    # at N1 = N2 = 160 only the coset kernels fit in LDS: every sense_xy_variant value
    # prepares layout 1 and gives the same bits, which match the oracle
    from hastycompute_amd import _capi as C

    rng = np.random.default_rng(16)
    shape = (12, 160, 160)
    Sn = maps(rng, 3, shape)
    co = rng.uniform(-np.pi, np.pi, (3, 6000))
    _, k = N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)
    x = rnd(rng, (2,) + shape)
    prev = C.get_option("sense_xy_variant")
    outs = []
    try:
        for v in (0, 1, 2, 3):
            C.set_option("sense_xy_variant", v)
            op = N.SenseNormalToeplitz(g(Sn), k, coil_batch=2, problem_batch=2)
            assert op.layout == 1
            outs.append(op.apply(g(x)).clone())
    finally:
        C.set_option("sense_xy_variant", prev)
    assert all(torch.equal(outs[0], o) for o in outs[1:])
    diag = O.toeplitz_diagonal(co.astype(np.float32).astype(np.float64), [1, *shape])
    got = h(outs[0])
    for p in range(2):
        e = O.rel_l2(got[p:p + 1], O.sense_normal_toeplitz(diag, x[p:p + 1], Sn, [0, 1, 2]))
        print(f"160: problem {p}: {e:.2e}", flush=True)
        assert e < 1e-5


def test_prepared_layout_survives_option_change(N):
    # kernels prepared for the register plane kernels (layout 2 at N1 = N2 = 128) keep
    # running them after the sense_xy_variant option changes, and coset-prepared kernels
    # (layout 1) keep the coset order; both agree with the oracle-checked default
    from hastycompute_amd import _capi as C

    rng = np.random.default_rng(8)
    shape = (8, 128, 128)
    S = g(maps(rng, 2, shape))
    co = rng.uniform(-np.pi, np.pi, (3, 3000))
    _, k = N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)
    x = g(rnd(rng, (1,) + shape))
    prev = C.get_option("sense_xy_variant")
    try:
        C.set_option("sense_xy_variant", 2)
        op2 = N.SenseNormalToeplitz(S, k, coil_batch=2)
        C.set_option("sense_xy_variant", 1)
        op1 = N.SenseNormalToeplitz(S, k, coil_batch=2)
        assert (op2.layout, op1.layout) == (2, 1)
        a = h(op2.apply(x))  # option now 1: still the register kernels
        b = h(op1.apply(x))
        C.set_option("sense_xy_variant", 2)
        c = h(op1.apply(x))  # option now 2: still the coset kernels
    finally:
        C.set_option("sense_xy_variant", prev)
    assert O.rel_l2(a, b) < 1e-5 and O.rel_l2(c, b) < 1e-5


@pytest.mark.parametrize("N0", [128, 160])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_zconv_variants_vs_oracle(N, variant, N0):
    # every z-convolution kernel at N0 = 128 (2N0 = 256): 16 or 8 columns per workgroup,
    # one or two LDS transpose buffers; at N0 = 160 (2N0 = 320) the 20 x 16 register form
    # with one (0) or two (2) buffers and the LDS Stockham form (1, 3); three problems with
    # their own kernels
    from hastycompute_amd import _capi as C

    rng = np.random.default_rng(40 + variant)
    shape = (N0, 32, 32)
    S = maps(rng, 3, shape)
    cos = [rng.uniform(-np.pi, np.pi, (3, 3000)) for _ in range(2)]
    kern, diags = [], []
    for co in cos:
        _, k = N.toeplitz_build(g(co, torch.float32), list(shape), diag=False)
        kern.append(k)
        diags.append(O.toeplitz_diagonal(co.astype(np.float32).astype(np.float64), [1, *shape]))
    x = rnd(rng, (3,) + shape)
    kidx = [1, 0, 1]
    prev = C.get_option("sense_zconv_variant")
    try:
        C.set_option("sense_zconv_variant", variant)
        op = N.SenseNormalToeplitz(g(S), torch.stack(kern), coil_batch=3, problem_batch=3)
        out = h(op.apply(g(x), kernel_index=kidx))
    finally:
        C.set_option("sense_zconv_variant", prev)
    for p in range(3):
        e = O.rel_l2(out[p:p + 1], O.sense_normal_toeplitz(diags[kidx[p]], x[p:p + 1], S, [0, 1, 2]))
        print(f"z-conv variant {variant} problem {p}: {e:.2e}", flush=True)
        assert e < 1e-5
