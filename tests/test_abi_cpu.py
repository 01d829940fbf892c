"""CPU checks of the boundary: the C-ABI library loads and exports every function that
include/hasty_hip.h declares, its host-only size queries answer, and the TorchScript
operator library registers the reference's module/class names and reports the
reference's argument errors. No kernel is launched here."""
import ctypes
import os

import pytest
import torch

from hastycompute_amd import _capi


def test_header_symbols_exported():
    names = _capi.header_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature and nothing extra is bound
    assert sorted(_capi._SIGS) == names


def test_host_queries():
    L = _capi.lib()
    assert L.hasty_version() >= 1
    n3 = _capi.i64arr([128, 128, 128])
    ws = L.hasty_sense_normal_workspace_size(n3, 16, 4)
    assert ws >= 4 * 16 * 8 * (2 * 128) ** 3 // 8  # at least the per-coil half-grid staging
    assert L.hasty_cg_workspace_size(4, 1000) > 0
    assert L.hasty_svt_workspace_size(4096, 100) > 0


def test_llr_shifts_host():
    """hasty_llr_shifts is host code: it must reproduce the oracle's restatement."""
    from oracle import oracle as O

    import numpy as np

    out = np.zeros((3, 5), dtype=np.int32)
    _capi.check(_capi.lib().hasty_llr_shifts(5, _capi.i64arr([8, 8, 8]),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    assert (out == O.llr_shifts(5, [8, 8, 8])).all()


def test_invalid_arguments_rejected_on_host():
    L = _capi.lib()
    plan = ctypes.c_void_p()
    rc = L.hasty_nufft_makeplan(3, 3, _capi.i64arr([8, 8, 8]), -1, 1, 1e-6, 0, ctypes.byref(plan))
    assert rc != 0 and b"type" in L.hasty_last_error().lower()
    rc = L.hasty_nufft_makeplan(2, 4, _capi.i64arr([8, 8, 8, 8]), -1, 1, 1e-6, 0, ctypes.byref(plan))
    assert rc != 0


@pytest.fixture(scope="module")
def ffi():
    from hastycompute_amd.ffi import hasty_interface, hasty_nufft, hasty_op, hasty_sense, hasty_svt  # noqa: F401

    return hasty_nufft, hasty_sense, hasty_svt, hasty_interface


def test_torchscript_modules_registered(ffi):
    for mod, classes in {"HastyNufft": ["NufftOptions", "Nufft", "NufftNormal"],
                         "HastySense": ["Sense", "SenseAdjoint", "SenseNormal"],
                         "HastyBatchedSense": ["BatchedSense", "BatchedSenseAdjoint", "BatchedSenseNormal"],
                         "HastySVT": ["Random3DBlocksSVT", "Normal3DBlocksSVT"],
                         "HastyInterface": ["FunctionLambda"],
                         "HastyOp": ["SenseNOp", "PowerIteration", "ConjugateGradient",
                                     "BatchedConjugateGradient"]}.items():
        ns = getattr(torch.classes, mod)
        for c in classes:
            assert getattr(ns, c) is not None
        assert isinstance(getattr(torch.ops, mod).doc(), str)


def test_reference_errors(ffi):
    hasty_nufft, hasty_sense, _, _ = ffi
    with pytest.raises(RuntimeError, match="Not available type"):
        hasty_nufft.NufftOptions(4)
    co = torch.zeros(3, 10)
    with pytest.raises(RuntimeError, match="cuda device"):
        hasty_nufft.Nufft(co, [1, 8, 8, 8], hasty_nufft.NufftOptions.type2())
    with pytest.raises(RuntimeError, match="cuda device"):
        hasty_sense.SenseNormal(co, [1, 8, 8, 8])


def test_function_lambda_runs(ffi):
    _, _, _, hasty_interface = ffi
    fl = hasty_interface.FunctionLambda(
        "def apply(a: List[Tensor], b: List[Tensor]):\n    for i in range(len(a)):\n        a[i].mul_(b[0])\n",
        "apply", [torch.full((4,), 3.0)])
    x = torch.arange(4.0)
    fl.apply([x])
    assert torch.equal(x, torch.arange(4.0) * 3)


def test_product_has_no_oracle_import():
    """Nothing in the package imports the oracle (it is test infrastructure only)."""
    pkg = os.path.dirname(os.path.abspath(_capi.__file__))
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                text = open(os.path.join(root, f), errors="ignore").read()
                assert "import oracle" not in text and "from oracle" not in text, f


def test_serving_request_stream_host_logic():
    """The Poisson request stream is seeded and the footprints follow the stated sizes."""
    from hastycompute_amd import serving

    a = serving.poisson_requests(20, 2.0, seed=3)
    b = serving.poisson_requests(20, 2.0, seed=3)
    assert [(r.n, r.ncoils, r.nspokes, r.arrival) for r in a] == [(r.n, r.ncoils, r.nspokes, r.arrival) for r in b]
    assert all(r.n in (96, 128, 160) and 8 <= r.ncoils <= 32 and 1000 <= r.nspokes <= 10000 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    assert serving.resident_bytes(128, 16) == 8 * 16 * 128 ** 3 + 4 * 256 ** 3 + 40 * 128 ** 3


def test_llr_grid_corners_match_the_oracle_grid():
    # the host enumeration behind hasty_llr_blocks_dev follows the reference's strided
    # grid and skip rule (svt.cpp:378-404), as the oracle restates it
    import numpy as np

    from hastycompute_amd import native
    from oracle import oracle as O

    for dims, strides, shape in (([3, 1, 48, 40, 36], [16, 16, 16], [16, 16, 16]),
                                 ([2, 2, 20, 24, 28], [4, 6, 5], [8, 8, 8])):
        shifts = np.array(O.llr_shifts(4, shape), dtype=np.int32).reshape(3, 4)
        ref = O.normal_blocks(dims[2:], strides, shape, 4, shifts)
        for it in range(4):
            got = native.llr_grid_corners(dims, strides, shape, shifts[:, it])
            assert [tuple(c) for c in got] == [tuple(c) for c in ref[it]], (dims, it)


def test_options_round_trip_on_host():
    """Every option the header documents for hasty_set_option can be read; the ranged ones
    accept their bounds and reject the values just outside (host code, no kernel)."""
    import re

    text = open(_capi.HEADER).read()
    block = text[text.index("/* Library options"):text.index("int hasty_set_option")]
    names = re.findall(r'^ \*   "(\w+)"', block, flags=re.M)
    assert len(names) >= 12
    L = _capi.lib()
    saved = {n: _capi.get_option(n) for n in names}
    ranges = {"sense_xy_variant": ([0, 3], [-1, 4]), "sense_zconv_batch": ([1, 16], [0, 17]),
              "nufft_locality": ([0, 4], [-1, 5]), "sense_zconv_variant": ([0, 3], [-1, 4]),
              "llr_max_sweeps": ([1, 15], [0, 16]), "llr_eig_threads": ([256, 512, 768], [128, 300, 1024]),
              "nufft_spread_tile": ([0, 12, 28], [11, 29]), "nufft_spread_chunk": ([0, 32, 8192], [31, 8193]),
              "nufft_spread_variant": ([0, 6], [-1, 7]), "fft_r16": ([0, 1], [-1, 2]),
              "fft_strided_lines": ([0, 64], [-1, 65])}
    try:
        for n, (good, bad) in ranges.items():
            assert n in names, n
            for v in good:
                _capi.set_option(n, v)
                assert _capi.get_option(n) == v, (n, v)
            before = _capi.get_option(n)
            for v in bad:
                assert L.hasty_set_option(n.encode(), v) == 1, (n, v)
                assert _capi.get_option(n) == before, (n, v)
        assert L.hasty_set_option(b"no_such_option", 1) == 1
    finally:
        for n, v in saved.items():
            _capi.set_option(n, v)
