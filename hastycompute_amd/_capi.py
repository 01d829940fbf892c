"""ctypes binding of libhasty_hip.so (the C ABI declared in include/hasty_hip.h).

Device pointers are passed as integers (``tensor.data_ptr()``) and streams as raw
``hipStream_t`` values (``torch.cuda.current_stream().cuda_stream``). Every call checks
the return code and raises :class:`HastyError` with ``hasty_last_error()``.
There is no fallback: if the library is missing, importing the product fails.
"""
from __future__ import annotations

import ctypes
import os
import re
import threading

import torch  # noqa: F401  (load torch's HIP runtime first so one runtime serves both)

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libhasty_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "hasty_hip.h")

_lock = threading.Lock()
_lib = None

c_int, c_i64, c_dbl, c_sz, c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_dbl = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "hasty_last_error": ([], ctypes.c_char_p),
    "hasty_version": ([], c_int),
    "hasty_nufft_makeplan": ([c_int, c_int, P_i64, c_int, c_int, c_dbl, c_int, ctypes.POINTER(c_vp)], c_int),
    "hasty_nufft_setpts": ([c_vp, c_i64, c_vp, c_vp, c_vp, c_vp], c_int),
    "hasty_nufft_execute": ([c_vp, c_vp, c_vp, c_vp], c_int),
    "hasty_nufft_set_scale": ([c_vp, c_dbl], c_int),
    "hasty_nufft_destroy": ([c_vp], c_int),
    "hasty_nufft_info": ([c_vp, P_i64, ctypes.POINTER(c_int), P_dbl], c_int),
    "hasty_toeplitz_build": ([c_vp, c_int, c_i64, c_int, P_i64, c_vp, c_vp, c_vp, c_dbl, c_vp], c_int),
    "hasty_toeplitz_apply": ([c_vp, c_vp, c_vp, c_vp, c_int, c_int, P_i64, c_int, c_vp, c_vp, c_vp], c_int),
    "hasty_sense_normal_workspace_size": ([P_i64, c_int, c_int], c_sz),
    "hasty_sense_normal_prepare": ([c_vp, c_vp, c_i64, P_i64, ctypes.POINTER(c_int), c_vp], c_int),
    "hasty_sense_normal_toeplitz": ([c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_int, c_int, P_i32, c_vp, c_int, P_i64,
                                     c_vp, c_sz, c_int, c_vp], c_int),
    "hasty_sense_normal_toeplitz_ex": ([c_vp, c_vp, c_vp, c_i64, c_vp, c_int, c_int, P_i32, c_vp, c_vp, c_int, c_int,
                                        P_i32, c_vp, c_int, P_i64, c_vp, c_sz, c_int, c_vp], c_int),
    "hasty_set_option": ([ctypes.c_char_p, c_i64], c_int),
    "hasty_get_option": ([ctypes.c_char_p, P_i64], c_int),
    "hasty_autotune_sense_normal": ([P_i64, c_int, c_int, P_i64, P_dbl, c_vp], c_int),
    "hasty_nufft_autotune": ([c_vp, c_vp, c_vp, P_i64, P_dbl, c_vp], c_int),
    "hasty_autotune_llr": ([P_i64, P_i64, P_i64, P_i64, P_dbl, c_vp], c_int),
    "hasty_timing_enable": ([c_int], c_int),
    "hasty_timing_entries": ([], c_int),
    "hasty_timing_read": ([c_int, ctypes.c_char_p, c_int, P_dbl, P_i64], c_int),
    "hasty_stream_probe": ([c_int, c_int, c_int, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), c_i64, c_vp], c_int),
    "hasty_graph_begin": ([c_vp], c_int),
    "hasty_graph_end": ([c_vp, ctypes.POINTER(c_vp)], c_int),
    "hasty_graph_launch": ([c_vp, c_vp], c_int),
    "hasty_graph_nodes": ([c_vp], c_i64),
    "hasty_graph_destroy": ([c_vp], c_int),
    "hasty_cg_workspace_size": ([c_i64, c_i64], c_sz),
    "hasty_cg_init": ([c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "hasty_cg_iterate": ([c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp], c_int),
    "hasty_cg_iterate_l2": ([c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp], c_int),
    "hasty_cg_init_precond": ([c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "hasty_cg_iterate_xr": ([c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp], c_int),
    "hasty_cg_iterate_pz": ([c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "hasty_circulant_precond_build": ([c_vp, c_i64, c_vp, c_int, c_i64, c_int, P_i64, c_vp, c_dbl, c_int, c_vp, c_dbl,
                                       c_vp], c_int),
    "hasty_circulant_precond_workspace_size": ([c_i64, c_int, P_i64, c_int, c_int], c_sz),
    "hasty_circulant_precond_apply": ([c_vp, c_vp, c_vp, c_i64, c_i64, P_i32, c_int, P_i64, c_int, c_int, c_vp, c_sz,
                                       c_vp], c_int),
    "hasty_dct3": ([c_vp, c_i64, P_i64, c_int, c_int, c_vp], c_int),
    "hasty_l1dct_prox": ([c_vp, c_i64, P_i64, c_int, c_dbl, c_vp], c_int),
    "hasty_llr_shifts": ([c_i64, P_i64, P_i32], c_int),
    "hasty_llr_workspace_size": ([P_i64, P_i64, P_i64], c_sz),
    "hasty_llr_normal_blocks": ([c_vp, P_i64, P_i64, P_i64, c_i64, P_i32, c_dbl, c_int, c_vp, c_sz, c_vp], c_int),
    "hasty_llr_blocks": ([c_vp, P_i64, P_i64, P_i32, c_i64, c_dbl, c_int, c_vp, c_sz, c_vp], c_int),
    "hasty_llr_blocks_workspace_size": ([P_i64, P_i64, c_i64], c_sz),
    "hasty_nufft_plan_shape": ([c_vp, P_i64, P_i64, P_i64], c_int),
    "hasty_llr_grid_corners": ([P_i64, P_i64, P_i64, P_i32, P_i32, c_i64, P_i64], c_int),
    "hasty_llr_blocks_dev": ([c_vp, P_i64, P_i64, c_vp, c_i64, c_dbl, c_int, c_vp, c_sz, c_vp], c_int),
    "hasty_svt_workspace_size": ([c_i64, c_i64], c_sz),
    "hasty_svt_matrix": ([c_vp, c_i64, c_i64, c_dbl, c_int, c_vp, c_sz, c_vp], c_int),
    "hasty_svt_matrix_noalloc": ([c_vp, c_i64, c_i64, c_dbl, c_int, c_vp], c_int),
}

ERROR_NAMES = {1: "invalid argument", 2: "HIP runtime error", 3: "not positive definite", 4: "unsupported",
               5: "out of memory"}


class HastyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def header_functions() -> list[str]:
    """Names of the functions include/hasty_hip.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hasty_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                    "(hipcc --offload-arch=gfx950); there is no CPU fallback")
            l = ctypes.CDLL(LIB_PATH)
            for name, (args, res) in _SIGS.items():
                fn = getattr(l, name)
                fn.argtypes = args
                fn.restype = res
            _lib = l
    return _lib


def last_error() -> str:
    return lib().hasty_last_error().decode(errors="replace")


def check(rc: int) -> None:
    if rc != 0:
        raise HastyError(rc, last_error() or ERROR_NAMES.get(rc, "error"))


def i64arr(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * max(1, len(vals)))(*vals)


def i32arr(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int32 * max(1, len(vals)))(*vals)


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def get_option(name: str) -> int:
    v = ctypes.c_int64()
    call("hasty_get_option", name.encode(), ctypes.byref(v))
    return int(v.value)


def set_option(name: str, value: int) -> None:
    call("hasty_set_option", name.encode(), int(value))
