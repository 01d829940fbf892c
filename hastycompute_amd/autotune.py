"""Launch-shape autotuner over the C ABI (hasty_autotune_sense_normal,
hasty_nufft_autotune, hasty_autotune_llr).

The library times the candidates with HIP events on the current stream, in a fixed
order starting from the defaults, keeps a challenger only when it is 2 % faster, and
caches the result per shape for the process; these wrappers return the choices as dicts
(the benchmark records them next to its numbers).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

from . import _capi as C


def sense_normal(nmodes3: Sequence[int], ncoil: int, nproblems: int) -> dict:
    """Plane kernels, problems per launch, the z-convolution kernel and problems per
    z-convolution launch, and coils per workspace of the batched Toeplitz SENSE normal
    operator. Sets the library options to the winners, so build SenseNormalToeplitz
    objects (which prepare their kernels) after this call, with the returned coil_batch
    and problem_batch."""
    choice = C.i64arr([0, 0, 0, 0, 0])
    ms = ctypes.c_double()
    C.call("hasty_autotune_sense_normal", C.i64arr(nmodes3), int(ncoil), int(nproblems), choice, ctypes.byref(ms),
           C.stream_ptr())
    return {"sense_xy_variant": int(choice[0]), "sense_zconv_batch": int(choice[1]),
            "problem_batch": int(choice[2]), "coil_batch": int(choice[3]), "sense_zconv_variant": int(choice[4]),
            "ms_per_problem": ms.value}


def nufft(plan, c, f) -> dict:
    """The 3-D type-1 spreader, spreading tile edge and points per workgroup of a
    NufftPlan, timed on its own points (c -> f; f is overwritten). The plan keeps the
    fastest."""
    v = ctypes.c_int64()
    ms = ctypes.c_double()
    C.call("hasty_nufft_autotune", plan._h, c.data_ptr(), f.data_ptr(), ctypes.byref(v), ctypes.byref(ms),
           C.stream_ptr())
    tile, chunk = ctypes.c_int64(), ctypes.c_int64()
    C.call("hasty_nufft_plan_shape", plan._h, ctypes.byref(tile), ctypes.byref(chunk), None)
    return {"nufft_spread_variant": int(v.value), "nufft_spread_tile": int(tile.value),
            "nufft_spread_chunk": int(chunk.value), "execute_ms": ms.value}


def llr(dims5: Sequence[int], block_strides: Sequence[int], block_shape: Sequence[int]) -> dict:
    """The Gram row split of the LLR prox for this image and block grid, then the
    eigensolver's workgroup size (the process-wide llr_eig_threads option)."""
    ns = ctypes.c_int64()
    ms = ctypes.c_double()
    C.call("hasty_autotune_llr", C.i64arr(dims5), C.i64arr(block_strides), C.i64arr(block_shape), ctypes.byref(ns),
           ctypes.byref(ms), C.stream_ptr())
    return {"llr_row_split": int(ns.value), "llr_eig_threads": C.get_option("llr_eig_threads"), "pass_ms": ms.value}
