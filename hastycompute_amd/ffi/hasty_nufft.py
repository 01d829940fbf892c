"""NUFFT operators (python/python/hastypy/ffi/hasty_nufft.py); all tensors on the GPU.

``coords`` is [ndim, M] in radians (the last row is the fastest image axis), ``nmodes`` is
[ntransf, N...]. Type 2 maps [ntransf, N...] -> [ntransf, M], type 1 the reverse; both are
scaled by 1/sqrt(prod N).
"""
from __future__ import annotations

import torch

from . import hasty_ffi
from .hasty_interface import FunctionLambda

hasty_ffi.load()
_mod = torch.classes.HastyNufft


class NufftOptions:
    def __init__(self, type: int, positive: bool | None = None, tol: float | None = None):
        self._nufftopts = _mod.NufftOptions(type, positive, tol)

    @staticmethod
    def type2():
        return NufftOptions(2, False, 1e-5)

    @staticmethod
    def type1():
        return NufftOptions(1, True, 1e-5)


class Nufft:
    def __init__(self, coord: torch.Tensor, nmodes: list[int], opts: NufftOptions):
        self._nufftopt = _mod.Nufft(coord, nmodes, opts._nufftopts)

    def apply(self, input: torch.Tensor, output: torch.Tensor):
        self._nufftopt.apply(input, output)


class NufftNormal:
    def __init__(self, coord: torch.Tensor, nmodes: list[int], forward_opts: NufftOptions, backward_opts: NufftOptions):
        self._nufftopt = _mod.NufftNormal(coord, nmodes, forward_opts._nufftopts, backward_opts._nufftopts)

    def apply(self, input: torch.Tensor, output: torch.Tensor, storage: torch.Tensor,
              callable: FunctionLambda | None = None):
        self._nufftopt.apply(input, output, storage, callable._funclam if callable is not None else None)
