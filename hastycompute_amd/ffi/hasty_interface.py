"""TorchScript callbacks (python/python/hastypy/ffi/hasty_interface.py)."""
from __future__ import annotations

import torch

from . import hasty_ffi

hasty_ffi.load()
_mod = torch.classes.HastyInterface


class FunctionLambda:
    """``entry(inout, captures)`` compiled from TorchScript source ``script``, e.g.::

        def apply(a: List[Tensor], b: List[Tensor]):
            for i in range(len(a)):
                a[i] *= b[0]
    """

    def __init__(self, script: str, entry: str, captures: list[torch.Tensor]):
        self._funclam = _mod.FunctionLambda(script, entry, captures)

    def apply(self, input: list[torch.Tensor]):
        self._funclam.apply(input)
