"""Operator objects of the C++ layer (lib/op/mriop.cpp:156-266, ``op::SenseNOp``).

The reference binds no Python class for ``SenseNOp``; this shim exposes the TorchScript
class ``torch.classes.HastyOp.SenseNOp`` with the operator's own method names.
"""
from __future__ import annotations

import torch

from . import hasty_ffi

hasty_ffi.load()
_op = torch.classes.HastyOp


class SenseNOp:
    """A^H A for one trajectory, coil-map stack and coil list (GPU tensors).

    apply(x) = A^H A x; apply_forward(x) = A x, k-space [len(coils), M];
    apply_backward(y) = A^H y, image [1, N...]; adjoint() is the operator itself.
    """

    def __init__(self, coords: torch.Tensor, nmodes: list[int], smaps: torch.Tensor, coils: list[int],
                 forward_opts=None, backward_opts=None):
        unwrap = lambda o: getattr(o, "_nufftopts", o)  # hasty_nufft.NufftOptions or the TorchScript object
        self._op = _op.SenseNOp(coords, nmodes, smaps, coils, unwrap(forward_opts), unwrap(backward_opts))

    @classmethod
    def _wrap(cls, op):
        obj = cls.__new__(cls)
        obj._op = op
        return obj

    def apply(self, x):
        if isinstance(x, (list, tuple)):
            return self._op.apply_list(list(x))
        return self._op.apply(x)

    def apply_forward(self, x: torch.Tensor) -> torch.Tensor:
        return self._op.apply_forward(x)

    def apply_backward(self, y: torch.Tensor) -> torch.Tensor:
        return self._op.apply_backward(y)

    def adjoint(self) -> "SenseNOp":
        return SenseNOp._wrap(self._op.adjoint())
