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


class PowerIteration:
    """op::PowerIteration (lib/op/opalgs.cpp:29-44): run(A, v, iters) replaces v by the
    normalised A^iters v and returns the last |A v| (the largest eigenvalue estimate)."""

    def __init__(self):
        self._alg = _op.PowerIteration()

    def run(self, A: SenseNOp, v: torch.Tensor, iters: int = 30) -> torch.Tensor:
        return self._alg.run(A._op, v, iters)


class ConjugateGradient:
    """op::ConjugateGradient (opalgs.cpp:50-96) on A = SenseNOp (+ l2 I, the reference's
    l2_reg_linop): run(x, iter, tol) updates x in place; raises "A was not positive
    definite" as the reference does."""

    def __init__(self, A: SenseNOp, b: torch.Tensor, l2: float = 0.0):
        self._alg = _op.ConjugateGradient(A._op, b, float(l2))

    def run(self, x: torch.Tensor, iter: int = 30, tol: float = 0.0) -> None:
        self._alg.run(x, iter, tol)


class BatchedConjugateGradient:
    """op::BatchedConjugateGradient (opalgs.ixx:84-136) over every (frame, encoding) of a
    hasty_sense.BatchedSenseNormal built without k-space data: the F * E problems of x, b
    [F, E, N...] (GPU) go through the device CG together. run returns the per-problem
    state [F * E, 8] (rzold, alpha, beta, done, not positive definite, pAp, iterations)."""

    def __init__(self, A, b: torch.Tensor):
        self._alg = _op.BatchedConjugateGradient(getattr(A, "_senseop", A), b)

    def run(self, x: torch.Tensor, iter: int = 30, tol: float = 0.0) -> torch.Tensor:
        return self._alg.run(x, iter, tol)
