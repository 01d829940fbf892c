"""Circulant preconditioning for the CG-SENSE loop on GPU-resident tensors.

`circulant_precond` corresponds to hastypy.base.torch_precond.circulant_precond
(python/python/hastypy/base/torch_precond.py:60-101) and `PrecondLinop` to
reconstruct_util.PrecondLinop (python/python/hastypy/reconstruct_util.py:436-464). Both
run through the C ABI (hasty_circulant_precond_build / _apply), nothing on the CPU.

The reference's Python builder multiplies an uncentred autocorrelation by the centred
PSF; the diagonal here is the consistent one of CirculantPreconditionerOp::build_diagonal
(lib/mri/precond.cpp:65-117, sigpy's method), returned in FFT order to match the
uncentred transforms PrecondLinop applies: out = fftn(ifftn(x) * P) per frame.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from .. import native
from .opalg import Linop, Vector


def circulant_precond(smaps: torch.Tensor, coord: torch.Tensor, weights: Optional[torch.Tensor] = None,
                      lamda: float = 0.0) -> torch.Tensor:
    """Diagonal [N...] (complex64, FFT order) of the circulant preconditioner of the SENSE
    normal operator with these maps and samples."""
    w = None
    if weights is not None:
        w = torch.real(weights).reshape(-1).to(device=smaps.device, dtype=torch.float32)
    return native.circulant_precond_diagonal(smaps.to(torch.complex64).contiguous(), coord.to(smaps.device),
                                             w, lamda=lamda, centered=False)


class PrecondLinop(Linop):
    """[F, E, N...] -> [F, E, N...]: frame f's preconditioner on each of its encodings."""

    def __init__(self, smaps: torch.Tensor, coord_vec: Sequence[torch.Tensor],
                 weights_vec: Optional[Sequence[torch.Tensor]] = None, clone: bool = True, lamda: float = 0.0):
        shape = tuple([len(coord_vec), 1] + list(smaps.shape[1:]))
        self.clone = clone
        self.Pvec = torch.stack([circulant_precond(smaps, coord_vec[i],
                                                   weights_vec[i] if weights_vec is not None else None, lamda)
                                 for i in range(len(coord_vec))])
        super().__init__(shape, shape)

    def _apply(self, input: Vector) -> Vector:
        x = input.get_tensor()
        F, E = x.shape[0], x.shape[1]
        xv = x.reshape((F * E,) + tuple(x.shape[2:])).contiguous()
        out = xv if not self.clone else torch.empty_like(xv)
        # fftn(ifftn(x) * P) = the adjointified apply with conj(P)
        native.circulant_precond_apply(xv, torch.conj(self.Pvec).resolve_conj().contiguous(), out=out,
                                       diag_index=[f for f in range(F) for _ in range(E)], centered=False,
                                       adjoint=True)
        res = out.reshape(x.shape)
        if not self.clone:
            x.copy_(res)
            return input
        return Vector(res)
