"""Batched SENSE linear operators (python/python/hastypy/base/sense.py:123-424) over
torch.classes.HastyBatchedSense, with the image layout [outer, inner, N0, N1, N2].

Unlike the reference (CPU tensors, per-frame transfers), outputs are allocated on the
device of the coil maps, so a solver's iterates stay in HBM.
"""
from __future__ import annotations

import numpy as np
import torch

from ..ffi import hasty_sense
from .opalg import Linop, Vector


def _dtype(coord_vec):
    if coord_vec[0].dtype == torch.float32:
        return torch.complex64
    if coord_vec[0].dtype == torch.float64:
        return torch.complex128
    raise RuntimeError("Non supported fp type for coord")


class _Batched(Linop):
    def __init__(self, coord_vec, smaps, coils, ninner_batches):
        self.nouter_batches = len(coord_vec)
        self.ninner_batches = ninner_batches
        self.ncoils = smaps.shape[0]
        self.coils = coils
        self.dimensions = tuple(smaps.shape[1:])
        self.dtype = _dtype(coord_vec)
        self.device = smaps.device
        self.nfreqs = [c.shape[1] for c in coord_vec]
        self.set_coil_randomization()

    def set_coil_randomization(self, randomize=False, num_rand_coils=16):
        self.randomize_coils = randomize
        self.num_rand_coils = num_rand_coils

    def get_coils(self):
        if self.randomize_coils:
            return [np.random.permutation(self.ncoils).astype(np.int32)[:self.num_rand_coils].tolist()
                    for _ in range(self.nouter_batches)]
        return self.coils


class BatchedSense(_Batched):
    """x [F, E, N...] -> list of [E, C, M_f] (with kdata: A x - y; with weights: W(.))."""

    def __init__(self, coord_vec, smaps, kdata_vec=None, weights_vec=None, streams=None, coils=None,
                 ninner_batches=1):
        super().__init__(coord_vec, smaps, coils, ninner_batches)
        self._senseop = hasty_sense.BatchedSense(coord_vec, smaps, kdata_vec, weights_vec, streams)
        ncl = self.ncoils if coils is None else len(coils[0])
        self.output = [torch.empty((ninner_batches, ncl, m), dtype=self.dtype, device=self.device)
                       for m in self.nfreqs]
        super(_Batched, self).__init__((self.nouter_batches, ninner_batches, *self.dimensions),
                                       [tuple(o.shape) for o in self.output])

    def _apply(self, input: Vector) -> Vector:
        self._senseop.apply(input.get_tensor(), self.output, self.get_coils())
        return Vector([o.clone() for o in self.output])


class BatchedSenseAdjoint(_Batched):
    """list of [E, C, M_f] -> [F, E, N...] (with kdata: A^H (y_in - y); weights: A^H W)."""

    def __init__(self, coord_vec, smaps, kdata_vec=None, weights_vec=None, streams=None, coils=None,
                 ninner_batches=1):
        super().__init__(coord_vec, smaps, coils, ninner_batches)
        self._senseop = hasty_sense.BatchedSenseAdjoint(coord_vec, smaps, kdata_vec, weights_vec, streams)
        ncl = self.ncoils if coils is None else len(coils[0])
        ishape = [(ninner_batches, ncl, m) for m in self.nfreqs]
        super(_Batched, self).__init__(ishape, (self.nouter_batches, ninner_batches, *self.dimensions))

    def _apply(self, input: Vector) -> Vector:
        out = torch.empty(self.oshape, dtype=self.dtype, device=self.device)
        self._senseop.apply(input.get_tensorlist(), out, self.get_coils())
        return Vector(out)


class BatchedSenseNormal(_Batched):
    """[F, E, N...] -> [F, E, N...]: A^H W (A x - y) per (outer, inner) batch, through the
    Toeplitz-embedded normal operator (one kernel per outer batch)."""

    def __init__(self, coord_vec, smaps, kdata_vec=None, weights_vec=None, streams=None, coils=None,
                 ninner_batches=1):
        super().__init__(coord_vec, smaps, coils, ninner_batches)
        self._senseop = hasty_sense.BatchedSenseNormal(coord_vec, smaps, kdata_vec, weights_vec, streams)
        shape = (self.nouter_batches, ninner_batches, *self.dimensions)
        super(_Batched, self).__init__(shape, shape)

    def _apply(self, input: Vector) -> Vector:
        out = torch.empty(self.oshape, dtype=self.dtype, device=self.device)
        self._senseop.apply(input.get_tensor().contiguous(), out, self.get_coils())
        return Vector(out)


class OuterInnerAutomorphism(Linop):
    """(nf, ne, ...) -> (nf * ne, 1, ...) (sense.py:402-409)."""

    def __init__(self, nouter_batches, ninner_batches, sp_dims):
        super().__init__((nouter_batches, ninner_batches) + tuple(sp_dims),
                         (nouter_batches * ninner_batches, 1) + tuple(sp_dims))

    def _apply(self, input: Vector) -> Vector:
        return Vector(input.get_tensor().view(self.oshape))


class InnerOuterAutomorphism(Linop):
    """(nf * ne, 1, ...) -> (nf, ne, ...) (sense.py:411-418)."""

    def __init__(self, nouter_batches, ninner_batches, sp_dims):
        super().__init__((nouter_batches * ninner_batches, 1) + tuple(sp_dims),
                         (nouter_batches, ninner_batches) + tuple(sp_dims))

    def _apply(self, input: Vector) -> Vector:
        return Vector(input.get_tensor().view(self.oshape))
