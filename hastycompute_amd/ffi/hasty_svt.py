"""Block singular value thresholding (python/python/hastypy/ffi/hasty_svt.py).

``input`` is [F, E, X, Y, Z] complex64 (CPU or GPU). Each block is the Casorati matrix
with rows (e, x, y, z) and one column per frame.
"""
from __future__ import annotations

import torch

from . import hasty_ffi

hasty_ffi.load()
_mod = torch.classes.HastySVT


class Random3DBlocksSVT:
    def __init__(self, streams: list[torch.Stream] | None = None):
        self._svtop = _mod.Random3DBlocksSVT(streams)

    def apply(self, input: torch.Tensor, nblocks: int, block_shape: list[int], thresh: float, soft: bool):
        self._svtop.apply(input, nblocks, block_shape, thresh, soft)

    def last_corners(self) -> torch.Tensor:
        """[nblocks, 3] corners (x, y, z) of the last apply, in the order they were
        thresholded (an addition: the reference's draws cannot be replayed)."""
        return self._svtop.last_corners()


class Normal3DBlocksSVT:
    def __init__(self, streams: list[torch.Stream] | None = None):
        self._svtop = _mod.Normal3DBlocksSVT(streams)

    def apply(self, input: torch.Tensor, block_strides: list[int], block_shape: list[int], block_iter: int,
              thresh: float, soft: bool):
        self._svtop.apply(input, block_strides, block_shape, block_iter, thresh, soft)
