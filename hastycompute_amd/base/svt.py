"""Block SVT operators (python/python/hastypy/base/svt.py:29-61) over
torch.classes.HastySVT, in place on [F, E, X, Y, Z] complex64 (GPU tensors stay on the
GPU)."""
from __future__ import annotations

from ..ffi import hasty_svt as ffisvt
from .opalg import Operator, Vector


class Random3DBlocksSVT(Operator):
    def __init__(self, input_shape, streams, nblocks: int, block_shape, thresh: float, soft: bool):
        self._svtop = ffisvt.Random3DBlocksSVT(streams)
        self.nblocks, self.block_shape, self.thresh, self.soft = nblocks, list(block_shape), thresh, soft
        super().__init__(input_shape, input_shape)

    def update_thresh(self, thresh):
        self.thresh = thresh

    def _apply(self, input: Vector):
        self._svtop.apply(input.get_tensor(), self.nblocks, self.block_shape, float(self.thresh), self.soft)
        return input


class Normal3DBlocksSVT(Operator):
    def __init__(self, input_shape, streams, block_strides, block_shape, block_iter: int, thresh: float,
                 soft: bool):
        self._svtop = ffisvt.Normal3DBlocksSVT(streams)
        self.block_strides, self.block_shape = list(block_strides), list(block_shape)
        self.block_iter, self.thresh, self.soft = block_iter, thresh, soft
        super().__init__(input_shape, input_shape)

    def update_thresh(self, thresh):
        self.thresh = thresh

    def _apply(self, input: Vector):
        self._svtop.apply(input.get_tensor(), self.block_strides, self.block_shape, self.block_iter,
                          float(self.thresh), self.soft)
        return input
