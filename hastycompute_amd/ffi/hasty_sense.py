"""SENSE operators (python/python/hastypy/ffi/hasty_sense.py).

``Sense``/``SenseAdjoint``/``SenseNormal`` take GPU tensors. The batched operators take
an outer batch (frames) of trajectories; their tensors may live on the CPU, as in the
reference, or on the GPU. ``BatchedSenseNormal`` evaluates
A^H W (A x - y) per (frame, encoding) with the Toeplitz-embedded normal operator.
"""
from __future__ import annotations

import torch

from . import hasty_ffi

hasty_ffi.load()
_sense = torch.classes.HastySense
_batched = torch.classes.HastyBatchedSense


class Sense:
    def __init__(self, coords: torch.Tensor, nmodes: list[int]):
        self._senseop = _sense.Sense(coords, nmodes)

    def apply(self, input, output, smaps, coils, imspace_storage=None, kspace_storage=None):
        self._senseop.apply(input, output, smaps, coils, imspace_storage, kspace_storage)


class SenseAdjoint:
    def __init__(self, coords: torch.Tensor, nmodes: list[int]):
        self._senseop = _sense.SenseAdjoint(coords, nmodes)

    def apply(self, input, output, smaps, coils, imspace_storage=None, kspace_storage=None):
        self._senseop.apply(input, output, smaps, coils, imspace_storage, kspace_storage)


class SenseNormal:
    def __init__(self, coords: torch.Tensor, nmodes: list[int]):
        self._senseop = _sense.SenseNormal(coords, nmodes)

    def apply(self, input, output, smaps, coils, imspace_storage=None, kspace_storage=None):
        self._senseop.apply(input, output, smaps, coils, imspace_storage, kspace_storage)


class BatchedSense:
    def __init__(self, coords: list[torch.Tensor], smaps: torch.Tensor, kdata: list[torch.Tensor] | None = None,
                 weights: list[torch.Tensor] | None = None, streams: list[torch.Stream] | None = None):
        self._senseop = _batched.BatchedSense(coords, smaps, kdata, weights, streams)

    def apply(self, input: torch.Tensor, output: list[torch.Tensor], coils: list[list[int]] | None = None):
        self._senseop.apply(input, output, coils)


class BatchedSenseAdjoint:
    def __init__(self, coords: list[torch.Tensor], smaps: torch.Tensor, kdata: list[torch.Tensor] | None = None,
                 weights: list[torch.Tensor] | None = None, streams: list[torch.Stream] | None = None):
        self._senseop = _batched.BatchedSenseAdjoint(coords, smaps, kdata, weights, streams)

    def apply(self, input: list[torch.Tensor], output: torch.Tensor, coils: list[list[int]] | None = None):
        self._senseop.apply(input, output, coils)


class BatchedSenseNormal:
    def __init__(self, coords: list[torch.Tensor], smaps: torch.Tensor, kdata: list[torch.Tensor] | None = None,
                 weights: list[torch.Tensor] | None = None, streams: list[torch.Stream] | None = None):
        self._senseop = _batched.BatchedSenseNormal(coords, smaps, kdata, weights, streams)

    def apply(self, input: torch.Tensor, output: torch.Tensor, coils: list[list[int]] | None = None):
        self._senseop.apply(input, output, coils)
