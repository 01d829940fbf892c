"""Reconstruction drivers (python/python/hastypy/base/recon.py:14-205).

FivePointFULL: gradient descent (FISTA) on the 5 encodings with all frames' samples
stacked, with an optional L1-DCT prox (recon.py:27-83).

FivePointLLR: FISTA-accelerated proximal gradient for a 5-point 4D-flow scan, the
gradient A^H W (A x - y) through the batched Toeplitz normal operator and the LLR prox
over [F, 5, N0, N1, N2] (Casorati rows (encoding, x, y, z), one column per frame).
The step size is 0.5 / lambda_max over a random eighth of the frame-encodings
(recon.py:125-141); the power iterations of that subset run as one batch here.
"""
from __future__ import annotations

import numpy as np
import torch

from . import opalg
from . import proximal as prox
from .opalg import Vector
from .sense import BatchedSenseAdjoint, BatchedSenseNormal, InnerOuterAutomorphism, OuterInnerAutomorphism
from .solvers import GradientDescent


class GriddingReconstruction:
    """A^H W y / sum_c |S_c|^2 (recon.py:14-24)."""

    def __init__(self, coord_vec, smaps, kdata_vec, weights_vec):
        self.coord_vec, self.smaps, self.kdata_vec, self.weights_vec = coord_vec, smaps, kdata_vec, weights_vec

    def run(self):
        op = BatchedSenseAdjoint(self.coord_vec, self.smaps, None, self.weights_vec)
        den = torch.sum(self.smaps.conj() * self.smaps, dim=0).unsqueeze(0)
        return op(Vector(self.kdata_vec)) / Vector(den)


def batched_max_eig(coord_vec, smaps, weights_vec=None, max_iter=8, generator=None, x0=None):
    """Power iteration (MaxEig, opalg.py:647-674) of each frame-encoding's normal
    operator A_p^H W_p A_p, all in one batch; returns the per-problem estimates [P].
    x0 [P, 1, N...] is the start (default: torch.rand, as MaxEig)."""
    P = len(coord_vec)
    op = BatchedSenseNormal(coord_vec, smaps, None, weights_vec)
    shape = (P, 1) + tuple(smaps.shape[1:])
    if x0 is not None:
        x = x0.to(smaps.device, torch.complex64).reshape(shape).clone()
    else:
        x = torch.rand(shape, dtype=torch.complex64, device=smaps.device, generator=generator)
    lam = torch.full((P,), float("inf"), device=smaps.device)
    for _ in range(max_iter):
        y = op(Vector(x)).get_tensor()
        lam = torch.linalg.vector_norm(y.reshape(P, -1), dim=1)
        x = y / lam.view(P, 1, 1, 1, 1).to(y.dtype)
    return lam


def stack_frame_datas_fivepoint(coord_vec, kdata_vec):
    """recon.py:27-44: per encoding, the coordinates of all frames concatenated along the
    sample axis (dim 1) and their k-space data along dim 2."""
    new_coord_vec, new_kdata_vec = [], []
    nframes = len(coord_vec) // 5
    for enc in range(5):
        new_coord_vec.append(torch.cat([coord_vec[f * 5 + enc] for f in range(nframes)], dim=1))
        new_kdata_vec.append(torch.cat([kdata_vec[f * 5 + enc] for f in range(nframes)], dim=2))
    return new_coord_vec, new_kdata_vec


class FivePointFULL:
    """recon.py:47-83: max eigenvalue of encoding 0's normal operator (5 power
    iterations), step 1 / max_eig, gradient descent with the batched normal operator and,
    for lamda != 0, ProximalL1DCT with threshold lamda / step (the prox's step argument is
    ignored, as ProximalOperator.__call__ does). max_eig may be given to skip the power
    iteration (it starts from a random vector)."""

    def __init__(self, smaps, coord_vec, kdata_vec, weights_vec=None, streams=None, solver="GD", lamda=0.0,
                 max_eig=None):
        self.smaps, self.coord_vec, self.kdata_vec, self.weights_vec = smaps, coord_vec, kdata_vec, weights_vec
        self.streams, self.solver = streams, solver
        if max_eig is None:
            self.max_eig_run()
        else:
            self.max_eig = torch.tensor(float(max_eig), dtype=torch.float32)
        self.stepsize = 1 / self.max_eig
        if self.solver == "GD":
            self.sensenormalop = BatchedSenseNormal(self.coord_vec, self.smaps, self.kdata_vec, self.weights_vec,
                                                    self.streams)
        else:
            raise RuntimeError(f"solver {solver!r} is not supported (GD is)")
        self.lamda = lamda / self.stepsize
        if float(self.lamda) != 0.0:
            self._dctprox = prox.ProximalL1DCT((5, 1) + tuple(self.smaps.shape[1:]), (True, True, True),
                                               float(self.lamda))
        else:
            self._dctprox = None

    def max_eig_run(self, print_info=False):
        w = None if self.weights_vec is None else [self.weights_vec[0]]
        maxeigop = BatchedSenseNormal([self.coord_vec[0]], self.smaps, None, w)
        self.max_eig = opalg.MaxEig(maxeigop, torch.complex64, max_iter=5).run(print_info=print_info).to(
            torch.float32)

    def run(self, image, iter=50, callback=None, accelerate=True, print_info=False):
        image = Vector(image)
        if self.solver == "GD":
            gd = GradientDescent(self.sensenormalop, image, self._dctprox, float(1 / self.max_eig),
                                 accelerate=accelerate, max_iter=iter)
            image = gd.run(callback, print_info=print_info)
        return image.get_tensor()


class FivePointLLR:
    def __init__(self, smaps, coord_vec, kdata_vec, svt_options: prox.SVTOptions | None, weights_vec=None,
                 grad_streams=None, prox_streams=None, solver="GD", tau=None, sigma=None, seed=None):
        self.smaps, self.coord_vec, self.kdata_vec, self.weights_vec = smaps, coord_vec, kdata_vec, weights_vec
        self.svt_options = svt_options
        self.nframes = len(coord_vec) // 5
        self.grad_streams, self.prox_streams = grad_streams, prox_streams
        self.solver = solver
        rng = np.random.default_rng(seed)
        gen = None
        if seed is not None:
            gen = torch.Generator(device=smaps.device)
            gen.manual_seed(seed)

        if self.solver == "GD":
            perm = rng.permutation(len(self.coord_vec))[:max(1, len(self.coord_vec) // 8)]
            sub_w = [self.weights_vec[p] for p in perm] if self.weights_vec is not None else None
            eigs = batched_max_eig([self.coord_vec[p] for p in perm], smaps, sub_w, max_iter=8, generator=gen)
            self.max_eig = float(eigs.max())
            self.stepsize = 0.5 / self.max_eig
            self.sensenormalop = BatchedSenseNormal(self.coord_vec, self.smaps, self.kdata_vec, self.weights_vec,
                                                    self.grad_streams)
        elif self.solver == "Gridding":
            pass
        else:
            raise RuntimeError(f"solver {solver!r} is not supported (GD and Gridding are)")

        sp = tuple(self.smaps.shape[1:])
        self.svtshape = (self.nframes, 5) + sp
        self.gradshape = (self.nframes * 5, 1) + sp
        self.tosvtshape = InnerOuterAutomorphism(self.nframes, 5, sp)
        self.togradshape = OuterInnerAutomorphism(self.nframes, 5, sp)
        if self.svt_options is not None:
            self.svtprox = prox.UnitaryTransform(prox.Proximal3DSVT(self.svtshape, self.svt_options, self.prox_streams),
                                                 self.tosvtshape, self.togradshape)
        else:
            self.svtprox = prox.ZeroFunc(self.gradshape)

    def run(self, image, iter=50, callback=None, accelerate=True, print_info=False):
        image = self.togradshape(Vector(image))
        if self.solver == "GD":
            gd = GradientDescent(self.sensenormalop, image, self.svtprox, self.stepsize, accelerate=accelerate,
                                 max_iter=iter)
            image = gd.run(callback, print_info=print_info)
        elif self.solver == "Gridding":
            image = GriddingReconstruction(self.coord_vec, self.smaps, self.kdata_vec, self.weights_vec).run()
        return self.tosvtshape(image).get_tensor()
