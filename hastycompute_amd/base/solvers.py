"""Iterative solvers (python/python/hastypy/base/solvers.py:11-251).

ConjugateGradient and GradientDescent keep the reference's update rules, stopping rules
and attribute names; with the operators of hastycompute_amd.base.sense every vector
stays on the GPU. For many independent CG problems in one batch use
hastycompute_amd.native.cg_solve (device CG with FP64 scalars).
"""
from __future__ import annotations

import math
import time

import torch

from . import opalg
from .opalg import IterativeAlg, Linop, Operator, Vector


class ConjugateGradient(IterativeAlg):
    """solvers.py:11-83: r = b - A x, z = P r, p = z; per iteration alpha = rz / pAp,
    x += alpha p, r -= alpha Ap, beta = rz_new / rz, p = z + beta p; stops at max_iter,
    on pAp <= 0 (not positive definite) or when sqrt(rz) <= tol."""

    def __init__(self, A: Linop, b: Vector, x: Vector, P: Linop | None = None, tol=0.0, max_iter=100):
        self.A, self.b, self.x, self.P, self.tol = A, b, x, P, tol
        self.r = self.b - self.A(self.x)
        z = self.r if self.P is None else self.P(self.r)
        self.p = z.clone() if max_iter > 1 else z
        self.positive_definite = True
        self.rzold = torch.real(opalg.vdot(self.r, z))
        self.resid = torch.sqrt(self.rzold)
        self.resids = [float(self.resid)]
        super().__init__(max_iter)

    def _update(self):
        Ap = self.A(self.p)
        pAp = torch.real(opalg.vdot(self.p, Ap))
        if pAp <= 0:
            self.positive_definite = False
            return
        alpha = Vector(self.rzold / pAp)
        self.x += self.p * alpha
        if self.iter < self.max_iter - 1:
            self.r -= Ap * alpha
            z = self.P(self.r) if self.P is not None else self.r
            rznew = torch.real(opalg.vdot(self.r, z))
            beta = Vector(rznew / self.rzold)
            self.p *= beta
            self.p += z
            self.rzold = rznew
        self.resid = torch.sqrt(self.rzold)
        self.resids.append(float(self.resid))

    def _done(self):
        return self.iter >= self.max_iter or not self.positive_definite or self.resid <= self.tol

    def run(self, callback=None, print_info=False):
        i = 0
        while not self.done():
            self.update()
            if print_info:
                print("CG Iter: ", i, "/", self.max_iter, ", Res: ", self.resids[-1])
            if callback is not None:
                callback(self.x, i)
            i += 1
        return self.x


class GradientDescent(IterativeAlg):
    """solvers.py:85-162: x -= alpha grad f(x); x = prox(x, alpha); with accelerate the
    FISTA momentum t' = (1 + sqrt(1 + 4 t^2)) / 2, z = x + ((t - 1) / t') (x - x_old).
    resids[i] = |x - x_old| (or |z| before the momentum step when accelerated)."""

    def __init__(self, gradf: Operator, x: Vector, prox=None, alpha=1.0, accelerate=False, max_iter=100, tol=0.0):
        self.gradf, self.x, self.alpha, self.accelerate, self.tol, self.prox = gradf, x, alpha, accelerate, tol, prox
        self.resids = [-1]
        self.rel_resids = [-1]
        self.grad_time = self.prox_time = 0.0
        if self.accelerate:
            self.z = self.x.clone()
            self.t = 1
        super().__init__(max_iter)

    def _update(self):
        x_old = self.x.clone()
        if self.accelerate:
            self.x = self.z
        t0 = time.time()
        self.x -= self.alpha * self.gradf(self.x)
        self.grad_time = time.time() - t0
        if self.prox is not None:
            t0 = time.time()
            self.x = self.prox(self.x, self.alpha)
            self.prox_time = time.time() - t0
        if self.accelerate:
            t_old = self.t
            self.t = (1.0 + math.sqrt(1.0 + 4.0 * t_old * t_old)) / 2.0
            self.z = self.x - x_old
            self.resids.append(opalg.norm(self.z))
            self.z = self.x + ((t_old - 1) / self.t) * self.z
        else:
            self.resids.append(opalg.norm(self.x - x_old))
        if self.tol != 0.0:
            self.rel_resids.append(self.resids[-1] / opalg.norm(self.x))

    def _done(self):
        return self.iter >= self.max_iter or (self.rel_resids[-1] < self.tol and self.rel_resids[-1] > 0.0)

    def run(self, callback=None, print_info=False):
        i = 0
        while not self.done() or i == 0:
            self.update()
            if print_info:
                msg = f"GD Iter: {i} / {self.max_iter}, Res: {float(self.resids[-1])}, GradTime: {self.grad_time}"
                if self.prox is not None:
                    msg += f", ProxTime: {self.prox_time}"
                print(msg)
            if callback is not None:
                callback(self.x, i)
            i += 1
        return self.x


class PrimalDualHybridGradient(IterativeAlg):
    """solvers.py:164-251, Chambolle-Pock with the reference's update order:
    y += sigma K(x_bar); y = proxfc(y, sigma); x -= tau K^H(y); x = proxg(x, tau);
    theta = 1 / sqrt(1 + 2 acceleration tau_min) (tau *= theta, sigma /= theta unless
    theta == 1); x_bar = x + theta (x - x_bar_previous). The extrapolation uses the
    previous x_bar, not the previous x (the reference stores x_old and does not use it),
    and the prox calls ignore their step argument as ProximalOperator.__call__ does.
    classic=True extrapolates from the previous x (Chambolle-Pock), which converges where
    the reference's order may not."""

    def __init__(self, proxfc, proxg, K: Linop, KH: Linop, x: Vector, y: Vector, tau: Vector, sigma: Vector,
                 acceleration: float, max_iter=100, tol=0.0, classic: bool = False):
        self.classic = classic
        self.proxfc, self.proxg, self.K, self.KH = proxfc, proxg, K, KH
        self.x, self.y = x, y
        self.acceleration = acceleration
        self.tau, self.sigma = tau, sigma
        self.tau_mul = opalg.min(tau).item()
        self.xtemp = self.x.clone()
        self.tol = tol
        self.resids = [-1]
        self.rel_resids = [-1]
        self.theta = 1.0
        super().__init__(max_iter)

    def _update(self):
        self.y += self.sigma * self.K(self.xtemp)
        self.y = self.proxfc(self.y, self.sigma)
        self.xold = self.x.clone()
        self.x -= self.tau * self.KH(self.y)
        self.x = self.proxg(self.x, self.tau)
        self.theta = 1.0 / math.sqrt(1.0 + 2.0 * self.acceleration * self.tau_mul)
        if not math.isclose(self.theta, 1.0):
            self.tau_mul *= self.theta
            self.tau *= self.theta
            self.sigma /= self.theta
        self.xtemp = self.x - (self.xold if self.classic else self.xtemp)
        self.resids.append(opalg.norm(self.xtemp).item())
        self.xtemp = self.x + self.theta * self.xtemp
        if self.tol != 0.0:
            self.rel_resids.append(self.resids[-1] / opalg.norm(self.x))

    def _done(self):
        return self.iter >= self.max_iter or (self.rel_resids[-1] < self.tol and self.rel_resids[-1] > 0.0)

    def run(self, callback=None, print_info=False):
        i = 0
        while not self.done() or i == 0:
            t0 = time.time()
            self.update()
            if print_info:
                print(f"PDHG Iter: {i} / {self.max_iter}, Res: {self.resids[-1]}, Theta: {self.theta}, "
                      f"UpdateTime: {time.time() - t0}")
            if callback is not None:
                callback(self.x, i)
            i += 1
        return self.x

