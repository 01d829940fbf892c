"""Iterative solvers with the interface of python/python/hastypy/base/solvers.py:11-251.

The class names, constructor arguments, result attributes (x, resids, rel_resids,
positive_definite, grad_time, prox_time, theta) and the arithmetic of each step are the
reference's contract (cited per class); the driving loop, bookkeeping and reporting are
this module's own. With the operators of hastycompute_amd.base.sense every vector stays
on the GPU. For many independent CG problems in one batch use
hastycompute_amd.native.cg_solve (device CG with FP64 scalars).
"""
from __future__ import annotations

import math
import time

import torch

from . import opalg
from .opalg import IterativeAlg, Linop, Operator, Vector


class _Solver(IterativeAlg):
    """One loop for every solver: `_update` advances one iteration, `_report` returns the
    progress line, and `run` calls `callback(x, i)` after each iteration. `min_iters`
    iterations run even when the stopping rule already holds (the reference's gradient
    solvers take their first step unconditionally)."""

    name = "Iter"
    min_iters = 0

    def _report(self, i: int) -> str:
        return f"{self.name} iteration {i} of {self.max_iter}: residual {float(self.resids[-1]):.6g}"

    def run(self, callback=None, print_info=False):
        i = 0
        while i < self.min_iters or not self.done():
            self.update()
            if print_info:
                print(self._report(i))
            if callback is not None:
                callback(self.x, i)
            i += 1
        return self.x


def _real_dot(a: Vector, b: Vector) -> torch.Tensor:
    return torch.real(opalg.vdot(a, b))


class ConjugateGradient(_Solver):
    """ConjugateGradient (solvers.py:11-83; the C++ loop is opalgs.cpp:50-96). Step:
    alpha = <r, z> / <p, A p>, x += alpha p; unless it is the last allowed iteration,
    r -= alpha A p, z = P r, p = z + (<r, z>_new / <r, z>) p. Stops after max_iter
    iterations, when <p, A p> <= 0 (positive_definite becomes False) or when
    sqrt(<r, z>) <= tol. resids holds sqrt(<r, z>) after each iteration (the initial one
    first)."""

    name = "CG"

    def __init__(self, A: Linop, b: Vector, x: Vector, P: Linop | None = None, tol=0.0, max_iter=100):
        self.A, self.b, self.x, self.P, self.tol = A, b, x, P, tol
        self.r = b - A(x)
        z = self._precondition(self.r)
        # with one iteration p is never updated, so it may share storage with z
        self.p = z.clone() if max_iter > 1 else z
        self.rz = _real_dot(self.r, z)
        self.positive_definite = True
        self.resids = [float(torch.sqrt(self.rz))]
        super().__init__(max_iter)

    @property
    def rzold(self):
        return self.rz

    @property
    def resid(self):
        return self.resids[-1]

    def _precondition(self, r: Vector) -> Vector:
        return r if self.P is None else self.P(r)

    def _update(self):
        Ap = self.A(self.p)
        curvature = _real_dot(self.p, Ap)
        if not curvature > 0:
            self.positive_definite = False
            return
        step = Vector(self.rz / curvature)
        self.x += self.p * step
        last = self.iter == self.max_iter - 1
        if not last:
            self.r -= Ap * step
            z = self._precondition(self.r)
            rz_next = _real_dot(self.r, z)
            self.p *= Vector(rz_next / self.rz)
            self.p += z
            self.rz = rz_next
        self.resids.append(float(torch.sqrt(self.rz)))

    def _done(self):
        return self.iter >= self.max_iter or not self.positive_definite or self.resid <= self.tol


class GradientDescent(_Solver):
    """GradientDescent (solvers.py:85-162). Step from x (or, accelerated, from the
    extrapolated point z): x = prox(x - alpha grad f(x), alpha). Accelerated (FISTA):
    t' = (1 + sqrt(1 + 4 t^2)) / 2 and z = x + ((t - 1) / t') (x - x_previous).
    resids[i] = |x - x_previous|; with tol != 0, rel_resids[i] = resids[i] / |x| and the
    solver stops once 0 < rel_resids[-1] < tol. The first step always runs."""

    name = "GD"
    min_iters = 1

    def __init__(self, gradf: Operator, x: Vector, prox=None, alpha=1.0, accelerate=False, max_iter=100, tol=0.0):
        self.gradf, self.x, self.prox = gradf, x, prox
        self.alpha, self.accelerate, self.tol = alpha, accelerate, tol
        self.resids = [-1]
        self.rel_resids = [-1]
        self.grad_time = 0.0
        self.prox_time = 0.0
        if accelerate:
            self.z = x.clone()
            self.t = 1
        super().__init__(max_iter)

    def _report(self, i: int) -> str:
        line = f"GD iteration {i} of {self.max_iter}: residual {float(self.resids[-1]):.6g}, gradient {self.grad_time:.4f} s"
        return line + (f", prox {self.prox_time:.4f} s" if self.prox is not None else "")

    def _update(self):
        previous = self.x.clone()
        point = self.z if self.accelerate else self.x
        tick = time.time()
        point -= self.alpha * self.gradf(point)
        self.grad_time = time.time() - tick
        if self.prox is not None:
            tick = time.time()
            point = self.prox(point, self.alpha)
            self.prox_time = time.time() - tick
        self.x = point
        change = self.x - previous
        self.resids.append(opalg.norm(change))
        if self.accelerate:
            t = self.t
            self.t = (1.0 + math.sqrt(1.0 + 4.0 * t * t)) / 2.0
            self.z = self.x + ((t - 1) / self.t) * change
        if self.tol != 0.0:
            self.rel_resids.append(self.resids[-1] / opalg.norm(self.x))

    def _done(self):
        r = self.rel_resids[-1]
        return self.iter >= self.max_iter or 0.0 < r < self.tol


class PrimalDualHybridGradient(_Solver):
    """PrimalDualHybridGradient (solvers.py:164-251), Chambolle-Pock with the reference's
    update order: y = proxfc(y + sigma K(x_bar), sigma); x = proxg(x - tau K^H(y), tau);
    theta = 1 / sqrt(1 + 2 acceleration tau_min) (tau *= theta, sigma /= theta unless
    theta == 1); x_bar = x + theta (x - x_bar_previous). The extrapolation starts from the
    previous x_bar, not the previous x (the reference stores x_old and does not use it),
    and the prox calls ignore their step argument as ProximalOperator.__call__ does.
    classic=True extrapolates from the previous x (Chambolle-Pock), which converges where
    the reference's order may not."""

    name = "PDHG"
    min_iters = 1

    def __init__(self, proxfc, proxg, K: Linop, KH: Linop, x: Vector, y: Vector, tau: Vector, sigma: Vector,
                 acceleration: float, max_iter=100, tol=0.0, classic: bool = False):
        self.proxfc, self.proxg, self.K, self.KH = proxfc, proxg, K, KH
        self.x, self.y, self.tau, self.sigma = x, y, tau, sigma
        self.acceleration, self.tol, self.classic = acceleration, tol, classic
        self.tau_mul = opalg.min(tau).item()
        self.xtemp = x.clone()
        self.resids = [-1]
        self.rel_resids = [-1]
        self.theta = 1.0
        self._last = 0.0
        super().__init__(max_iter)

    def _report(self, i: int) -> str:
        return (f"PDHG iteration {i} of {self.max_iter}: residual {self.resids[-1]:.6g}, theta {self.theta:.6g}, "
                f"{self._last:.4f} s")

    def _update(self):
        tick = time.time()
        self.y += self.sigma * self.K(self.xtemp)
        self.y = self.proxfc(self.y, self.sigma)
        x_prev = self.x.clone()
        self.x -= self.tau * self.KH(self.y)
        self.x = self.proxg(self.x, self.tau)
        self.theta = 1.0 / math.sqrt(1.0 + 2.0 * self.acceleration * self.tau_mul)
        if not math.isclose(self.theta, 1.0):
            self.tau_mul *= self.theta
            self.tau *= self.theta
            self.sigma /= self.theta
        step = self.x - (x_prev if self.classic else self.xtemp)
        self.resids.append(opalg.norm(step).item())
        self.xtemp = self.x + self.theta * step
        if self.tol != 0.0:
            self.rel_resids.append(self.resids[-1] / opalg.norm(self.x))
        self._last = time.time() - tick

    def _done(self):
        r = self.rel_resids[-1]
        return self.iter >= self.max_iter or 0.0 < r < self.tol
