"""Proximal operators (python/python/hastypy/base/proximal.py:13-221).

As in the reference, ProximalOperator.__call__(input, alpha) ignores alpha
(proximal.py:39-40): the effective threshold of Proximal3DSVT is its base_alpha, the
absolute `thresh` of its SVTOptions, whatever the solver's step size.
"""
from __future__ import annotations

from .. import native
from . import opalg
from .opalg import Linop, Vector
from .svt import Normal3DBlocksSVT, Random3DBlocksSVT


class ProximalOperator:
    def __init__(self, shape, base_alpha, inplace):
        self.base_alpha = base_alpha if base_alpha is not None else 1.0
        self.inplace = inplace
        self.shape = shape

    def get_base_alpha(self):
        return self.base_alpha

    def set_base_alpha(self, base_alpha):
        self.base_alpha = base_alpha

    def apply(self, input: Vector, alpha=None) -> Vector:
        if opalg.get_shape(input) != self.shape:
            raise RuntimeError("Incompatible input shape for ProximalOperator")
        mul = self.base_alpha * (1.0 if alpha is None else alpha)
        output = self._apply(mul, input)
        if opalg.get_shape(output) != self.shape:
            raise RuntimeError("Incompatible output shape for ProximalOperator")
        return output

    def __call__(self, input: Vector, alpha=None) -> Vector:
        return self.apply(input)


class ZeroFunc(ProximalOperator):
    """Proximal operator of f = 0 (the identity)."""

    def __init__(self, shape):
        super().__init__(shape, None, inplace=True)

    def _apply(self, alpha, input):
        return input


class UnitaryTransform(ProximalOperator):
    """prox of phi(Q x) for unitary Q: Q^H prox_phi(Q x) (proximal.py:81-92)."""

    def __init__(self, prox: ProximalOperator, unitary: Linop, unitary_adjoint: Linop, base_alpha=None, inplace=False):
        if unitary.ishape != unitary_adjoint.oshape:
            raise RuntimeError("Unitary input shape and UnitaryAdjoint output shape did not match")
        self.prox, self.unitary, self.unitary_adjoint = prox, unitary, unitary_adjoint
        super().__init__(unitary.ishape, base_alpha, inplace)

    def _apply(self, alpha, input):
        return self.unitary_adjoint(self.prox(self.unitary(input), alpha))


class ProximalL1DCT(ProximalOperator):
    """prox of lam |DCT(x)|_1 over the axes of `axis` (proximal.py:134-183): per volume
    dctn, c <- exp(i angle c) max(|c| - alpha, 0), idctn (scipy's unnormalised DCT-II,
    norm=None), on the GPU through hasty_l1dct_prox."""

    def __init__(self, shape, axis, lamda, inplace=True):
        self.axis = tuple(bool(a) for a in axis)
        super().__init__(shape, lamda, inplace)

    @staticmethod
    def tensor_apply(input, lamda, axis):
        native.l1dct_prox_(input, axis, float(lamda))

    def _apply(self, alpha, input):
        output = input if self.inplace else input.clone()
        for t in output.get_tensorlist():
            ProximalL1DCT.tensor_apply(t, alpha, self.axis)
        return output


class L2Reg(ProximalOperator):
    """prox of phi(x) + (rho/2)|x - a|^2 (proximal.py:108-134)."""

    def __init__(self, shape, prox: ProximalOperator | None = None, rho=1.0, a=None, base_alpha=None, inplace=False):
        self.prox, self.rho, self.a = prox, rho, a
        super().__init__(shape, base_alpha, inplace)

    def _apply(self, alpha, input):
        output = input if self.inplace else input.clone()
        mult = alpha * self.rho
        denom = 1.0 + mult
        if self.a is not None:
            output += mult * self.a
        output /= denom
        if self.prox is not None:
            return self.prox(output, alpha / denom)
        return output


class SVTOptions:
    def __init__(self, block_shapes=(16, 16, 16), block_strides=(16, 16, 16), block_iter: int = 4, random=False,
                 nblocks=1, thresh: float = 0.01, soft: bool = True):
        self.block_shape = list(block_shapes)
        self.block_strides = list(block_strides)
        self.block_iter = block_iter
        self.random = random
        self.nblocks = nblocks
        self.thresh = thresh
        self.soft = soft


class Proximal3DSVT(ProximalOperator):
    """LLR prox over [F, E, X, Y, Z] (proximal.py:186-221) with the GPU block SVT."""

    def __init__(self, shape, svt_options: SVTOptions, streams=None, inplace=True):
        if svt_options.random:
            self._svtop = Random3DBlocksSVT(shape, streams, svt_options.nblocks, svt_options.block_shape,
                                            svt_options.thresh, svt_options.soft)
        else:
            self._svtop = Normal3DBlocksSVT(shape, streams, svt_options.block_strides, svt_options.block_shape,
                                            svt_options.block_iter, svt_options.thresh, svt_options.soft)
        super().__init__(shape, svt_options.thresh, inplace)

    def _apply(self, alpha, input):
        output = input if self.inplace else input.clone()
        self._svtop.update_thresh(alpha)
        self._svtop._apply(output)
        return output
