"""Vectors, operators and iterative-algorithm scaffolding
(python/python/hastypy/base/opalg.py).

A Vector is a tensor or a list of Vectors; arithmetic is elementwise and recursive.
`norm` and `vdot` of a list follow the reference: the norm of a list is the SUM of
its children's norms (opalg.py:434-441), not the norm of the concatenation.
"""
from __future__ import annotations

import torch


def _wrap(x):
    if isinstance(x, Vector):
        return x
    if isinstance(x, (int, float, complex)):
        # a zero-dim double tensor: it keeps the scalar's precision, and type promotion
        # still gives the result the dtype of the data it multiplies (float32 / complex64
        # data stay single precision)
        return Vector(torch.tensor(x, dtype=torch.complex128 if isinstance(x, complex) else torch.float64))
    return Vector(x)


def _binary(a: "Vector", b, fn):
    b = _wrap(b)
    if a.istensor() and b.istensor():
        return Vector(fn(a.tensor, b.tensor))
    if a.islist() and b.islist():
        return Vector([_binary(x, y, fn) for x, y in zip(a.children, b.children)])
    if a.islist():  # list op scalar tensor
        return Vector([_binary(x, b, fn) for x in a.children])
    return Vector([_binary(a, y, fn) for y in b.children])


def _inplace(a: "Vector", b, fn):
    b = _wrap(b)
    if a.istensor():
        fn(a.tensor, b.tensor if b.istensor() else b)
    elif b.islist():
        for x, y in zip(a.children, b.children):
            _inplace(x, y, fn)
    else:
        for x in a.children:
            _inplace(x, b, fn)
    return a


class Vector:
    def __init__(self, input):
        self.children = None
        self.tensor = None
        if isinstance(input, list):
            self.children = [i if isinstance(i, Vector) else Vector(i) for i in input]
        elif isinstance(input, torch.Tensor):
            self.tensor = input
        elif isinstance(input, float):
            self.tensor = torch.tensor(input, dtype=torch.float32)
        else:
            raise RuntimeError("Was not Vector, list or Tensor")

    def istensor(self):
        return self.tensor is not None

    def islist(self):
        return self.children is not None

    def clone(self):
        if self.istensor():
            return Vector(self.tensor.detach().clone())
        return Vector([c.clone() for c in self.children])

    def get_tensor(self):
        if self.istensor():
            return self.tensor
        raise RuntimeError("Can't get_tensor() when self is list")

    def get_tensorlist(self):
        if self.istensor():
            return [self.tensor]
        out = []
        for c in self.children:
            out.extend(c.get_tensorlist())
        return out

    def __getitem__(self, key):
        return self.children[key] if self.islist() else Vector(self.tensor[key])

    @staticmethod
    def scale(input: "Vector", scale):
        return input * scale

    def __add__(self, o):
        return _binary(self, o, torch.add)

    def __sub__(self, o):
        return _binary(self, o, torch.sub)

    def __mul__(self, o):
        return _binary(self, o, torch.mul)

    def __truediv__(self, o):
        return _binary(self, o, torch.div)

    def __pow__(self, o):
        return _binary(self, o, torch.pow)

    def __iadd__(self, o):
        return _inplace(self, o, lambda a, b: a.add_(b.to(a.device) if torch.is_tensor(b) else b))

    def __isub__(self, o):
        return _inplace(self, o, lambda a, b: a.sub_(b.to(a.device) if torch.is_tensor(b) else b))

    def __imul__(self, o):
        return _inplace(self, o, lambda a, b: a.mul_(b.to(a.device) if torch.is_tensor(b) else b))

    def __itruediv__(self, o):
        return _inplace(self, o, lambda a, b: a.div_(b.to(a.device) if torch.is_tensor(b) else b))

    def __radd__(self, o):
        return self + o

    def __rsub__(self, o):
        return _wrap(o) - self

    def __rmul__(self, o):
        return self * o

    def __rtruediv__(self, o):
        return _wrap(o) / self

    def __neg__(self):
        return self * -1.0

    def __pos__(self):
        return self


def get_shape(inp):
    if isinstance(inp, Vector):
        return tuple(inp.tensor.shape) if inp.istensor() else [get_shape(c) for c in inp.children]
    if isinstance(inp, torch.Tensor):
        return tuple(inp.shape)
    return inp


def _make(shape, fn):
    if isinstance(shape, list):
        return Vector([_make(s, fn) for s in shape])
    return Vector(fn(shape))


def rand(shape, dtype, device=None):
    return _make(shape, lambda s: torch.rand(s, dtype=dtype, device=device))


def zeros(shape, dtype, device=None):
    return _make(shape, lambda s: torch.zeros(s, dtype=dtype, device=device))


def ones(shape, dtype, device=None):
    return _make(shape, lambda s: torch.ones(s, dtype=dtype, device=device))


def empty(shape, dtype, device=None):
    return _make(shape, lambda s: torch.empty(s, dtype=dtype, device=device))


def vdot(a: Vector, b: Vector) -> torch.Tensor:
    if a.istensor():
        return torch.vdot(a.tensor.flatten(), b.tensor.flatten())
    return sum((vdot(x, y) for x, y in zip(a.children, b.children)), torch.tensor(0.0))


def norm(a: Vector) -> torch.Tensor:
    if a.istensor():
        return torch.linalg.vector_norm(a.tensor)
    return sum((norm(c) for c in a.children), torch.tensor(0.0))


def min(a: Vector) -> torch.Tensor:
    """Smallest entry over all leaves (opalg.py:443-450)."""
    if a.istensor():
        return torch.min(a.tensor)
    out = min(a.children[0])
    for c in a.children[1:]:
        out = torch.minimum(out, min(c))
    return out


def max(a: Vector) -> torch.Tensor:
    """Largest entry over all leaves (opalg.py:452-459)."""
    if a.istensor():
        return torch.max(a.tensor)
    out = max(a.children[0])
    for c in a.children[1:]:
        out = torch.maximum(out, max(c))
    return out


class Operator:
    def __init__(self, ishape, oshape):
        self.ishape = ishape
        self.oshape = oshape

    def reinitialize(self, ishape, oshape):
        self.ishape, self.oshape = ishape, oshape

    def _apply(self, input: Vector) -> Vector:
        raise NotImplementedError

    def apply(self, input: Vector) -> Vector:
        if get_shape(input) != self.ishape:
            raise RuntimeError("Incompatible input shape for Operator")
        output = self._apply(input)
        if get_shape(output) != self.oshape:
            raise RuntimeError("Incompatible output shape for Operator")
        return output

    def __call__(self, input: Vector) -> Vector:
        return self.apply(input)

    def __mul__(self, other: "Operator"):
        return CompositeOp(self, other)

    def __add__(self, other: "Operator"):
        return AdditiveOp(self, other)


class Linop(Operator):
    pass


class IdentityOp(Operator):
    def __init__(self, ishape):
        super().__init__(ishape, ishape)

    def _apply(self, input):
        return input


class CompositeOp(Operator):
    def __init__(self, left: Operator, right: Operator):
        if left.ishape != right.oshape:
            raise RuntimeError("Incompatible left and right ishape/oshape")
        self.left, self.right = left, right
        super().__init__(right.ishape, left.oshape)

    def _apply(self, input):
        return self.left(self.right(input))


class AdditiveOp(Operator):
    def __init__(self, left: Operator, right: Operator):
        if left.ishape != right.ishape or left.oshape != right.oshape:
            raise RuntimeError("Incompatible left and right ishape/oshape")
        self.left, self.right = left, right
        super().__init__(right.ishape, left.oshape)

    def _apply(self, input):
        return self.left(input) + self.right(input)


class ScaleOp(Operator):
    def __init__(self, op: Operator, scale):
        self.op, self.scale = op, scale
        super().__init__(op.ishape, op.oshape)

    def _apply(self, input):
        return Vector.scale(self.op(input), self.scale)


class MultiplyOp(Operator):
    def __init__(self, mult: Vector):
        self.mult = mult
        shape = get_shape(mult)
        super().__init__(shape, shape)

    def _apply(self, input):
        return self.mult * input


class IterativeAlg:
    def __init__(self, max_iter):
        self.max_iter = max_iter
        self.iter = 0

    def _update(self):
        raise NotImplementedError

    def _done(self):
        return self.iter >= self.max_iter

    def update(self):
        self._update()
        self.iter += 1

    def done(self):
        return self._done()


class MaxEig(IterativeAlg):
    """Power iteration (opalg.py:647-674; op::PowerIteration, lib/op/opalgs.cpp:29-44):
    x random (torch.rand), per iteration y = A x, lambda = |y|, x = y / lambda."""

    def __init__(self, A: Linop, dtype=torch.float32, max_iter=30, device=None, generator=None):
        self.A = A
        if generator is not None:
            self.x = _make(A.ishape, lambda s: torch.rand(s, dtype=dtype, device=device, generator=generator))
        else:
            self.x = rand(A.ishape, dtype=dtype, device=device)
        self.max_eig = torch.tensor(float("inf"))
        super().__init__(max_iter)

    def _update(self):
        y = self.A(self.x)
        self.max_eig = norm(y)
        self.x = y / Vector(self.max_eig)

    def run(self, callback=None, print_info=False):
        i = 0
        while not self.done():
            self.update()
            if print_info:
                print("MaxEig Iter: ", i, "/", self.max_iter, " MaxEig: ", float(self.max_eig))
            if callback is not None:
                callback(self.x, i)
            i += 1
        return self.max_eig
