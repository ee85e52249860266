"""NamedLinop: a linear operator whose input/output axes carry names.

Same contract as the reference's base class
(src/torchlinops/linops/namedlinop.py:34-569):
  * subclasses provide static ``fn`` / ``adj_fn`` (and optionally
    ``normal_fn``); ``forward`` calls ``fn`` (on ``self.stream`` if set);
  * ``.H`` / ``.N`` build (and cache, per ``config.cache_adjoint_normal``)
    the adjoint and normal operators; ``adjoint()`` swaps fn/adj_fn on a
    shallow copy; ``normal(inner)`` returns ``A^H inner A`` as a Chain, or a
    linop that runs ``normal_fn`` when there is no relevant inner;
  * ``@`` composes, ``+`` / ``-`` / scalar ``*`` build Add / Scalar chains.
"""

from __future__ import annotations

from copy import copy, deepcopy
from functools import partial
from typing import Optional

import torch
import torch.nn as nn

from .. import config
from ..nameddim import NamedShape as NS

__all__ = ["NamedLinop"]


class NamedLinop(nn.Module):
    is_container: bool = False

    def __init__(self, shape: NS, name: Optional[str] = None, stream=None):
        super().__init__()
        self._shape = shape
        self._suffix = ""
        self._name = name
        self.stream = stream
        self.reset_adjoint_and_normal()

    # ------------------------------------------------------------ apply
    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if self.stream is not None and x.is_cuda:
            with torch.cuda.stream(self.stream):
                return self.fn(self, x)
        return self.fn(self, x)

    @staticmethod
    def fn(linop, x, /):
        return x

    @staticmethod
    def adj_fn(linop, x, /):
        return x

    @staticmethod
    def normal_fn(linop, x, /):
        return linop.adj_fn(linop, linop.fn(linop, x))

    @staticmethod
    def split(linop, tile):
        return copy(linop)

    def size(self, dim) -> Optional[int]:
        return None

    @property
    def dims(self) -> set:
        return set(self.ishape).union(set(self.oshape))

    # ------------------------------------------------- adjoint / normal
    @property
    def H(self) -> "NamedLinop":
        if config.cache_adjoint_normal:
            if self._adjoint is None:
                adj = self.adjoint()
                adj._adjoint = [self]
                self._adjoint = [adj]
            return self._adjoint[0]
        return self.adjoint()

    def adjoint(self) -> "NamedLinop":
        adj = copy(self)
        adj._shape = adj._shape.H
        adj.fn, adj.adj_fn = adj.adj_fn, adj.fn
        adj._update_suffix(adjoint=True)
        return adj

    @property
    def N(self) -> "NamedLinop":
        if config.cache_adjoint_normal:
            if self._normal is None:
                self._normal = [self.normal()]
            return self._normal[0]
        return self.normal()

    def normal(self, inner=None) -> "NamedLinop":
        if config.inner_not_relevant(inner):
            normal = copy(self)
            normal._shape = self._shape.N
            table = _NormalFunctions(self)
            normal.fn = table.forward
            normal.adj_fn = table.forward
            normal.normal_fn = table.normal
            normal.adjoint = partial(_normal_adjoint, normal)
            normal._update_suffix(normal=True)
            return normal
        pre = copy(self)
        pre.oshape = inner.ishape
        post = self.adjoint()
        post.ishape = inner.oshape
        return post @ inner @ pre

    def _update_suffix(self, adjoint: bool = False, normal: bool = False):
        if adjoint:
            self._suffix = self._suffix[:-2] if self._suffix.endswith(".H") else self._suffix + ".H"
        elif normal:
            self._suffix += ".N"

    def reset_adjoint_and_normal(self):
        self._adjoint = None
        self._normal = None

    # -------------------------------------------------------- algebra
    def flatten(self) -> list:
        return [self]

    def compose(self, inner: "NamedLinop") -> "NamedLinop":
        from .chain import Chain

        return Chain(*(inner.flatten() + self.flatten()))

    def __matmul__(self, right):
        if isinstance(right, NamedLinop):
            return self.compose(right)
        if isinstance(right, torch.Tensor):
            return self(right)
        return NotImplemented

    def __rmatmul__(self, left):
        if not isinstance(left, NamedLinop):
            raise ValueError(f"__rmatmul__ with non-linop {type(left)} is undefined.")
        return left.compose(self)

    def __add__(self, right):
        from .chain import Add

        return Add(self, right)

    def __radd__(self, left):
        from .chain import Add

        return Add(left, self)

    def __mul__(self, right):
        if isinstance(right, (int, float, complex, torch.Tensor)):
            from .basic import Scalar

            return self.compose(Scalar(weight=right, ioshape=self.ishape))
        return NotImplemented

    def __rmul__(self, left):
        if isinstance(left, (int, float, complex, torch.Tensor)):
            from .basic import Scalar

            return Scalar(weight=left, ioshape=self.oshape).compose(self)
        return NotImplemented

    def __neg__(self):
        return (-1) * self

    def __sub__(self, right):
        from .chain import Add

        return Add(self, -right)

    # --------------------------------------------------------- shapes
    @property
    def name(self):
        return self._name if self._name is not None else type(self).__name__

    @name.setter
    def name(self, v):
        self._name = v

    @property
    def repr_name(self):
        return self.name + self._suffix

    def __repr__(self):
        return f"{self.repr_name}({self.ishape} -> {self.oshape})"

    @property
    def shape(self):
        return self._shape

    @shape.setter
    def shape(self, v):
        self._shape = v

    @property
    def ishape(self):
        return self._shape.ishape

    @ishape.setter
    def ishape(self, v):
        self._shape.ishape = v

    @property
    def oshape(self):
        return self._shape.oshape

    @oshape.setter
    def oshape(self, v):
        self._shape.oshape = v

    def __copy__(self):
        cls = type(self)
        new = cls.__new__(cls)
        new.__dict__ = self.__dict__.copy()
        new._parameters = new._parameters.copy()
        new._modules = new._modules.copy()
        new._buffers = new._buffers.copy()
        new._shape = deepcopy(self._shape)
        new.reset_adjoint_and_normal()
        return new


class _NormalFunctions:
    """fn / adj_fn / normal_fn of A^H A built from A (picklable, no lambdas)."""

    def __init__(self, linop):
        self.linop = linop

    def forward(self, _, x, *args, **kwargs):
        return self.linop.normal_fn(self.linop, x, *args, **kwargs)

    def normal(self, _, x, *args, **kwargs):
        y = self.linop.normal_fn(self.linop, x, *args, **kwargs)
        return self.linop.normal_fn(self.linop, y, *args, **kwargs)


def _normal_adjoint(self):
    adj = copy(self)
    adj._shape = adj._shape.H
    return adj
