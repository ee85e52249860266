"""Chain: sequential composition, and Add: sum of linops.

Chain follows src/torchlinops/linops/chain.py:26-232 (``Chain(A, B, C)``
applies A then B then C; adjoint reverses and adjoints each member;
``normal(inner)`` folds ``linop.normal(inner)`` from the last member to the
first, which lets FFT / Pad / Diagonal shortcuts and Toeplitz embedding
compose).  Add follows src/torchlinops/linops/add.py (sum of outputs).
"""

from __future__ import annotations

from copy import copy

import torch
import torch.nn as nn

from .. import config
from ..nameddim import NamedShape as NS, isequal, resolve_wildcards
from .namedlinop import NamedLinop

__all__ = ["Chain", "Add"]


class Chain(NamedLinop):
    is_container = True

    def __init__(self, *linops, name=None):
        if len(linops) == 0:
            raise ValueError("Chain must contain at least one linop.")
        super().__init__(NS(linops[0].ishape, linops[-1].oshape), name=name)
        self.linops = nn.ModuleList(list(linops))
        if config.shape_inference:
            for a, b in zip(self.linops[:-1], self.linops[1:]):
                r = resolve_wildcards(b.ishape, a.oshape)
                if tuple(r) != tuple(b.ishape):
                    b.ishape = r
        cur = self.ishape
        for i, linop in enumerate(self.linops):
            if not isequal(linop.ishape, cur)[0]:
                raise ValueError(
                    f"Mismatched shape: expected {linop.ishape}, got {cur} at input to "
                    f"{linop}. Full stack: {self}, index {i}"
                )
            cur = linop.oshape

    @staticmethod
    def fn(chain, x, /):
        for linop in chain.linops:
            x = linop(x)
        return x

    @staticmethod
    def adj_fn(chain, x, /):
        for linop in reversed(chain.linops):
            x = linop.H(x)
        return x

    @staticmethod
    def normal_fn(chain, x, /):
        return chain.adj_fn(chain, chain.fn(chain, x))

    @staticmethod
    def split(chain, tile):
        out = copy(chain)
        out.linops = nn.ModuleList(
            [type(l).split(l, {d: tile.get(d, slice(None)) for d in l.dims}) for l in chain.linops]
        )
        return out

    def size(self, dim):
        out = None
        for linop in self.linops:
            s = linop.size(dim)
            if s is not None:
                if out is not None and out != s:
                    raise ValueError(f"Conflicting linop sizes found: {out} and {s} for dim {dim}")
                out = s
        return out

    @property
    def dims(self):
        return set().union(*[l.dims for l in self.linops])

    def adjoint(self):
        adj = copy(self)
        adj.linops = nn.ModuleList([l.adjoint() for l in reversed(self.linops)])
        adj._update_suffix(adjoint=True)
        return adj

    def normal(self, inner=None):
        for linop in reversed(self.linops):
            inner = linop.normal(inner)
        return inner

    @property
    def shape(self):
        return NS(self.linops[0].ishape, self.linops[-1].oshape)

    @shape.setter
    def shape(self, v):
        self.ishape = v.ishape
        self.oshape = v.oshape

    @property
    def ishape(self):
        return self.linops[0].ishape

    @ishape.setter
    def ishape(self, v):
        self.linops[0].ishape = v

    @property
    def oshape(self):
        return self.linops[-1].oshape

    @oshape.setter
    def oshape(self, v):
        self.linops[-1].oshape = v

    def flatten(self):
        return list(self.linops)

    def __getitem__(self, idx):
        sel = self.linops[idx]
        return sel if isinstance(sel, NamedLinop) else Chain(*sel, name=self._name)

    def __len__(self):
        return len(self.linops)

    def __repr__(self):
        inner = "\n".join("  " + repr(l).replace("\n", "\n  ") for l in self.linops)
        return f"{self.repr_name}(\n{inner}\n)"


class Add(NamedLinop):
    """Sum of linops with identical input and output shapes."""

    is_container = True

    def __init__(self, *linops):
        if len(linops) == 0:
            raise ValueError("Add needs at least one linop")
        super().__init__(NS(linops[0].ishape, linops[0].oshape))
        for l in linops[1:]:
            if not (isequal(l.ishape, self.ishape)[0] and isequal(l.oshape, self.oshape)[0]):
                raise ValueError(f"Add: incompatible shapes {l.shape} vs {self.shape}")
        self.linops = nn.ModuleList(list(linops))

    @staticmethod
    def fn(add, x, /):
        y = add.linops[0](x)
        for l in add.linops[1:]:
            y = y + l(x)
        return y

    @staticmethod
    def adj_fn(add, x, /):
        y = add.linops[0].H(x)
        for l in add.linops[1:]:
            y = y + l.H(x)
        return y

    @staticmethod
    def normal_fn(add, x, /):
        return add.adj_fn(add, add.fn(add, x))

    def adjoint(self):
        adj = copy(self)
        adj._shape = adj._shape.H
        adj.linops = nn.ModuleList([l.adjoint() for l in self.linops])
        return adj

    def normal(self, inner=None):
        if config.inner_not_relevant(inner):
            return super().normal(None)
        return super().normal(inner)

    def size(self, dim):
        for l in self.linops:
            s = l.size(dim)
            if s is not None:
                return s
        return None
