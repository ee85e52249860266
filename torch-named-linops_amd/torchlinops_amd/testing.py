"""Contract every linop must satisfy (mirrors src/torchlinops/testing.py:10-123).

Subclass ``BaseNamedLinopTests`` and provide a ``linop_input_output``
fixture returning (A, x, y) with x in A's domain and y in its range.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Literal

import pytest
import torch

from .alg.cg import inner

__all__ = ["BaseNamedLinopTests"]


class BaseNamedLinopTests(ABC):
    equality_check: Literal["exact", "approx"] = "exact"
    isclose_kwargs: dict = {}

    @pytest.fixture
    @abstractmethod
    def linop_input_output(self):
        ...

    def _close(self, a, b):
        if self.equality_check == "exact":
            return bool(a == b)
        return bool(torch.isclose(a, b, **self.isclose_kwargs).all())

    def test_input_mutation(self, linop_input_output):
        A, x, y = linop_input_output
        xc, yc = x.clone(), y.clone()
        A(x)
        assert (xc == x).all()
        A.H(y)
        assert (yc == y).all()
        A.N(x)
        assert (xc == x).all()

    def test_adjoint(self, linop_input_output):
        A, x, y = linop_input_output
        assert self._close(inner(y, A(x)), inner(A.H(y), x))
        assert A.H.shape == A.shape.adjoint()

    def test_normal(self, linop_input_output):
        A, x, _ = linop_input_output
        assert self._close(A.H(A(x)).flatten(), A.N(x).flatten()) or torch.isclose(
            A.H(A(x)), A.N(x), **self.isclose_kwargs
        ).all()

    def test_adjoint_level2(self, linop_input_output):
        A, x, y = linop_input_output
        self.test_adjoint((A.H, y, x))
        self.test_normal((A.H, y, y))

    def test_normal_level2(self, linop_input_output):
        A, x, y = linop_input_output
        self.test_adjoint((A.N, x, x))
        self.test_normal((A.N, x, x))

    def test_split(self, linop_input_output):
        A, x, y = linop_input_output
        for dim in set(A.ishape + A.oshape):
            try:
                assert A.split(A, {dim: slice(0, 2)}) is not None
            except (KeyError, ValueError, NotImplementedError):
                pass

    def test_size(self, linop_input_output):
        A, x, y = linop_input_output
        for dim in A.dims:
            s = A.size(dim)
            assert s is None or isinstance(s, int)

    def test_adj_fn(self, linop_input_output):
        A, x, y = linop_input_output
        assert torch.isclose(A.H(y), A.adj_fn(A, y), **self.isclose_kwargs).all()

    def test_normal_fn(self, linop_input_output):
        A, x, y = linop_input_output
        assert torch.isclose(A.N(x), A.normal_fn(A, x.clone()), **self.isclose_kwargs).all()

    def test_backprop(self, linop_input_output):
        A, x, y = linop_input_output
        x = x.clone().requires_grad_(True)
        out = A(x)
        (out.real.sum() if torch.is_complex(out) else out.sum()).backward()
        assert x.grad.allclose(A.H(torch.ones_like(out)), rtol=1e-3)
