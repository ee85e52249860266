"""Polynomial preconditioners P(T) = sum_k c_k T^k.

Mirrors src/torchlinops/alg/poly.py:27-244 (after sidward/ppcs):
``polynomial_preconditioner(T, degree, norm, lower_eig, upper_eig)``
returns the linop ``c_0 I + (c_1 I + (...) T) T`` and the three
coefficient rules are the same:

  * ``l_2``    minimises  int_a^b (1 - x p(x))^2 dx  (Johnson, Micchelli
               and Paul, DOI 10.1137/0720025); the normal equations have the
               closed-form moments (b^(j+k+3) - a^(j+k+3)) / (j+k+3);
  * ``l_inf``  minimises  sup_[a,b] |1 - x p(x)|:  1 - x p(x) =
               T_{d+1}((b + a - 2x)/(b - a)) / T_{d+1}((b + a)/(b - a))
               (Chebyshev; Shewchuk, eq. 50);
  * ``ifista`` c_k = binom(d + 1, k + 1) (-1)^k  (DOI 10.1137/140970537).

The reference derives the first two symbolically (sympy and the Chebyshev
package); here they are evaluated with numpy's polynomial classes, with the
same positivity / injectivity checks at the stationary points and ends of
the interval.  ``apply_polynomial`` evaluates P(T) x by Horner's rule
directly on tensors (d applications of T, no linop graph), for use as the
``M`` of ``DeviceCG``.
"""

from __future__ import annotations

from typing import Callable, Literal

import numpy as np
import torch
from numpy.polynomial import Chebyshev, Polynomial
from scipy.special import binom

from ..linops import Identity, NamedLinop

__all__ = ["polynomial_preconditioner", "l_2_opt", "l_inf_opt", "ifista_coeffs",
           "apply_polynomial"]


def _extreme_values(poly: Polynomial, lower: float, upper: float) -> np.ndarray:
    """poly at its real stationary points inside [lower, upper] and the ends."""
    pts = [lower, upper]
    if poly.degree() > 1:
        for r in poly.deriv().roots():
            if abs(r.imag) < 1e-12 and lower <= r.real <= upper:
                pts.append(r.real)
    return poly(np.array(pts))


def l_2_opt(degree: int, lower: float = 0.0, upper: float = 1.0, verbose: bool = False):
    """Coefficients (lowest degree first, float32) of the degree-``degree``
    p minimising int_lower^upper (1 - x p(x))^2 dx, and p itself."""
    assert degree >= 0
    j = np.arange(degree + 1)
    e = j[:, None] + j[None, :] + 3
    mat = (upper ** e - lower ** e) / e
    vec = (upper ** (j + 2) - lower ** (j + 2)) / (j + 2)
    res = np.asarray(np.linalg.pinv(mat) @ vec, dtype=np.float32)
    poly = Polynomial(res.astype(np.float64))
    if verbose:
        print(f"L-2 optimized polynomial, degree {degree}, spectrum [{lower}, {upper}]: {poly}")
    if degree > 0:
        assert _extreme_values(poly, lower, upper).min() > 1e-8, "Polynomial is not positive."
    return res, poly


def l_inf_opt(degree: int, lower: float = 0.0, upper: float = 1.0, verbose: bool = False):
    """Coefficients (lowest degree first, float32) of the Chebyshev p with
    1 - x p(x) = T_{d+1}((upper + lower - 2x)/(upper - lower)) / T_{d+1}(at x = 0)."""
    assert degree >= 0
    t = Chebyshev.basis(degree + 1).convert(kind=Polynomial)
    arg = Polynomial([(upper + lower) / (upper - lower), -2.0 / (upper - lower)])
    P = t(arg)
    P = P / P(0.0)
    q = Polynomial([1.0]) - P  # zero constant term: divisible by x
    poly = Polynomial(q.coef[1:] if q.coef.size > 1 else [0.0])
    if verbose:
        print(f"L-infinity optimized polynomial, degree {degree}, "
              f"spectrum [{lower}, {upper}]: {poly}")
    if degree > 0:
        assert np.abs(_extreme_values(poly, lower, upper)).min() > 1e-8, \
            "Polynomial not injective."
    return np.asarray(poly.coef, dtype=np.float32), poly


def ifista_coeffs(degree: int) -> np.ndarray:
    """c_k = binom(degree + 1, k + 1) (-1)^k, k = 0..degree."""
    return np.array([binom(degree + 1, k + 1) * (-1) ** k for k in range(degree + 1)])


def _coeffs(degree, norm, lower_eig, upper_eig):
    if norm == "l_2":
        return l_2_opt(degree, lower_eig, upper_eig)[0]
    if norm == "l_inf":
        return l_inf_opt(degree, lower_eig, upper_eig)[0]
    if norm == "ifista":
        return ifista_coeffs(degree)
    raise ValueError(f"Unknown polynomial preconditioning norm option: {norm}")


def polynomial_preconditioner(T: NamedLinop, degree: int,
                              norm: Literal["l_2", "l_inf", "ifista"] = "l_2",
                              lower_eig: float = 0.0, upper_eig: float = 1.0) -> NamedLinop:
    """The linop P(T) = sum_k c_k T^k (Identity for ``degree < 0``)."""
    Id = Identity()
    if degree < 0:
        return Id
    c = _coeffs(degree, norm, lower_eig, upper_eig)

    def horner(c):
        if c.size == 1:
            return float(c[0]) * Id
        return float(c[0]) * Id + horner(c[1:]) @ T @ Id

    return horner(np.asarray(c))


def apply_polynomial(T: Callable, coeffs, x: torch.Tensor) -> torch.Tensor:
    """P(T) x = sum_k c_k T^k x by Horner's rule: y = c_d x, then
    y = c_k x + T(y) for k = d-1 .. 0."""
    c = [float(v) for v in np.asarray(coeffs).ravel()]
    y = c[-1] * x
    for ck in reversed(c[:-1]):
        y = T(y).add_(x, alpha=ck)
    return y
