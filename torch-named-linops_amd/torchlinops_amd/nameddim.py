"""Named dimensions and named shapes (the subset the gridding path needs).

Semantics follow the reference's ``torchlinops.nameddim``:
  * ``NamedDimension``  src/torchlinops/nameddim/_nameddim.py:93-191
    (name + integer index, string equality, ``next_unused``, ``infer``)
  * ``Dim``             _nameddim.py:12-90 (split "NxNyNz" at capitals)
  * ``NamedShape``      _namedshape.py:13-144 and
    ``NamedDimCollection`` _nameddimcollection.py:13-325: input/output
    shapes plus auxiliary shapes that share dimension objects, so renaming
    a dim through one shape renames it in every shape of the linop;
    ``.H`` swaps input and output, ``.N`` re-indexes output dims that are
    not shared with the input.
  * wildcard matching   _matching.py:60-200 ("..." = any run, "()" = one)
"""

from __future__ import annotations

import re
from copy import deepcopy
from typing import Iterable, Optional, Sequence

__all__ = [
    "NamedDimension",
    "ND",
    "Dim",
    "ELLIPSES",
    "ANY",
    "NamedShape",
    "NS",
    "Shape",
    "isequal",
    "is_any",
    "get_nd_shape",
    "resolve_wildcards",
]

ELLIPSES = "..."
ANY = "()"

Shape = Sequence


def Dim(s: Optional[str] = None) -> tuple:
    """Split a compact dim string: a new dim starts at each capital letter,
    at "(" and after "()" / "(k)" tokens.

    >>> Dim("NxNyNz")
    ('Nx', 'Ny', 'Nz')
    >>> Dim("()A(2)B")
    ('()', 'A', '(2)', 'B')
    """
    if not s:
        return tuple()
    tokens = re.findall(r"\(\d*\)|[A-Z][^A-Z(]*|[^A-Z(]+", s)
    if "".join(tokens) != s:
        raise ValueError(f"Cannot parse dim string {s!r}")
    if any(t[0].isdigit() for t in tokens):
        raise ValueError(f"Dim cannot start with a digit in dim string {s}")
    return tuple(tokens)


class NamedDimension:
    """A dimension name with an integer index (``A``, ``A1``, ...).

    Equality and hashing use the printed form, so ``ND("A", 1) == "A1"``.
    """

    __slots__ = ("name", "i")

    def __init__(self, name: str, i: int = 0):
        object.__setattr__(self, "name", name)
        object.__setattr__(self, "i", i)

    def __setattr__(self, k, v):
        raise AttributeError("NamedDimension is immutable")

    @classmethod
    def infer(cls, dim):
        if isinstance(dim, cls):
            return dim
        if isinstance(dim, (tuple, list)):
            return type(dim)(cls.infer(d) for d in dim)
        dim = str(dim)
        if dim == ELLIPSES:
            return cls(ELLIPSES)
        m = re.fullmatch(r"\((\d+)\)", dim)
        if m:
            return cls(ANY, int(m.group(1)))
        if len(dim) == 2 and dim[1].isdigit():
            return cls(dim[0], int(dim[1]))
        return cls(dim)

    def __repr__(self):
        if self.name == ANY and self.i > 0:
            return f"({self.i})"
        return self.name + ("" if self.i == 0 else str(self.i))

    def __add__(self, k: int):
        if self.name == ELLIPSES:
            return self
        return type(self)(self.name, self.i + k)

    def next_unused(self, avoid: Optional[Iterable] = None):
        if avoid is None:
            return self + 1
        if self.name == ELLIPSES:
            return self
        cur = self
        avoid = list(avoid)
        while cur in avoid:
            cur = cur + 1
        return cur

    def __eq__(self, other):
        return repr(self) == (repr(other) if isinstance(other, NamedDimension) else other)

    def __hash__(self):
        return hash(repr(self))

    def __copy__(self):
        return self

    def __deepcopy__(self, memo):
        return self


ND = NamedDimension


def is_any(dim) -> bool:
    s = repr(dim) if isinstance(dim, NamedDimension) else str(dim)
    return s == ANY or re.fullmatch(r"\(\d+\)", s) is not None


def _is_ell(d) -> bool:
    return d == ELLIPSES


def _tok_eq(a, b) -> bool:
    return a == b or is_any(a) or is_any(b)


def isequal(shape1: Sequence, shape2: Sequence) -> tuple:
    """Whether two dim sequences can denote the same axes: "..." matches any
    run (possibly empty) of the other sequence, "()" any single dim.
    Returns (bool, None) -- the boolean is what callers use."""
    s1, s2 = list(shape1), list(shape2)
    n1, n2 = len(s1), len(s2)
    # dp[i][j]: s1[:i] and s2[:j] can match
    dp = [[False] * (n2 + 1) for _ in range(n1 + 1)]
    dp[0][0] = True
    for i in range(n1 + 1):
        for j in range(n2 + 1):
            if not dp[i][j]:
                continue
            if i < n1 and _is_ell(s1[i]):
                dp[i + 1][j] = True
                if j < n2:
                    dp[i][j + 1] = True  # ellipsis in s1 absorbs s2[j]
            if j < n2 and _is_ell(s2[j]):
                dp[i][j + 1] = True
                if i < n1:
                    dp[i + 1][j] = True
            if i < n1 and j < n2 and _tok_eq(s1[i], s2[j]):
                dp[i + 1][j + 1] = True
    return dp[n1][n2], None


def resolve_wildcards(shape: Sequence, target: Sequence) -> tuple:
    """Replace a single "..." in ``shape`` by the dims of ``target`` it
    covers (prefix/suffix aligned); unchanged when not resolvable."""
    shape = tuple(shape)
    if ELLIPSES not in shape or ELLIPSES in tuple(target):
        return shape
    i = shape.index(ELLIPSES)
    pre, post = shape[:i], shape[i + 1:]
    if ELLIPSES in post or len(pre) + len(post) > len(target):
        return shape
    mid = tuple(target)[len(pre): len(target) - len(post)]
    return pre + mid + post


def get_nd_shape(ndim: int, kspace: bool = False) -> tuple:
    names = ("Kx", "Ky", "Kz") if kspace else ("Nx", "Ny", "Nz")
    if ndim not in (1, 2, 3):
        raise ValueError(f"dim should be 1, 2, or 3 but got {ndim}")
    return names[:ndim]


def _align(old: tuple, new: tuple) -> list:
    """For each entry of ``old``, the slice of ``new`` it corresponds to.
    A concrete old dim takes exactly one new entry; an old "..." takes a
    (possibly empty) run; a new "..." may stand for several old dims."""
    n_o, n_n = len(old), len(new)
    out = [None] * n_o
    i = 0
    while i < n_o and i < n_n and not _is_ell(old[i]) and not _is_ell(new[i]):
        out[i] = [i]
        i += 1
    k = 0
    while (k < n_o - i and k < n_n - i and not _is_ell(old[n_o - 1 - k])
           and not _is_ell(new[n_n - 1 - k])):
        out[n_o - 1 - k] = [n_n - 1 - k]
        k += 1
    mid_o = list(range(i, n_o - k))
    mid_n = list(range(i, n_n - k))
    if not mid_o and not mid_n:
        return out
    if len(mid_o) == 1 and _is_ell(old[mid_o[0]]):
        out[mid_o[0]] = mid_n
        return out
    if len(mid_n) == 1 and _is_ell(new[mid_n[0]]):
        for j in mid_o:
            out[j] = [mid_n[0]]
        return out
    if len(mid_o) == len(mid_n):
        for a, b in zip(mid_o, mid_n):
            out[a] = [b]
        return out
    raise ValueError(f"Updated shape {new} not compatible with current: {old}")


class NamedShape:
    """Input/output dims of a linop plus auxiliary named shapes that share
    dim objects with them (renaming a dim renames it everywhere)."""

    def __init__(self, ishape=None, oshape=None, **other_shapes):
        if isinstance(ishape, NamedShape):
            src = ishape
            object.__setattr__(self, "_dims", list(src._dims))
            object.__setattr__(self, "_idx", dict(src._idx))
            return
        object.__setattr__(self, "_dims", [])
        object.__setattr__(self, "_idx", {})
        if oshape is None:
            oshape = ishape if ishape is not None else (ELLIPSES,)
        if ishape is None:
            ishape = (ELLIPSES,)
        self._add("ishape", ishape)
        self._add("oshape", oshape)
        for k, v in other_shapes.items():
            self._add(k, v)

    # --------------------------------------------------------- storage
    def _slot(self, d) -> int:
        d = ND.infer(d)
        for k, x in enumerate(self._dims):
            if x == d:
                return k
        self._dims.append(d)
        return len(self._dims) - 1

    def _add(self, name, data):
        if name in self._idx:
            raise ValueError(f"{name} already in shape")
        if isinstance(data, (tuple, list)):
            data = _coalesce(tuple(ND.infer(d) for d in data))
            self._idx[name] = tuple(self._slot(d) for d in data)
        elif isinstance(data, dict):
            self._idx[name] = {self._slot(k): v for k, v in data.items()}
        else:
            self._idx[name] = self._slot(data)

    def _get(self, name):
        v = self._idx[name]
        if isinstance(v, tuple):
            return tuple(self._dims[k] for k in v)
        if isinstance(v, dict):
            return {self._dims[k]: x for k, x in v.items()}
        return self._dims[v]

    def _set(self, name, new):
        if name not in self._idx:
            self._add(name, new)
            return
        cur = self._idx[name]
        if not isinstance(cur, tuple):
            if isinstance(cur, dict):
                raise ValueError("cannot rename a mapping shape")
            self._dims[cur] = ND.infer(new)
            return
        new = _coalesce(tuple(ND.infer(d) for d in new))
        old = self._get(name)
        spans = _align(old, new)
        renames = {}
        slots = []
        for k_old, (od, span) in enumerate(zip(old, spans)):
            slot = cur[k_old]
            if _is_ell(od):
                slots.extend(self._slot(new[j]) for j in span)
                continue
            nd = new[span[0]]
            if _is_ell(nd) or is_any(nd):
                # concrete dim replaced by a wildcard: only in this shape
                slots.append(self._slot(nd))
                continue
            if is_any(od):
                slots.append(self._slot(nd))
                continue
            prev = renames.get(slot)
            if prev is not None and prev != nd:
                raise ValueError(
                    f"Detected inconsistent mapping: dim {od} updated to both {prev} and {nd}"
                )
            renames[slot] = nd
            slots.append(slot)
        for slot, nd in renames.items():
            self._dims[slot] = nd
        self._idx[name] = tuple(_coalesce_idx(slots, self._dims))

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        idx = object.__getattribute__(self, "_idx")
        if name in idx:
            return self._get(name)
        raise AttributeError(f"{name} not in shape {list(idx)}")

    def __setattr__(self, name, value):
        if name.startswith("_"):
            object.__setattr__(self, name, value)
        else:
            self._set(name, value)

    def __getitem__(self, name):
        if name not in self._idx:
            raise KeyError(f"{name} not in shape {list(self._idx)}")
        return self._get(name)

    def __setitem__(self, name, value):
        self._set(name, value)

    def __contains__(self, name) -> bool:
        return name in self._idx

    @property
    def shapes(self) -> dict:
        return {k: self._get(k) for k in self._idx}

    @property
    def other_shapes(self) -> dict:
        return {k: v for k, v in self.shapes.items() if k not in ("ishape", "oshape")}

    def update(self, mapping: dict):
        for k, v in mapping.items():
            self._set(k, v)

    # ------------------------------------------------------------ algebra
    def adjoint(self) -> "NamedShape":
        new = deepcopy(self)
        i, o = new._idx["ishape"], new._idx["oshape"]
        new._idx["ishape"], new._idx["oshape"] = o, i
        return new

    def normal(self) -> "NamedShape":
        new = deepcopy(self)
        ishape = new.ishape
        oshape = []
        for d in ishape:
            oshape.append(d if d in self.oshape else d.next_unused(ishape))
        new._idx["oshape"] = tuple(new._slot(d) for d in oshape)
        return new

    @property
    def H(self) -> "NamedShape":
        return self.adjoint()

    @property
    def N(self) -> "NamedShape":
        return self.normal()

    def __add__(self, right: "NamedShape") -> "NamedShape":
        new = NamedShape(tuple(self.ishape) + tuple(right.ishape),
                         tuple(self.oshape) + tuple(right.oshape))
        for k, v in {**self.other_shapes, **right.other_shapes}.items():
            new._set(k, v)
        return new

    def __radd__(self, left):
        if left is None:
            return self
        return left.__add__(self)

    def __eq__(self, other):
        return (isequal(self.ishape, other.ishape)[0]
                and isequal(self.oshape, other.oshape)[0])

    def __repr__(self):
        return f"{self.ishape} -> {self.oshape}"

    def __deepcopy__(self, memo):
        new = NamedShape.__new__(NamedShape)
        object.__setattr__(new, "_dims", list(self._dims))
        object.__setattr__(
            new, "_idx",
            {k: (dict(v) if isinstance(v, dict) else v) for k, v in self._idx.items()},
        )
        return new


NS = NamedShape


def _coalesce(dims: tuple) -> tuple:
    out = []
    for d in dims:
        if _is_ell(d) and out and _is_ell(out[-1]):
            continue
        out.append(d)
    return tuple(out)


def _coalesce_idx(slots, dims):
    out = []
    for s in slots:
        if _is_ell(dims[s]) and out and _is_ell(dims[out[-1]]):
            continue
        out.append(s)
    return out
