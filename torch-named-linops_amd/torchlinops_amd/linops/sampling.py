"""Sampling linop: a tensor read at integer grid locations.

Follows src/torchlinops/linops/sampling.py:17-139: shape
``(batch, input) -> (batch, output)``, one integer index tensor per
sampled axis held as frozen Parameters, indices checked against
``[0, size - 1]`` (ValueError), ``from_mask`` / ``from_stacked_idx``
constructors, ``locs`` stacked on the last axis, splitting along the
output axes.  ``fn`` / ``adj_fn`` are the HIP gather and cell-major
scatter-add of ``functional.index``; the raveled index and its
cell-grouped map are built once per device and reused by every call.
"""

from __future__ import annotations

from typing import Optional

import torch
import torch.nn as nn

from ..functional import _index as FI
from ..nameddim import ELLIPSES, NamedShape as NS
from .namedlinop import NamedLinop

__all__ = ["Sampling"]


def _names(shape) -> tuple:
    """A named-dim tuple, ``("...",)`` when unspecified."""
    return ("...",) if shape is None else tuple(shape)


def _checked_indices(idx, input_size: tuple) -> list:
    """Integer index tensors (slices expanded), each within ``[0, size-1]``
    of its axis; ValueError texts as the reference's constructor."""
    if len(idx) != len(input_size):
        raise ValueError(
            f"Input size {input_size} doesn't match index with length {len(idx)}."
        )
    out = list(FI.ensure_tensor_indexing(idx, input_size))
    for axis, (t, size) in enumerate(zip(out, input_size)):
        lo, hi = (int(v) for v in torch.aminmax(t)) if t.numel() else (0, 0)
        if lo < 0 or hi >= size:
            raise ValueError(
                f"Sampling index must lie within range [0, {size - 1}] but got range "
                f"[{lo}, {hi}] for dim {axis}"
            )
    return out


class Sampling(NamedLinop):
    """``(batch, input) -> (batch, output)``: the input read at one integer
    index tensor per sampled axis."""

    def __init__(self, idx, input_size, output_shape=None, input_shape=None,
                 batch_shape=None):
        self.input_size = tuple(int(s) for s in input_size)
        checked = _checked_indices(idx, self.input_size)
        dims = {"batch_shape": _names(batch_shape), "input_shape": _names(input_shape),
                "output_shape": _names(output_shape)}
        super().__init__(NS(dims["batch_shape"])
                         + NS(dims["input_shape"], dims["output_shape"]))
        for name, names in dims.items():
            self._shape[name] = names
        self.idx = nn.ParameterList([nn.Parameter(t, requires_grad=False) for t in checked])
        self._maps = {}

    @property
    def locs(self):
        """Indices stacked on the last axis (as Interpolate's locs)."""
        return torch.stack(tuple(self.idx), dim=-1)

    @classmethod
    def from_mask(cls, mask, *args, **kwargs):
        return cls(FI.mask2idx(mask.bool()), *args, **kwargs)

    @classmethod
    def from_stacked_idx(cls, idx, *args, dim=-1, **kwargs):
        return cls(FI.canonicalize_idx(idx, dim=-1), *args, **kwargs)

    def index_map(self, device) -> FI.IndexMap:
        """Raveled index and cell-grouped point map on ``device``, cached
        per device and index storage."""
        key = (str(device),) + tuple((i.data_ptr(), i._version) for i in self.idx)
        imap = self._maps.get(key)
        if imap is None:
            imap = FI.index_map(tuple(self.idx), self.input_size, device)
            self._maps = {key: imap}
        return imap

    @staticmethod
    def fn(sampling, x, /):
        return FI.index_with_map(x, sampling.index_map(x.device))

    @staticmethod
    def adj_fn(sampling, x, /):
        return FI.index_adjoint_with_map(x, sampling.index_map(x.device))

    @staticmethod
    def split(sampling, tile):
        """A tile over the output (point) dims selects those points of every
        index tensor; with unnamed output dims the linop cannot be split."""
        out_dims = tuple(sampling._shape.output_shape)
        if out_dims == (ELLIPSES,):
            return sampling
        obatch = tuple(tile.get(d, slice(None)) for d in sampling.oshape)
        return type(sampling)(
            sampling.split_idx(None, obatch, sampling.idx),
            sampling.input_size,
            out_dims,
            sampling._shape.input_shape,
            sampling._shape.batch_shape,
        )

    def split_idx(self, ibatch, obatch, idx):
        """Slice each index tensor by the trailing entries of ``obatch``
        (one per index-tensor axis)."""
        point_axes = idx[0].ndim
        if point_axes == 0:
            return idx
        sel = tuple(obatch)[len(obatch) - point_axes:]
        return [t[sel] for t in idx]

    def register_shape(self, name, shape: tuple):
        self._shape[name] = shape

    def size(self, dim):
        sizes = dict(zip(self._shape.input_shape, self.input_size))
        sizes.update(zip(self._shape.output_shape, self.locs.shape[:-1]))
        return sizes.get(dim)
