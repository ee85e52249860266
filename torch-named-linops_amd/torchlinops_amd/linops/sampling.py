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


class Sampling(NamedLinop):
    def __init__(self, idx, input_size, output_shape=None, input_shape=None,
                 batch_shape=None):
        dim = len(input_size)
        if len(idx) != dim:
            raise ValueError(
                f"Input size {input_size} doesn't match index with length {len(idx)}."
            )
        self.input_size = tuple(int(s) for s in input_size)
        batch_shape = tuple(batch_shape) if batch_shape is not None else ("...",)
        input_shape = tuple(input_shape) if input_shape is not None else ("...",)
        output_shape = tuple(output_shape) if output_shape is not None else ("...",)
        super().__init__(NS(batch_shape) + NS(input_shape, output_shape))
        self._shape.batch_shape = batch_shape
        self._shape.input_shape = input_shape
        self._shape.output_shape = output_shape
        idx = FI.ensure_tensor_indexing(idx, self.input_size)
        for d, (t, s) in enumerate(zip(idx, self.input_size)):
            if (t < 0).any() or (t >= s).any():
                raise ValueError(
                    f"Sampling index must lie within range [0, {s - 1}] but got range "
                    f"[{t.min().item()}, {t.max().item()}] for dim {d}"
                )
        self.idx = nn.ParameterList([nn.Parameter(i, requires_grad=False) for i in idx])
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
        ibatch = tuple(tile.get(d, slice(None)) for d in sampling.ishape)
        obatch = tuple(tile.get(d, slice(None)) for d in sampling.oshape)
        if tuple(sampling._shape.output_shape) == (ELLIPSES,):
            return sampling
        return type(sampling)(
            sampling.split_idx(ibatch, obatch, sampling.idx),
            sampling.input_size,
            sampling._shape.output_shape,
            sampling._shape.input_shape,
            sampling._shape.batch_shape,
        )

    def split_idx(self, ibatch, obatch, idx):
        n = len(idx[0].shape)
        if n > 0:
            sl = tuple(obatch[-n:])
            return [i[sl] for i in idx]
        return idx

    def register_shape(self, name, shape: tuple):
        self._shape[name] = shape

    def size(self, dim):
        out = list(self._shape.output_shape)
        inp = list(self._shape.input_shape)
        if dim in out:
            return self.locs.shape[out.index(dim)]
        if dim in inp:
            return self.input_size[inp.index(dim)]
        return None
