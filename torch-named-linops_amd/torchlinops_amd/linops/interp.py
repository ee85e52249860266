"""Interpolate, ArrayToBlocks and BlocksToArray linops over the HIP path.

Interpolate follows src/torchlinops/linops/interp.py:16-137 (shape
``(batch, grid) -> (batch, locs_batch)``, locs held as a frozen Parameter,
default Kaiser-Bessel beta 1.0); ArrayToBlocks / BlocksToArray follow
src/torchlinops/linops/array_to_blocks.py:16-205 (adjoints of each other).
"""

from __future__ import annotations

from copy import copy
from typing import Optional

import torch
import torch.nn as nn

from .. import functional as F
from ..nameddim import ELLIPSES, NamedShape as NS
from .namedlinop import NamedLinop

__all__ = ["Interpolate", "ArrayToBlocks", "BlocksToArray"]


class Interpolate(NamedLinop):
    def __init__(self, locs, grid_size, batch_shape=None, locs_batch_shape=None,
                 grid_shape=None, width: float = 4.0, kernel: str = "kaiser_bessel",
                 norm: int = 1, pad_mode: str = "circular",
                 kernel_params: Optional[dict] = None):
        if locs_batch_shape is not None and len(locs_batch_shape) > locs.ndim - 1:
            raise ValueError(
                f"locs_batch_shape has length longer than batch dim of locs. "
                f"locs_batch_shape: {locs_batch_shape}, locs: {tuple(locs.shape)}"
            )
        batch_shape = tuple(batch_shape) if batch_shape is not None else ("...",)
        locs_batch_shape = tuple(locs_batch_shape) if locs_batch_shape is not None else ("...",)
        grid_shape = tuple(grid_shape) if grid_shape is not None else ("...",)
        super().__init__(NS(batch_shape) + NS(grid_shape, locs_batch_shape))
        self._shape.batch_shape = batch_shape
        self._shape.locs_batch_shape = locs_batch_shape
        self._shape.grid_shape = grid_shape
        self.locs = nn.Parameter(locs, requires_grad=False)
        self.grid_size = tuple(grid_size)
        kp = {"beta": 1.0}
        if kernel_params:
            kp.update(kernel_params)
        self.interp_params = {"width": width, "kernel": kernel, "norm": norm,
                              "pad_mode": pad_mode, "kernel_params": kp}

    @staticmethod
    def fn(interp, x, /):
        return F.interpolate(x, interp.locs, **interp.interp_params)

    @staticmethod
    def adj_fn(interp, x, /):
        return F.interpolate_adjoint(x, interp.locs, interp.grid_size, **interp.interp_params)

    @staticmethod
    def split(interp, tile):
        obatch = tuple(tile.get(d, slice(None)) for d in interp.oshape)
        return type(interp)(interp.split_locs(obatch, interp.locs), interp.grid_size,
                            interp._shape.batch_shape, interp._shape.locs_batch_shape,
                            interp._shape.grid_shape, **interp.interp_params)

    def split_locs(self, obatch, /, locs):
        lbs = self._shape.locs_batch_shape
        if tuple(lbs) == (ELLIPSES,):
            return locs
        n = len(lbs)
        return locs[tuple(obatch[-n:]) + (slice(None),)]

    def size(self, dim):
        lbs = list(self._shape.locs_batch_shape)
        gs = list(self._shape.grid_shape)
        if dim in lbs:
            return self.locs.shape[lbs.index(dim)]
        if dim in gs:
            return self.grid_size[gs.index(dim)]
        return None


class ArrayToBlocks(NamedLinop):
    def __init__(self, grid_size, block_size, stride, mask=None, batch_shape=None,
                 array_shape=None, blocks_shape=None):
        self.grid_size = tuple(grid_size)
        self.ndim = len(self.grid_size)
        self.block_size = tuple(block_size)
        self.stride = tuple(stride)
        self.batch_shape = tuple(batch_shape) if batch_shape is not None else ("...",)
        self.array_shape = tuple(array_shape) if array_shape is not None else ("...",)
        self.blocks_shape = tuple(blocks_shape) if blocks_shape is not None else ("...",)
        super().__init__(NS(self.batch_shape) + NS(self.array_shape, self.blocks_shape))
        self.mask = nn.Parameter(mask, requires_grad=False) if mask is not None else None

    @staticmethod
    def fn(a, x, /):
        return F.array_to_blocks(x, a.block_size, a.stride, a.mask)

    @staticmethod
    def adj_fn(a, x, /):
        return F.blocks_to_array(x, a.grid_size, a.block_size, a.stride, a.mask)

    @staticmethod
    def normal_fn(a, x, /):
        return a.adj_fn(a, a.fn(a, x))

    @staticmethod
    def split(linop, tile):
        return copy(linop)

    def adjoint(self):
        return BlocksToArray(self.grid_size, self.block_size, self.stride, self.mask,
                             self.batch_shape, self.blocks_shape, self.array_shape)

    def size(self, dim):
        if dim in self.ishape[-self.ndim:]:
            return self.grid_size[list(self.ishape).index(dim) - len(self.ishape)]
        return None


class BlocksToArray(NamedLinop):
    def __init__(self, grid_size, block_size, stride, mask=None, batch_shape=None,
                 blocks_shape=None, array_shape=None):
        self.grid_size = tuple(grid_size)
        self.ndim = len(self.grid_size)
        self.block_size = tuple(block_size)
        self.stride = tuple(stride)
        self.batch_shape = tuple(batch_shape) if batch_shape is not None else ("...",)
        self.blocks_shape = tuple(blocks_shape) if blocks_shape is not None else ("...",)
        self.array_shape = tuple(array_shape) if array_shape is not None else ("...",)
        super().__init__(NS(self.batch_shape) + NS(self.blocks_shape, self.array_shape))
        self.mask = nn.Parameter(mask, requires_grad=False) if mask is not None else None

    @staticmethod
    def fn(b, x, /):
        return F.blocks_to_array(x, b.grid_size, b.block_size, b.stride, b.mask)

    @staticmethod
    def adj_fn(b, x, /):
        if tuple(x.shape[-b.ndim:]) != b.grid_size:
            raise RuntimeError(
                f"BlocksToArray expected input with full size {b.grid_size} but got {tuple(x.shape)}"
            )
        return F.array_to_blocks(x, b.block_size, b.stride, b.mask)

    @staticmethod
    def split(linop, tile):
        return copy(linop)

    def adjoint(self):
        return ArrayToBlocks(self.grid_size, self.block_size, self.stride, self.mask,
                             self.batch_shape, self.array_shape, self.blocks_shape)

    def size(self, dim):
        if dim in self.oshape[-self.ndim:]:
            return self.grid_size[list(self.oshape).index(dim) - len(self.oshape)]
        return None
