"""Elementary linops on the NUFFT chain: Identity, Diagonal, Scalar, Pad
(centred zero pad / crop), FFT (centred, orthonormal) and Dense (einsum).

Reference counterparts (src/torchlinops/linops/): identity.py:16-51,
diagonal.py:15-182, scalar.py:14-42, pad_last.py:13-199, fft.py:13-131,
dense.py:17-414.  These run as torch ops (rocFFT for the FFT).
"""

from __future__ import annotations

from copy import copy
from typing import Optional

import torch
import torch.fft as tfft
import torch.nn as nn
import torch.nn.functional as TF
from einops import einsum

from .. import config
from ..nameddim import ND, NamedShape as NS, get_nd_shape, is_any
from .namedlinop import NamedLinop

__all__ = ["Identity", "Diagonal", "Scalar", "Pad", "FFT", "Dense",
           "pad_to_size", "crop_slice_from_pad"]


class Identity(NamedLinop):
    def __init__(self, ishape=("...",), oshape=None):
        super().__init__(NS(ishape, oshape))

    def adjoint(self):
        return self

    def normal(self, inner=None):
        return self if inner is None else inner

    @staticmethod
    def fn(linop, x, /):
        return x

    @staticmethod
    def adj_fn(linop, x, /):
        return x

    @staticmethod
    def normal_fn(linop, x, /):
        return x

    @staticmethod
    def split(linop, tile):
        return linop


class Diagonal(NamedLinop):
    """x -> w * x (adjoint conj(w), normal |w|^2); ishape == oshape."""

    def __init__(self, weight: torch.Tensor, ioshape=None, broadcast_dims=None):
        if ioshape is not None and weight.ndim > len(ioshape):
            raise ValueError(
                f"All dimensions must be named or broadcastable, but got weight shape "
                f"{tuple(weight.shape)} and ioshape {ioshape}"
            )
        super().__init__(NS(ioshape))
        self.weight = nn.Parameter(weight, requires_grad=False)
        bd = list(broadcast_dims) if broadcast_dims is not None else []
        bd += [d for d in self.ishape if is_any(d)]
        self._shape.broadcast_dims = bd

    @property
    def broadcast_dims(self):
        return self._shape.broadcast_dims

    @NamedLinop.ishape.setter
    def ishape(self, v):
        self._shape.ishape = v
        self._shape.oshape = v

    @NamedLinop.oshape.setter
    def oshape(self, v):
        self._shape.oshape = v
        self._shape.ishape = v

    @staticmethod
    def fn(diag, x, /):
        return x * diag.weight

    @staticmethod
    def adj_fn(diag, x, /):
        return x * torch.conj(diag.weight)

    @staticmethod
    def normal_fn(diag, x, /):
        return x * torch.abs(diag.weight) ** 2

    def adjoint(self):
        adj = copy(self)
        adj.weight = nn.Parameter(self.weight.conj(), requires_grad=False)
        return adj

    def normal(self, inner=None):
        if inner is None:
            n = copy(self)
            n.weight = nn.Parameter(torch.abs(self.weight) ** 2, requires_grad=False)
            return n
        return super().normal(inner)

    def size(self, dim):
        if dim in self.ishape:
            nb = len(self.ishape) - self.weight.ndim
            k = list(self.ishape).index(dim)
            if k < nb or dim in self.broadcast_dims:
                return None
            return self.weight.shape[k - nb]
        return None


class Scalar(Diagonal):
    def __init__(self, weight, ioshape=None):
        if not isinstance(weight, torch.Tensor):
            weight = torch.tensor(weight)
        super().__init__(weight, ioshape=ioshape if ioshape is not None else ("...",))

    def size(self, dim):
        return None


def pad_to_size(grid_size, padded_size) -> list:
    """Centred padding list for torch.nn.functional.pad (last axis first):
    odd total pad goes right for an even image and left for an odd image
    (pad_last.py:139-185)."""
    if len(grid_size) != len(padded_size):
        raise ValueError(f"Dimension mismatch: cannot pad from size {grid_size} to size {padded_size}.")
    pads = []
    for n, p in zip(grid_size, padded_size):
        tp = p - n
        if tp % 2:
            lr = (tp // 2 + 1, tp // 2) if n % 2 else (tp // 2, tp // 2 + 1)
        else:
            lr = (tp // 2, tp // 2)
        pads.append(list(lr))
    pads.reverse()
    return sum(pads, start=[])


def crop_slice_from_pad(pad) -> list:
    out = []
    for i in range(len(pad) // 2):
        lo, hi = pad[2 * i], pad[2 * i + 1]
        out.append(slice(lo, -hi if hi else None))
    out.reverse()
    return out


class Pad(NamedLinop):
    """Centred zero-pad of the last D axes; adjoint crops."""

    def __init__(self, pad_im_size, im_size, in_shape=None, out_shape=None, batch_shape=None):
        if len(pad_im_size) != len(im_size):
            raise ValueError(
                f"Padded and unpadded dims should be the same length. padded: {pad_im_size} unpadded: {im_size}"
            )
        self.in_im_shape = ND.infer(tuple(in_shape) if in_shape is not None else get_nd_shape(len(im_size)))
        self.out_im_shape = (tuple(d.next_unused(self.in_im_shape) for d in self.in_im_shape)
                             if out_shape is None else ND.infer(tuple(out_shape)))
        batch_shape = tuple(batch_shape) if batch_shape is not None else ("...",)
        super().__init__(NS(batch_shape) + NS(self.in_im_shape, self.out_im_shape))
        self.D = len(im_size)
        self.im_size = tuple(im_size)
        self.pad_im_size = tuple(pad_im_size)
        self.in_im_size = tuple(im_size)
        self.out_im_size = tuple(pad_im_size)
        self.pad = pad_to_size(self.im_size, self.pad_im_size)
        self.crop_slice = crop_slice_from_pad(self.pad)

    @staticmethod
    def fn(p, x, /):
        if tuple(x.shape[-p.D:]) != p.im_size:
            raise ValueError(f"Mismatched shapes: expected {p.im_size} but got {tuple(x.shape[-p.D:])}")
        return TF.pad(x, p.pad)

    @staticmethod
    def adj_fn(p, y, /):
        if tuple(y.shape[-p.D:]) != p.pad_im_size:
            raise ValueError(f"Mismatched shapes: expected {p.pad_im_size} but got {tuple(y.shape[-p.D:])}")
        return y[(Ellipsis, *p.crop_slice)]

    @staticmethod
    def normal_fn(p, x, /):
        return x

    def adjoint(self):
        adj = super().adjoint()
        adj.name = "Crop" if self.name == "Pad" else ("Pad" if self.name == "Crop" else adj.name)
        adj.in_im_shape, adj.out_im_shape = self.out_im_shape, self.in_im_shape
        adj.in_im_size, adj.out_im_size = self.out_im_size, self.in_im_size
        return adj

    def size(self, dim):
        if dim in self.ishape[-self.D:]:
            return self.in_im_size[list(self.ishape[-self.D:]).index(dim)]
        if dim in self.oshape[-self.D:]:
            return self.out_im_size[list(self.oshape[-self.D:]).index(dim)]
        return None


class FFT(NamedLinop):
    """n-D FFT over the last ``ndim`` axes; ``centered`` wraps it in
    ifftshift / fftshift (fft.py:84-101); ortho norm makes it unitary."""

    def __init__(self, ndim, batch_shape=None, grid_shapes=None, norm="ortho", centered=False):
        self.ndim = ndim
        self.dim = tuple(range(-ndim, 0))
        if grid_shapes is None:
            grid_shapes = (get_nd_shape(ndim), get_nd_shape(ndim, kspace=True))
        elif len(grid_shapes) != 2:
            raise ValueError(f"grid_shapes should consist of two shape tuples but got {grid_shapes}")
        if len(grid_shapes[0]) != len(grid_shapes[1]):
            raise ValueError("Input and output shapes of FFT must have same length")
        batch_shape = tuple(batch_shape) if batch_shape is not None else ("...",)
        super().__init__(NS(batch_shape) + NS(tuple(grid_shapes[0]), tuple(grid_shapes[1])))
        self._shape.batch_shape = batch_shape
        self._shape.input_grid_shape = tuple(grid_shapes[0])
        self._shape.output_grid_shape = tuple(grid_shapes[1])
        self.norm = norm
        self.centered = centered

    @staticmethod
    def fn(f, x, /):
        if f.centered:
            x = tfft.ifftshift(x, dim=f.dim)
        x = tfft.fftn(x, dim=f.dim, norm=f.norm)
        return tfft.fftshift(x, dim=f.dim) if f.centered else x

    @staticmethod
    def adj_fn(f, x, /):
        if f.centered:
            x = tfft.ifftshift(x, dim=f.dim)
        x = tfft.ifftn(x, dim=f.dim, norm=f.norm)
        return tfft.fftshift(x, dim=f.dim) if f.centered else x

    @staticmethod
    def normal_fn(f, x, /):
        return x

    def normal(self, inner=None):
        if inner is None:
            return Identity(self.ishape)
        return super().normal(inner)


def _einstr(shape) -> str:
    parts = []
    for s in shape:
        s = str(s)
        if s == "()":
            parts.append("any0")
        elif s.startswith("(") and s.endswith(")"):
            parts.append(f"any{s[1:-1]}")
        else:
            parts.append(s)
    return " ".join(parts)


class Dense(NamedLinop):
    """y = einsum(x, W) over named axes (dense.py:17-414); used for the
    Toeplitz point-spread kernel."""

    def __init__(self, weight, weightshape, ishape, oshape, broadcast_dims=None):
        super().__init__(NS(tuple(ishape), tuple(oshape)))
        self.weight = nn.Parameter(weight, requires_grad=False)
        self._shape.weightshape = tuple(weightshape)
        self._shape.broadcast_dims = list(broadcast_dims or [])

    @property
    def weightshape(self):
        return self._shape.weightshape

    @staticmethod
    def fn(d, x, /):
        return einsum(x, d.weight, f"{_einstr(d.ishape)},{_einstr(d.weightshape)}->{_einstr(d.oshape)}")

    @staticmethod
    def adj_fn(d, x, /):
        return einsum(x, d.weight.conj(), f"{_einstr(d.oshape)},{_einstr(d.weightshape)}->{_einstr(d.ishape)}")

    def adjoint(self):
        adj = copy(self)
        adj.weight = nn.Parameter(self.weight.conj(), requires_grad=False)
        adj._shape = adj._shape.H
        return adj

    def size(self, dim):
        if dim in self._shape.broadcast_dims:
            return None
        if dim in self.weightshape:
            return self.weight.shape[list(self.weightshape).index(dim)]
        return None
