"""Integer-index sampling: ``index`` / ``index_adjoint`` and their helpers.

Signatures, shapes and errors follow the reference's
src/torchlinops/functional/_index/index.py:
  * ``index``                 17-49  (gather over the last D axes)
  * ``multi_index``           52-72  (raveled gather; negative or large
                                      indices wrap per axis, ``% size``)
  * ``index_adjoint``         75-99  (IndexError outside [-size, size))
  * ``multi_grid``           140-180 (index_add_ into zeros)
  * ``ravel``                183-198
  * ``mask2idx`` / ``canonicalize_idx`` / ``slice2range`` /
    ``ensure_tensor_indexing``  102-137, 201-277
The gather and the scatter-add run in libtnl (tnl_index /
tnl_index_adjoint) for tensors on a HIP device; host tensors take
``index_select`` / a sequential ``index_add_`` (``_host.py``).  The scatter-add is a cell-major sum over a CSR map
of the points (``IndexMap``), so it is deterministic, and for cells hit by
at most ``HEAVY_MIN`` points it equals the CPU's sequential index_add_ bit
for bit (busier cells are summed by a workgroup each in a fixed order);
``deterministic_backward`` is accepted for signature compatibility and
changes nothing.
"""

from __future__ import annotations

from dataclasses import dataclass
from math import prod
from typing import Optional

import torch
from torch import Tensor
from torch.autograd import Function

from .._lib import TNL_F32, TNL_F64, TNL_I8, TNL_I16, TNL_I32, TNL_I64, TNL_U8, check, lib, stream_of
from ._host import gather_host, scatter_add_host

__all__ = [
    "index",
    "index_adjoint",
    "multi_index",
    "multi_grid",
    "ravel",
    "mask2idx",
    "canonicalize_idx",
    "slice2range",
    "ensure_tensor_indexing",
    "IndexMap",
    "index_map",
    "HEAVY_MIN",
]

# Cells hit by more points than this are summed by a workgroup each.
HEAVY_MIN = 64


# ---- index helpers (host logic on the index tensors) ---------------------


def ravel(x: Tensor, shape: tuple, dim: int) -> Tensor:
    """Row-major cell number of the index vectors along ``dim`` of ``x``,
    each component taken ``% size`` (non-negative), as index.py:183-198."""
    out = 0
    sizes = tuple(int(s) for s in shape)
    nxt = sizes[1:] + (1,)
    for i in range(x.shape[dim]):
        out = (out + torch.remainder(torch.select(x, dim, i), sizes[i])) * nxt[i]
    return out


def mask2idx(mask: Tensor) -> tuple:
    """Boolean mask -> tuple of integer index tensors of its True entries."""
    if mask.dtype != torch.bool:
        raise ValueError(f"Input tensor must be of boolean dtype, but got {mask.dtype}")
    return torch.nonzero(mask, as_tuple=True)


def canonicalize_idx(idx: Tensor, dim: int = -1) -> tuple:
    """``[B1..., D, B2...]`` integer tensor -> D-tuple of ``[B1..., B2...]``."""
    return tuple(torch.select(idx, dim, i) for i in range(idx.shape[dim]))


def slice2range(slice_obj: slice, n: int) -> range:
    """``slice`` -> ``range`` over an axis of size ``n`` (unnormalised
    start/stop, as index.py:201-222).

    >>> tuple(slice2range(slice(None, None, None), 4))
    (0, 1, 2, 3)
    >>> tuple(slice2range(slice(None, None, -1), 3))
    (2, 1, 0)
    """
    step = slice_obj.step if slice_obj.step is not None else 1
    if slice_obj.start is not None:
        start = slice_obj.start
    else:
        start = 0 if step > 0 else n - 1
    if slice_obj.stop is not None:
        stop = slice_obj.stop
    else:
        stop = n if step > 0 else -1
    return range(start, stop, step)


def ensure_tensor_indexing(idx: tuple, tshape) -> tuple:
    """Left-pad ``idx`` with full slices to ``len(tshape)`` and turn every
    slice into an integer range tensor shaped to broadcast over the axes
    after it (index.py:238-277)."""
    idx = list(idx)
    if len(tshape) < len(idx):
        raise ValueError(f"Cannot broadcast idx {idx} to tshape {tshape}")
    idx = [slice(None)] * (len(tshape) - len(idx)) + idx
    out = []
    for d, (size, i) in enumerate(zip(tshape, idx)):
        if isinstance(i, Tensor):
            out.append(i)
        elif isinstance(i, slice):
            r = torch.tensor(slice2range(i, size))
            out.append(r.view(-1, *((1,) * (len(tshape) - d - 1))))
        else:
            raise ValueError(
                f"idx must contain only tensors or slice() objects but got {i}"
            )
    return tuple(out)


@dataclass
class IndexMap:
    """Raveled cell per point and the points grouped by cell.

    ``flat[p]`` is point p's cell in ``[0, ngrid)``; ``order`` lists the
    points sorted by cell (stable: ascending within a cell) and
    ``starts[c]:starts[c+1]`` is cell c's run in ``order``; ``heavy``
    lists the cells with more than ``HEAVY_MIN`` points.  The grouping is
    built on first use by the scatter-add (the gather needs only
    ``flat``).  Built once per index set; ``index``/``index_adjoint`` of
    the Sampling linop reuse it.
    """

    flat: Tensor
    out_shape: tuple
    grid_size: tuple
    order: Optional[Tensor] = None
    starts: Optional[Tensor] = None
    heavy: Optional[Tensor] = None

    @property
    def ngrid(self) -> int:
        return prod(self.grid_size)

    @property
    def npts(self) -> int:
        return self.flat.numel()

    def grouped(self) -> "IndexMap":
        """Build ``order`` / ``starts`` / ``heavy`` if not built yet."""
        if self.order is None:
            ngrid = self.ngrid
            self.order = torch.sort(self.flat, stable=True).indices.contiguous()
            counts = torch.bincount(self.flat, minlength=ngrid)
            starts = torch.zeros(ngrid + 1, dtype=torch.int64, device=self.flat.device)
            starts[1:] = torch.cumsum(counts, 0)
            self.starts = starts
            self.heavy = torch.nonzero(counts > HEAVY_MIN).reshape(-1).contiguous()
        return self


def index_map(idx: tuple, grid_size: tuple, device=None, flat: Tensor = None) -> IndexMap:
    """Build the ``IndexMap`` of a tuple of broadcastable integer index
    tensors over the last ``len(grid_size)`` axes of a grid."""
    grid_size = tuple(int(s) for s in grid_size)
    if flat is None:
        idx = torch.broadcast_tensors(*idx)
        out_shape = tuple(idx[0].shape)
        stacked = torch.stack([i.to(device=device, dtype=torch.int64) for i in idx])
        flat = ravel(stacked, grid_size, dim=0).reshape(-1).contiguous()
    else:
        out_shape = tuple(flat.shape)
        flat = flat.to(device=device, dtype=torch.int64).reshape(-1).contiguous()
    return IndexMap(flat, out_shape, grid_size)


# ---- HIP launches --------------------------------------------------------


_INT_CODES = {torch.int32: TNL_I32, torch.int64: TNL_I64, torch.int16: TNL_I16,
              torch.int8: TNL_I8, torch.uint8: TNL_U8}


def _real_dtype_code(dtype: torch.dtype) -> tuple:
    if dtype in (torch.float32, torch.complex64):
        return TNL_F32, dtype.is_complex
    if dtype in (torch.float64, torch.complex128):
        return TNL_F64, dtype.is_complex
    if dtype in _INT_CODES:
        return _INT_CODES[dtype], False
    raise TypeError(
        f"index_adjoint supports float32/float64/complex64/complex128 and int8/uint8/int16/"
        f"int32/int64, got {dtype}"
    )


def _gather(vals: Tensor, imap: IndexMap) -> Tensor:
    nd = len(imap.grid_size)
    if vals.ndim < nd or tuple(vals.shape[vals.ndim - nd:]) != imap.grid_size:
        raise ValueError(
            f"Input value with shape {tuple(vals.shape)} does not end in the indexed "
            f"grid {imap.grid_size}"
        )
    batch = tuple(vals.shape[: vals.ndim - nd])
    x = vals.contiguous()
    nbatch = prod(batch)
    if not vals.is_cuda:
        flat = imap.flat.to(vals.device)
        out = gather_host(x.reshape(nbatch, imap.ngrid), flat)
        return out.reshape(batch + imap.out_shape)
    out = torch.empty(batch + imap.out_shape, dtype=vals.dtype, device=vals.device)
    if out.numel() == 0:
        return out
    with torch.cuda.device(vals.device):
        check(lib().tnl_index(x.data_ptr(), imap.flat.data_ptr(), out.data_ptr(),
                              x.element_size(), nbatch, imap.ngrid, imap.npts,
                              stream_of(x)))
    return out


def _scatter(vals: Tensor, imap: IndexMap) -> Tensor:
    nl = len(imap.out_shape)
    batch = tuple(vals.shape[: vals.ndim - nl])
    if tuple(vals.shape[vals.ndim - nl:]) != imap.out_shape:
        raise ValueError(
            f"x and idx should correspond in last {nl} dimensions but got x: "
            f"{tuple(vals.shape)} and idx: {imap.out_shape}"
        )
    x = vals.contiguous()
    nbatch = prod(batch)
    if not vals.is_cuda:
        flat = imap.flat.to(vals.device)
        out = scatter_add_host(x.reshape(nbatch, imap.npts), flat, imap.ngrid)
        return out.reshape(batch + imap.grid_size)
    code, cplx = _real_dtype_code(vals.dtype)
    out = torch.empty(batch + imap.grid_size, dtype=vals.dtype, device=vals.device)
    if out.numel() == 0:
        return out
    imap.grouped()
    with torch.cuda.device(vals.device):
        nheavy = imap.heavy.numel()
        check(lib().tnl_index_adjoint(x.data_ptr(), imap.order.data_ptr(),
                                      imap.starts.data_ptr(),
                                      imap.heavy.data_ptr() if nheavy else None,
                                      nheavy, HEAVY_MIN, out.data_ptr(), code,
                                      int(cplx), nbatch, imap.ngrid, imap.npts,
                                      stream_of(x)))
    return out


class _IndexFn(Function):
    @staticmethod
    def forward(ctx, vals, imap):
        ctx.imap = imap
        return _gather(vals, imap)

    @staticmethod
    def backward(ctx, g):
        return _scatter(g, ctx.imap), None


class _IndexAdjointFn(Function):
    @staticmethod
    def forward(ctx, vals, imap):
        ctx.imap = imap
        return _scatter(vals, imap)

    @staticmethod
    def backward(ctx, g):
        return _gather(g, ctx.imap), None


def index_with_map(vals: Tensor, imap: IndexMap) -> Tensor:
    """``index`` over a prebuilt ``IndexMap``: ``[N..., *grid] -> [N..., I...]``."""
    return _IndexFn.apply(vals, imap)


def index_adjoint_with_map(vals: Tensor, imap: IndexMap) -> Tensor:
    """``index_adjoint`` over a prebuilt ``IndexMap``: ``[N..., I...] -> [N..., *grid]``."""
    return _IndexAdjointFn.apply(vals, imap)


# ---- public API (mirrors the reference) ----------------------------------


def index(vals: Tensor, idx: tuple, deterministic_backward: bool = False) -> Tensor:
    """Gather ``vals[..., idx[0], ..., idx[D-1]]`` over the last D axes.

    Integer index tensors broadcast together and wrap per axis
    (``% size``), as the reference's ``multi_index``.  With slices in
    ``idx`` the result follows torch's advanced indexing (the reference's
    deterministic branch): the cell each output element takes is found by
    indexing a grid of cell numbers the same way.
    """
    if len(vals.shape) < len(idx):
        raise ValueError(
            f"Input value with shape {vals.shape} cannot be indexed with index "
            f"tensors of length {len(idx)}"
        )
    nd = len(idx)
    grid_size = tuple(vals.shape[vals.ndim - nd:])
    if any(isinstance(i, slice) for i in idx):
        cells = torch.arange(prod(grid_size), device=vals.device).view(grid_size)
        dev_idx = tuple(i.to(vals.device) if isinstance(i, Tensor) else i for i in idx)
        imap = index_map(None, grid_size, vals.device, flat=cells[dev_idx])
        out = _IndexFn.apply(vals, imap)
        pos = [d for d, i in enumerate(idx) if isinstance(i, Tensor)]
        nb = vals.ndim - nd
        if nb and pos and pos != list(range(pos[0], pos[-1] + 1)):
            # torch puts the broadcast axes of non-adjacent index tensors
            # in front of every sliced axis, the batch axes included
            nadv = len(torch.broadcast_shapes(*(idx[d].shape for d in pos)))
            out = out.movedim(tuple(range(nb, nb + nadv)), tuple(range(nadv)))
        return out
    imap = index_map(idx, grid_size, vals.device)
    return _IndexFn.apply(vals, imap)


def multi_index(x: Tensor, ndims: int, idx: Tensor, raveled: bool = False) -> Tensor:
    """Gather over the last ``ndims`` axes with stacked ``idx [ndims, I...]``
    (or raveled ``[I...]``): ``[N..., D...] -> [N..., I...]``."""
    grid_size = tuple(x.shape[x.ndim - ndims:])
    if raveled:
        imap = index_map(None, grid_size, x.device, flat=idx)
    else:
        assert ndims == idx.shape[0], (
            "idx must have same last dimension as number of indexed dimensions"
        )
        imap = index_map(tuple(idx), grid_size, x.device)
    return _IndexFn.apply(x, imap)


def index_adjoint(vals: Tensor, idx: tuple, grid_size: tuple) -> Tensor:
    """Scatter-add ``vals [N..., I...]`` into zeros ``[N..., *grid_size]``
    at the cells ``idx`` names (adjoint of ``index``)."""
    grid_size = tuple(int(s) for s in grid_size)
    for d, (dim_idx, dim_size) in enumerate(zip(idx, grid_size)):
        if (dim_idx >= dim_size).any() or (dim_idx < -dim_size).any():
            raise IndexError(f"Out-of-bounds index for grid of shape {grid_size}: idx[{d}]")
    idx_stacked = torch.stack(tuple(idx), dim=0)
    return multi_grid(vals, idx_stacked, grid_size)


def multi_grid(x: Tensor, idx: Tensor, final_size: tuple, raveled: bool = False,
               ravel_dim: int = 0) -> Tensor:
    """Scatter-add ``x [N..., I...]`` into ``[N..., *final_size]`` at
    ``idx`` (stacked along ``ravel_dim``, or raveled ``[I...]``)."""
    final_size = tuple(int(s) for s in final_size)
    if raveled:
        flat = idx
    else:
        if len(final_size) != idx.shape[ravel_dim]:
            raise ValueError(
                f"final_size should be of dimension {idx.shape[-1]} but got {final_size}"
            )
        flat = ravel(idx.to(x.device), final_size, dim=ravel_dim)
    imap = index_map(None, final_size, x.device, flat=flat)
    if tuple(x.shape[x.ndim - flat.ndim:]) != tuple(flat.shape):
        raise ValueError(
            f"x and idx should correspond in last {flat.ndim} dimensions but got x: "
            f"{tuple(x.shape)} and idx: {tuple(flat.shape)}"
        )
    return _IndexAdjointFn.apply(x, imap)
