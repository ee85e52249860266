"""Host (CPU) path of the gridding ops, for tensors that are not on a HIP
device.

The reference dispatches every op on ``x.is_cuda``: device tensors go to its
GPU kernels, host tensors to a torch fallback (``grid_torch``
src/torchlinops/functional/_interp/grid.py:571-639, ``ungrid_torch``
ungrid.py:624-684, ``_unfold_torch`` _unfold/unfold.py:531-561,
``_fold_torch`` _unfold/fold.py:489-517, ``index_select`` / ``index_add_``
in _index/index.py:17-99).  This module is that branch for this package:
device tensors always take libtnl's HIP kernels and never reach it.

Conventions are the ones the HIP kernels implement (so a host result and a
device result agree to rounding):
  * taps ``g = ceil(t - w/2) + k``, ``k = 0..ceil(w)``, kept iff
    ``|t - g| <= w/2``; circular padding wraps ``g mod N``, zero padding drops
    taps outside ``[0, N)``;
  * ``norm=1`` weight ``prod_d K((t_d - g_d) / (w_d/2))``; ``norm=2`` weight
    ``K(||(t - g) / (w/2)||)``, zero when that norm exceeds 1;
  * Kaiser-Bessel ``K(x) = I0(beta sqrt(1 - x^2))`` (unnormalised), spline
    ``max(1 - |x|, 0)``; non-finite weights (non-finite coordinates) are 0;
  * arithmetic in the values' precision (float32 / complex64 or float64 /
    complex128), points taken in chunks so the tap tensors stay small;
  * fold sums the covering blocks of each cell in lexicographic block order,
    the reference's sequential order, so the result is bit-identical to it.
"""

from __future__ import annotations

from itertools import product
from math import ceil, prod

import torch

__all__ = ["ungrid_host", "grid_host", "unfold_host", "fold_host", "gather_host",
           "scatter_add_host"]

# points per chunk: a 3-D chunk holds chunk x 125 tap weights and cell
# indices plus nbatch x chunk x 125 gathered values
_CHUNK = 1 << 14


def _window(kernel: str, x: torch.Tensor, beta: float) -> torch.Tensor:
    if kernel == "kaiser_bessel":
        z = beta * torch.sqrt(torch.clamp(1.0 - x * x, min=0.0))
        return torch.special.i0(z)
    return torch.clamp(1.0 - x.abs(), min=0.0)


def _taps(locs: torch.Tensor, grid_size, width, kernel, norm, pad_mode, beta, wdtype):
    """Cells ``[P, G]`` (int64, raveled) and weights ``[P, G]`` of a chunk of
    points ``locs [P, D]`` (taps in lexicographic order, last axis fastest)."""
    P, nd = locs.shape
    t = locs.to(wdtype)
    cells = torch.zeros(P, *([1] * nd), dtype=torch.int64)
    keep = torch.ones(P, *([1] * nd), dtype=torch.bool)
    wt = torch.ones(P, *([1] * nd), dtype=wdtype)
    s2 = torch.zeros(P, *([1] * nd), dtype=wdtype)
    for d in range(nd):
        half = width[d] / 2.0
        ntap = int(ceil(width[d])) + 1
        lower = torch.ceil(t[:, d] - half)
        lower = torch.where(torch.isfinite(lower), lower, torch.zeros_like(lower))
        g = lower[:, None] + torch.arange(ntap, dtype=wdtype)[None, :]  # [P, ntap]
        diff = t[:, d, None] - g
        dn = diff / half
        k = (diff.abs() <= half)
        gi = g.to(torch.int64)
        n = int(grid_size[d])
        if pad_mode == "zero":
            k = k & (gi >= 0) & (gi < n)
        gi = torch.remainder(gi, n)
        shape = [P] + [1] * nd
        shape[1 + d] = ntap
        cells = cells * n + gi.view(shape)
        keep = keep & k.view(shape)
        if norm == 1:
            wt = wt * _window(kernel, dn, beta).view(shape)
        else:
            s2 = s2 + (dn * dn).view(shape)
    if norm == 2:
        r = torch.sqrt(s2)
        wt = torch.where(r <= 1.0, _window(kernel, r, beta), torch.zeros_like(r))
    wt = torch.where(keep & torch.isfinite(wt), wt, torch.zeros_like(wt))
    cells = torch.where(keep, cells, torch.zeros_like(cells))
    return cells.reshape(P, -1), wt.reshape(P, -1)


def _real_dtype(dtype: torch.dtype) -> torch.dtype:
    if dtype in (torch.complex64, torch.float32):
        return torch.float32
    if dtype in (torch.complex128, torch.float64):
        return torch.float64
    raise TypeError(f"interpolation computes in float32/float64 (or complex); got {dtype}")


def ungrid_host(vals: torch.Tensor, locs2d: torch.Tensor, grid_size, width, kernel, norm,
                pad_mode, beta) -> torch.Tensor:
    """``out[n, p] = sum_taps w(p, g) vals[n, g]`` for ``vals [nbatch, *grid]``,
    ``locs2d [npts, D]``; returns ``[nbatch, npts]``."""
    nbatch = vals.shape[0]
    npts = locs2d.shape[0]
    wdtype = _real_dtype(vals.dtype)
    flat = vals.reshape(nbatch, -1)
    out = torch.zeros(nbatch, npts, dtype=vals.dtype)
    for p0 in range(0, npts, _CHUNK):
        p1 = min(npts, p0 + _CHUNK)
        cells, w = _taps(locs2d[p0:p1], grid_size, width, kernel, norm, pad_mode, beta,
                         wdtype)
        out[:, p0:p1] = (flat[:, cells] * w).sum(dim=-1)
    return out


def grid_host(vals: torch.Tensor, locs2d: torch.Tensor, grid_size, width, kernel, norm,
              pad_mode, beta) -> torch.Tensor:
    """``out[n, g] += w(p, g) vals[n, p]`` for ``vals [nbatch, npts]``; returns
    ``[nbatch, *grid_size]`` (contributions added point by point in order)."""
    nbatch, npts = vals.shape
    wdtype = _real_dtype(vals.dtype)
    out = torch.zeros(nbatch, prod(grid_size), dtype=vals.dtype)
    for p0 in range(0, npts, _CHUNK):
        p1 = min(npts, p0 + _CHUNK)
        cells, w = _taps(locs2d[p0:p1], grid_size, width, kernel, norm, pad_mode, beta,
                         wdtype)
        contrib = vals[:, p0:p1, None] * w[None]
        out.index_add_(1, cells.reshape(-1), contrib.reshape(nbatch, -1))
    return out.reshape(nbatch, *grid_size)


def unfold_host(x: torch.Tensor, block_size, stride) -> torch.Tensor:
    """``[nbatch, *im] -> [nbatch, *nblocks, *block]`` as strided windows
    (``Tensor.unfold`` appends each axis' window after the block-index axes,
    which is exactly this layout); a pure copy."""
    nd = len(block_size)
    y = x
    for d in range(nd):
        y = y.unfold(1 + d, int(block_size[d]), int(stride[d]))
    return y.contiguous()


def fold_host(x: torch.Tensor, im_size, block_size, stride, nblocks) -> torch.Tensor:
    """``[nbatch, *nblocks, *block] -> [nbatch, *im]``: each cell sums its
    covering blocks in lexicographic block order.

    Along axis d a cell at ``i`` is covered by blocks ``k`` with
    ``k s <= i < k s + b``, i.e. ``k = kmin(i) .. kmax(i)``.  The covering
    blocks of a cell are the product of those ranges, so lexicographic block
    order is the lexicographic order of the per-axis ranks ``m_d = k_d -
    kmin_d``.  Walking the rank tuples in that order and adding, for every
    cell at once, its block of that rank (or +0.0 where there is none, which
    leaves the sum unchanged) reproduces the sequential sum bit for bit."""
    nd = len(im_size)
    nbatch = x.shape[0]
    out = torch.zeros(nbatch, *im_size, dtype=x.dtype)
    ks, offs, oks, nr = [], [], [], []
    for d in range(nd):
        i = torch.arange(int(im_size[d]))
        s, b, nb = int(stride[d]), int(block_size[d]), int(nblocks[d])
        kmin = torch.clamp(-((b - 1 - i) // s), min=0)  # ceil((i - b + 1) / s), >= 0
        kmax = torch.clamp(i // s, max=nb - 1)
        ks.append((kmin, kmax))
        nr.append(-(-b // s))
        offs.append(i)
    for ranks in product(*(range(r) for r in nr)):
        idx, valid = [], None
        for d in range(nd):
            kmin, kmax = ks[d]
            k = kmin + ranks[d]
            ok = k <= kmax
            k = torch.where(ok, k, torch.zeros_like(k))
            o = torch.where(ok, offs[d] - k * int(stride[d]), torch.zeros_like(k))
            shape = [1] * nd
            shape[d] = -1
            idx.append((k.view(shape), o.view(shape)))
            v = ok.view(shape)
            valid = v if valid is None else valid & v
        kidx = tuple(k for k, _ in idx)
        oidx = tuple(o for _, o in idx)
        term = x[(slice(None), *kidx, *oidx)]
        out += torch.where(valid, term, torch.zeros((), dtype=x.dtype))
    return out


def gather_host(vals: torch.Tensor, flat: torch.Tensor) -> torch.Tensor:
    """``vals [nbatch, ngrid] -> [nbatch, npts]`` at the raveled cells."""
    return vals.index_select(1, flat)


def scatter_add_host(vals: torch.Tensor, flat: torch.Tensor, ngrid: int) -> torch.Tensor:
    """``vals [nbatch, npts]`` added into zeros ``[nbatch, ngrid]`` point by
    point in ascending order (``index_add_`` on the host is sequential)."""
    out = torch.zeros(vals.shape[0], ngrid, dtype=vals.dtype)
    return out.index_add_(1, flat, vals)
