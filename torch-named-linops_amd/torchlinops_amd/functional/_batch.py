"""Many independent interpolation problems in one persistent launch.

``BatchedInterp`` holds a set of problems of one dimensionality, kernel,
norm and padding, each with its own trajectory, grid size, width, beta and
channel count.  Their grid-side data live in one flat buffer and their
point-side data in another (``grid_views`` / ``pts_views`` give each
problem's ``[nbatch, *grid]`` / ``[nbatch, npts]`` view).  ``ungrid`` and
``grid`` run every problem in one persistent launch each
(``tnl_ungrid_batched`` / ``tnl_grid_batched``), with no host work per call,
so a solver loop around them can be captured in a graph.  Results equal the
per-problem ``interpolate`` / ``interpolate_adjoint`` with the same tiling.
The reference launches per problem (ungrid.py:86-119, grid.py:91-124).
"""

from __future__ import annotations

import ctypes
from math import prod
from typing import Optional, Sequence

import torch

from .._lib import (
    KERNELS,
    PAD_MODES,
    BatchLayout,
    BatchProblem,
    check,
    f32_array,
    i64_array,
    lib,
    require_device,
    stream_of,
)
from ._interp import _width_tuple, get_kernel_params
from ._plan import build_plan

__all__ = ["BatchedInterp"]


class BatchedInterp:
    def __init__(
        self,
        locs: Sequence[torch.Tensor],
        grid_sizes: Sequence[Sequence[int]],
        widths,
        nbatches: Sequence[int],
        kernel: str = "kaiser_bessel",
        norm: int = 1,
        pad_mode: str = "circular",
        kernel_params=None,
        dtype: torch.dtype = torch.complex64,
        tuning: Optional[dict] = None,
    ):
        n = len(locs)
        if n < 1 or len(grid_sizes) != n or len(nbatches) != n:
            raise ValueError("locs, grid_sizes and nbatches must have one entry per problem")
        if kernel not in KERNELS:
            raise ValueError(f"Unrecognized kernel {kernel}")
        if pad_mode not in PAD_MODES:
            raise ValueError(f"Unrecognized pad_mode {pad_mode}")
        if dtype not in (torch.float32, torch.complex64):
            raise TypeError("batched interpolation computes in float32/complex64")
        # widths: one number for all problems, or one entry (number or
        # per-axis tuple) per problem
        if isinstance(widths, (int, float)):
            widths = [widths] * n
        if len(widths) != n:
            raise ValueError("widths must be a number or have one entry per problem")
        if isinstance(kernel_params, (list, tuple)):
            kps = [get_kernel_params(kernel, k) for k in kernel_params]
        else:
            kps = [get_kernel_params(kernel, kernel_params)] * n
        require_device(*locs, op="BatchedInterp")
        self.ndim = locs[0].shape[-1]
        self.device = locs[0].device
        self.dtype = dtype
        self.is_complex = dtype == torch.complex64
        self.grid_sizes = [tuple(int(g) for g in gs) for gs in grid_sizes]
        self.nbatches = [int(c) for c in nbatches]
        self.npts = [prod(l.shape[:-1]) for l in locs]
        self.kernel, self.norm, self.pad_mode = kernel, int(norm), pad_mode
        pad_code = PAD_MODES[pad_mode]
        self.plans, self.locs2d = [], []
        probs = (BatchProblem * n)()
        goff = poff = 0
        self.grid_offsets, self.pts_offsets = [], []
        for i in range(n):
            l = locs[i]
            if l.shape[-1] != self.ndim or len(self.grid_sizes[i]) != self.ndim:
                raise ValueError("every problem of a batch must have the same ndim")
            w = _width_tuple(widths[i], self.ndim)
            l2 = l.reshape(self.npts[i], self.ndim).to(torch.float32).contiguous()
            plan = build_plan(l2, self.ndim, self.grid_sizes[i], w, pad_code, tuning)
            self.plans.append(plan)
            self.locs2d.append(l2)
            p = probs[i]
            p.nbatch = self.nbatches[i]
            p.npts = self.npts[i]
            for d in range(self.ndim):
                p.grid_size[d] = self.grid_sizes[i][d]
                p.width[d] = w[d]
            p.beta = float(kps[i].get("beta", 1.0)) if kernel == "kaiser_bessel" else 1.0
            p.grid_offset = goff
            p.pts_offset = poff
            p.layout = ctypes.pointer(plan.layout)
            p.plan = plan.ptr
            self.grid_offsets.append(goff)
            self.pts_offsets.append(poff)
            goff += self.nbatches[i] * prod(self.grid_sizes[i])
            poff += self.nbatches[i] * self.npts[i]
        self.grid_numel, self.pts_numel = goff, poff
        self._probs = probs
        self.layout = BatchLayout()
        check(lib().tnl_batch_layout_query(
            ctypes.addressof(probs), n, self.ndim, KERNELS[kernel], self.norm, pad_code,
            int(self.is_complex), self.layout))
        self.workspace = torch.empty(int(self.layout.bytes), dtype=torch.uint8,
                                     device=self.device)
        with torch.cuda.device(self.device):
            check(lib().tnl_batch_build(ctypes.addressof(probs), self.layout,
                                        self.workspace.data_ptr(),
                                        stream_of(self.workspace)))

    # ---- buffers
    def new_grid(self) -> torch.Tensor:
        return torch.empty(self.grid_numel, dtype=self.dtype, device=self.device)

    def new_points(self) -> torch.Tensor:
        return torch.empty(self.pts_numel, dtype=self.dtype, device=self.device)

    def grid_views(self, flat: torch.Tensor):
        return [flat[o:o + c * prod(g)].view(c, *g) for o, c, g in
                zip(self.grid_offsets, self.nbatches, self.grid_sizes)]

    def pts_views(self, flat: torch.Tensor):
        return [flat[o:o + c * m].view(c, m) for o, c, m in
                zip(self.pts_offsets, self.nbatches, self.npts)]

    def _check(self, t: torch.Tensor, numel: int, what: str):
        if t.dtype != self.dtype or t.numel() != numel or not t.is_contiguous():
            raise RuntimeError(f"{what}: expected a contiguous {self.dtype} buffer of "
                               f"{numel} elements, got {t.dtype} {tuple(t.shape)}")
        if t.device != self.device:
            raise RuntimeError(f"{what}: buffer on {t.device}, batch on {self.device}")

    def _raw(self, t: torch.Tensor) -> int:
        return (torch.view_as_real(t) if self.is_complex else t).data_ptr()

    # ---- the two launches
    def ungrid(self, grid_flat: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        self._check(grid_flat, self.grid_numel, "ungrid input")
        out = self.new_points() if out is None else out
        self._check(out, self.pts_numel, "ungrid output")
        with torch.cuda.device(self.device):
            check(lib().tnl_ungrid_batched(self._raw(grid_flat), self._raw(out), self.layout,
                                           self.workspace.data_ptr(), stream_of(out)))
        return out

    def grid(self, pts_flat: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        self._check(pts_flat, self.pts_numel, "grid input")
        out = self.new_grid() if out is None else out
        self._check(out, self.grid_numel, "grid output")
        with torch.cuda.device(self.device):
            check(lib().tnl_grid_batched(self._raw(pts_flat), self._raw(out), self.layout,
                                         self.workspace.data_ptr(), stream_of(out)))
        return out

    # ---- the same work launched problem by problem (same plans)
    def _each(self, adjoint: bool, src: torch.Tensor, dst: torch.Tensor) -> None:
        fn = lib().tnl_grid if adjoint else lib().tnl_ungrid
        sv = self.pts_views(src) if adjoint else self.grid_views(src)
        dv = self.grid_views(dst) if adjoint else self.pts_views(dst)
        with torch.cuda.device(self.device):
            for i, plan in enumerate(self.plans):
                p = self._probs[i]
                check(fn(self._raw(sv[i]), self.locs2d[i].data_ptr(), self._raw(dv[i]),
                         self.nbatches[i], self.npts[i],
                         self.ndim, i64_array(self.grid_sizes[i]),
                         f32_array(list(p.width)[:self.ndim]), KERNELS[self.kernel], self.norm,
                         PAD_MODES[self.pad_mode], int(self.is_complex), float(p.beta),
                         plan.layout, plan.ptr, stream_of(dst)))

    def ungrid_each(self, grid_flat: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """``ungrid`` as one launch per problem (for comparison)."""
        self._check(grid_flat, self.grid_numel, "ungrid input")
        self._check(out, self.pts_numel, "ungrid output")
        self._each(False, grid_flat, out)
        return out

    def grid_each(self, pts_flat: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """``grid`` as one launch per problem (for comparison)."""
        self._check(pts_flat, self.pts_numel, "grid input")
        self._check(out, self.grid_numel, "grid output")
        self._each(True, pts_flat, out)
        return out
