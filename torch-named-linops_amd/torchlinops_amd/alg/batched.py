"""Conjugate gradients over many independent NUFFT normal equations.

``BatchedNUFFTNormal`` applies A_i^H A_i for the NUFFTs A_i of independent
problems (own image size, coil count, trajectory and kernel width) held
back to back in one flat image buffer.  Per problem it follows the NUFFT
chain of src/torchlinops/linops/nufft.py:123-184 (apodize, centred pad,
centred FFT, interpolate, scale) and its adjoint; the interpolation and
gridding of all problems run in one persistent launch each
(``functional.BatchedInterp``).

``SegmentedCG`` is the CG of src/torchlinops/alg/pcg.py:67-87 run on every
problem at once with per-problem scalars kept on the device (the
``tnl_seg_*`` kernels: float64 segmented dot products, fused updates).  A
block of iterations can be captured once in a graph and replayed.
"""

from __future__ import annotations

from math import prod
from typing import Callable, Optional, Sequence

import torch

from .._lib import SegLayout, check, i64_array, lib, stream_of
from ..functional._batch import BatchedInterp
from ..linops.nufft import NUFFT, shifted_layout
from ..linops.toeplitz import crop_wrapped, pad_wrapped

__all__ = ["BatchedNUFFTNormal", "SegmentedCG"]


class BatchedNUFFTNormal:
    def __init__(
        self,
        locs: Sequence[torch.Tensor],
        im_sizes: Sequence[Sequence[int]],
        ncoils: Sequence[int],
        oversamp: float = 2.0,
        widths=4.0,
        tuning: Optional[dict] = None,
    ):
        n = len(locs)
        if isinstance(widths, (int, float)):
            widths = [float(widths)] * n
        self.device = locs[0].device
        self.ndim = locs[0].shape[-1]
        self.dim = tuple(range(-self.ndim, 0))
        self.im_sizes = [tuple(int(s) for s in im) for im in im_sizes]
        self.ncoils = [int(c) for c in ncoils]
        self.padded = [tuple(int(s * oversamp) for s in im) for im in self.im_sizes]
        betas, plocs = [], []
        self.apod, self.apod_fwd, self.index = [], [], []
        for i in range(n):
            w = float(widths[i])
            beta = float(NUFFT.beta(w, oversamp))
            betas.append({"beta": beta})
            # no centring shifts: the image goes to the wrapped cells and the
            # locations are relabelled (see shifted_layout)
            index, shift = shifted_layout(self.im_sizes[i], self.padded[i], self.device)
            self.index.append(index)
            pl = NUFFT.prep_locs(locs[i], self.im_sizes[i], self.padded[i])
            pad = torch.tensor(self.padded[i], dtype=torch.float32, device=pl.device)
            plocs.append(torch.remainder(pl - shift.to(pl.device), pad).contiguous())
            apod = NUFFT.apodize_weights(self.im_sizes[i], self.padded[i], oversamp, w, beta)
            scale = w**self.ndim * (prod(self.im_sizes[i]) / prod(self.padded[i])) ** 0.5
            # forward scale 1/scale and adjoint scale 1/scale fold into the
            # adjoint's weight, the orthonormal FFT scalings 1/sqrt(prod P)
            # into both (the FFTs run unscaled)
            rp = prod(self.padded[i]) ** 0.5
            self.apod.append((apod / (scale**2 * rp)).to(self.device, torch.float32).contiguous())
            self.apod_fwd.append((apod / rp).to(self.device, torch.float32).contiguous())
        # Grids and images are laid out grouped by padded size, largest grids
        # first and, within a group, the most coil-points first (the
        # persistent batched kernels then take the longest work first): each
        # group's FFT is one batched transform over its coils, and
        # consecutive problems with equal image size and width share one
        # weighted pad / crop launch.  Measured on 64 uneven problems: this
        # order 267 ms for 20 CG iterations, ascending size 271 ms, grouping
        # by width inside a size 285 ms.  ``image_views`` still returns
        # the problems in the caller's order; ``seg_elems`` (and so
        # SegmentedCG's per-segment results) follow the memory order
        # ``order`` (``position[i]`` is problem i's segment).
        work = [self.ncoils[i] * locs[i].numel() for i in range(n)]
        order = sorted(range(n), key=lambda i: (-prod(self.padded[i]), -work[i]))
        self.order = order
        self.position = [0] * n
        for pos, i in enumerate(order):
            self.position[i] = pos
        self.interp = BatchedInterp([plocs[i] for i in order], [self.padded[i] for i in order],
                                    [widths[i] for i in order], [self.ncoils[i] for i in order],
                                    kernel_params=[betas[i] for i in order], tuning=tuning)
        self.G = self.interp.new_grid()
        self.P = self.interp.new_points()
        self.im_offsets = [0] * n
        off = 0
        for i in order:
            self.im_offsets[i] = off
            off += self.ncoils[i] * prod(self.im_sizes[i])
        self.numel = off
        self.seg_elems = [self.ncoils[i] * prod(self.im_sizes[i]) for i in order]
        # FFT groups: [first position, end position, padded size]
        self._groups = []
        for pos, i in enumerate(order):
            if self._groups and self._groups[-1][2] == self.padded[i]:
                self._groups[-1][1] = pos + 1
            else:
                self._groups.append([pos, pos + 1, self.padded[i]])
        # weight groups: one pad / crop launch each
        self._wgroups = []
        for fg, (a, b, pad) in enumerate(self._groups):
            row = 0
            for pos in range(a, b):
                i = order[pos]
                key = (self.im_sizes[i], float(widths[i]))
                c = self.ncoils[i]
                if self._wgroups and self._wgroups[-1]["fg"] == fg and \
                        self._wgroups[-1]["key"] == key:
                    self._wgroups[-1]["coils"] += c
                else:
                    self._wgroups.append(dict(
                        fg=fg, key=key, row=row, coils=c, im=self.im_sizes[i], pad=pad,
                        xoff=self.im_offsets[i], goff=self.interp.grid_offsets[pos],
                        wf=self.apod_fwd[i], wa=self.apod[i]))
                row += c

    def new_image(self) -> torch.Tensor:
        return torch.empty(self.numel, dtype=torch.complex64, device=self.device)

    def image_views(self, flat: torch.Tensor):
        """Per-problem [coils, *im] views of a flat image buffer, in the
        caller's problem order."""
        return [flat[o:o + c * prod(im)].view(c, *im) for o, c, im in
                zip(self.im_offsets, self.ncoils, self.im_sizes)]

    def _group_views(self, flat: torch.Tensor):
        """One [sum of coils, *padded] view of the grid buffer per size group."""
        offs, nb, gs = self.interp.grid_offsets, self.interp.nbatches, self.interp.grid_sizes
        views = []
        for a, b, pad in self._groups:
            end = offs[b - 1] + nb[b - 1] * prod(gs[b - 1])
            views.append(flat[offs[a]:end].view(sum(nb[a:b]), *pad))
        return views

    def __call__(self, x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        out = self.new_image() if out is None else out
        for w in self._wgroups:
            # apodized images at the wrapped cells (zeros elsewhere)
            c, im, pad = w["coils"], w["im"], w["pad"]
            xg = x[w["xoff"]:w["xoff"] + c * prod(im)].view(c, *im)
            gg = self.G[w["goff"]:w["goff"] + c * prod(pad)].view(c, *pad)
            pad_wrapped(xg, im, pad, weight=w["wf"], out=gg)
        for g in self._group_views(self.G):  # unscaled FFTs, one per size group
            g.copy_(torch.fft.fftn(g, dim=self.dim, norm="backward"))
        self.interp.ungrid(self.G, out=self.P)
        self.interp.grid(self.P, out=self.G)
        spec = [torch.fft.ifftn(g, dim=self.dim, norm="forward")
                for g in self._group_views(self.G)]
        for w in self._wgroups:
            c, im, pad = w["coils"], w["im"], w["pad"]
            og = out[w["xoff"]:w["xoff"] + c * prod(im)].view(c, *im)
            crop_wrapped(spec[w["fg"]][w["row"]:w["row"] + c], im, pad, weight=w["wa"], out=og)
        return out


class SegmentedCG:
    """CG for (A_i + lam I) x_i = y_i on every segment of a flat buffer.

    ``normal(x, out)`` applies all A_i; ``seg_elems`` are the element counts
    of the segments (problems).  ``x``, ``r``, ``p``, ``Ap`` are flat
    buffers of the problem dtype; ``rs`` holds the per-problem |r|^2.
    """

    def __init__(self, normal: Callable, seg_elems: Sequence[int], device,
                 dtype: torch.dtype = torch.complex64, lam: float = 0.0,
                 block_floats: int = 16384):
        self.normal = normal
        self.lam = float(lam)
        self.device = torch.device(device)
        self.dtype = dtype
        comp = 2 if dtype.is_complex else 1
        self.nseg = len(seg_elems)
        numel = int(sum(seg_elems))
        seg_len = i64_array([int(e) * comp for e in seg_elems])
        self.layout = SegLayout()
        check(lib().tnl_seg_layout_query(seg_len, self.nseg, int(block_floats), self.layout))
        self.ws = torch.empty(int(self.layout.bytes), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(lib().tnl_seg_build(seg_len, self.layout, self.ws.data_ptr(),
                                      stream_of(self.ws)))
        mk = lambda: torch.zeros(numel, dtype=dtype, device=self.device)  # noqa: E731
        self.x, self.r, self.p, self.Ap = mk(), mk(), mk(), mk()
        dd = lambda: torch.zeros(self.nseg, dtype=torch.float64, device=self.device)  # noqa: E731
        self.rs, self.pAp, self.rs_new, self.rs0 = dd(), dd(), dd(), dd()
        self._graph = None
        self._graph_iters = 0

    def _f(self, t: torch.Tensor) -> int:
        return (torch.view_as_real(t) if t.is_complex() else t).data_ptr()

    def _dot(self, a, b, out):
        check(lib().tnl_seg_dot(self._f(a), self._f(b), out.data_ptr(), self.layout,
                                self.ws.data_ptr(), stream_of(out)))

    def _apply(self, x, out):
        self.normal(x, out)
        if self.lam:
            out.add_(x, alpha=self.lam)

    def reset(self, y: torch.Tensor, x0: Optional[torch.Tensor] = None) -> None:
        with torch.cuda.device(self.device):
            if x0 is None:
                self.x.zero_()
                self.r.copy_(y)
            else:
                self.x.copy_(x0)
                self._apply(self.x, self.Ap)
                torch.sub(y, self.Ap, out=self.r)
            self.p.copy_(self.r)
            self._dot(self.r, self.r, self.rs)
            self.rs0.copy_(self.rs)

    def step(self) -> None:
        """One CG iteration on every problem (no host synchronisation)."""
        with torch.cuda.device(self.device):
            self._apply(self.p, self.Ap)
            self._dot(self.p, self.Ap, self.pAp)
            s = stream_of(self.x)
            check(lib().tnl_seg_cg_update(self._f(self.x), self._f(self.r), self._f(self.p),
                                          self._f(self.Ap), self.rs.data_ptr(),
                                          self.pAp.data_ptr(), self.rs_new.data_ptr(),
                                          self.layout, self.ws.data_ptr(), s))
            check(lib().tnl_seg_cg_direction(self._f(self.p), self._f(self.r),
                                             self.rs.data_ptr(), self.rs_new.data_ptr(),
                                             self.layout, self.ws.data_ptr(), s))

    def capture(self, iters: int) -> None:
        """Record ``iters`` iterations in a graph (replayed by ``replay``).

        One eager application of the operator first creates every FFT plan
        and workspace outside the capture; the solver state is not touched.
        """
        scratch = torch.empty_like(self.Ap)
        self._apply(self.p, scratch)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):  # records only; nothing runs until replay
            for _ in range(iters):
                self.step()
        self._graph, self._graph_iters = g, iters

    def replay(self) -> None:
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph.replay()

    def solve(self, y: torch.Tensor, iters: int = 20, x0: Optional[torch.Tensor] = None,
              use_graph: bool = True) -> torch.Tensor:
        self.reset(y, x0)
        if use_graph:
            if self._graph is None or self._graph_iters != iters:
                self.capture(iters)
            self.replay()
        else:
            for _ in range(iters):
                self.step()
        return self.x

    def relative_residuals(self) -> torch.Tensor:
        """|r_i| / |r_i at reset| per problem (host tensor)."""
        return (self.rs / self.rs0.clamp_min(1e-300)).sqrt().cpu()
