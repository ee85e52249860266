"""NUFFT linop (type II) and its Toeplitz normal.

Follows src/torchlinops/linops/nufft.py:
  * "interpolate" mode: chain Apodize -> Pad(int(N*os)) -> FFT(centred,
    ortho) -> Interpolate(KB, beta) -> Scalar(1 / (w^D sqrt(prod N /
    prod N_pad))) (nufft.py:123-184);
  * "sampling" mode: Pad -> FFT -> Sampling at integer padded-grid
    locations, no apodization or scaling (185-203); applied fused as a
    scaled wrapped pad, an uncentred FFT and the gather at relabelled
    cells;
  * prep_locs (241-334), beta (336-346), apodize_weights (348-365);
  * adjoint (217-224); normal with ``toeplitz=True`` (226-239) builds
    Crop . F^H . (K *) . F . Pad with K the Fourier transform of the point
    spread function on a 2x grid (toeplitz_psf, 395-494), applied as one
    ToeplitzNormal (shift-free pad / crop copies around rocFFT; option
    ``toeplitz_fused=False`` returns the reference's chain).
The interpolation runs on the HIP kernels through Interpolate.
"""

from __future__ import annotations

from copy import copy
from itertools import product
from math import prod
from typing import Optional

import torch
import torch.nn as nn

from ..nameddim import ND, NamedShape as NS, get_nd_shape
from .basic import (FFT, Dense, Diagonal, Identity, Pad, Scalar, crop_slice_from_pad,
                    pad_to_size)
from .chain import Chain
from .interp import Interpolate
from .sampling import Sampling
from .toeplitz import ToeplitzNormal

__all__ = ["NUFFT", "toeplitz_psf", "cfftn", "cifftn"]


def cfftn(x, dim=None, norm="ortho"):
    """Centred n-D FFT: ifftshift -> fftn -> fftshift (utils/_fft.py:7-31)."""
    x = torch.fft.ifftshift(x, dim=dim)
    x = torch.fft.fftn(x, dim=dim, norm=norm)
    return torch.fft.fftshift(x, dim=dim)


def cifftn(x, dim=None, norm="ortho"):
    x = torch.fft.ifftshift(x, dim=dim)
    x = torch.fft.ifftn(x, dim=dim, norm=norm)
    return torch.fft.fftshift(x, dim=dim)


def shifted_layout(im_size, padded, device):
    """Index tensors and location shift that replace the centring shifts of
    cfftn / cifftn inside a NUFFT normal operator.

    For an axis of n image cells centred on a padded axis of N cells
    (crop slice [c0, c0 + n)), fftshift(fftn(ifftshift(pad(x)))) equals
    fftn of the image written at the wrapped cells (i - N // 2) mod N,
    relabelled by N // 2; under circular padding, interpolating the
    relabelled grid at t equals interpolating the unshifted one at
    (t - N // 2) mod N, and gridding likewise.  The adjoint crops from the
    same wrapped cells.  Returns (index, shift): a tuple of broadcastable
    index tensors (one per axis) for ``g[..., *index]`` and the per-axis
    shift N // 2 to subtract from the locations (modulo N)."""
    ndim = len(im_size)
    crop = crop_slice_from_pad(pad_to_size(tuple(im_size), tuple(padded)))
    index = []
    for d, (sl, N) in enumerate(zip(crop, padded)):
        start, stop, _ = sl.indices(N)
        i = (torch.arange(start, stop, device=device) - N // 2) % N
        shape = [1] * ndim
        shape[d] = -1
        index.append(i.view(shape))
    shift = torch.tensor([N // 2 for N in padded], dtype=torch.float32, device=device)
    return tuple(index), shift


class NUFFT(Chain):
    def __init__(self, locs, grid_size, output_shape, input_shape=None, input_kshape=None,
                 batch_shape=None, oversamp: float = 1.25, width: float = 4.0,
                 mode: str = "interpolate", do_prep_locs: bool = True,
                 apodize_weights: Optional[torch.Tensor] = None, **options):
        if mode not in ("interpolate", "sampling"):
            raise ValueError(f"Unrecognized NUFFT mode: {mode}")
        device = locs.device
        self.mode = mode
        self.options = options
        self.input_shape = ND.infer(tuple(input_shape) if input_shape is not None
                                    else get_nd_shape(len(grid_size)))
        self.input_kshape = ND.infer(tuple(input_kshape) if input_kshape is not None
                                     else get_nd_shape(len(grid_size), kspace=True))
        self.output_shape = ND.infer(tuple(output_shape))
        self.batch_shape = ND.infer(tuple(batch_shape) if batch_shape is not None else ("...",))
        batched_input = NS(self.batch_shape) + NS(self.input_shape)
        ndim = len(grid_size)
        padded_size = tuple(int(i * oversamp) for i in grid_size)
        pad = Pad(padded_size, grid_size, in_shape=self.input_shape, batch_shape=self.batch_shape)
        fft = FFT(ndim=locs.shape[-1], centered=True, norm="ortho", batch_shape=self.batch_shape,
                  grid_shapes=(pad.out_im_shape, self.input_kshape))
        grid_shape = fft._shape.output_grid_shape
        locs_p = (self.prep_locs(locs, grid_size, padded_size, nufft_mode=mode)
                  if do_prep_locs else locs)
        if mode == "sampling":
            if locs_p.is_complex() or locs_p.is_floating_point():
                raise ValueError(
                    f"Sampling linop requries integer-type locs but got {locs_p.dtype}"
                )
            interp = Sampling.from_stacked_idx(locs_p, dim=-1, input_size=padded_size,
                                               output_shape=self.output_shape,
                                               input_shape=grid_shape,
                                               batch_shape=self.batch_shape)
            super().__init__(pad, fft, interp, name="NUFFT")
            self._finish(locs, grid_size, oversamp, width, pad, fft, interp)
            if options.get("fused", True):
                # Pad . centred FFT . Sampling as a wrapped pad scaled by the
                # orthonormal 1/sqrt(prod N_pad), an unscaled FFT and the
                # gather at cells relabelled by -N_pad/2 (exact for cells)
                fw = torch.full(tuple(grid_size), prod(padded_size) ** -0.5,
                                dtype=torch.float32, device=device)
                self.register_buffer("_fused_weight", fw, persistent=False)
                self._fused_map = None
                self._fused_src = None
                self._fused_on = True
            return
        beta = self.beta(width, oversamp)
        weight = (self.apodize_weights(grid_size, padded_size, oversamp, width, beta).to(device)
                  if apodize_weights is None else apodize_weights)
        if weight.isnan().any() or weight.isinf().any():
            raise ValueError(
                f"Nan/Inf values detected in apodization weight (width={width}, oversamp={oversamp})."
            )
        apodize = Diagonal(weight, batched_input.ishape)
        apodize.name = "Apodize"
        interp = Interpolate(locs_p, padded_size, batch_shape=self.batch_shape,
                             locs_batch_shape=self.output_shape, grid_shape=grid_shape,
                             width=width, kernel="kaiser_bessel", kernel_params=dict(beta=beta))
        scale_factor = width**ndim * (prod(grid_size) / prod(padded_size)) ** 0.5
        scale = Scalar(weight=1.0 / scale_factor, ioshape=interp.oshape)
        super().__init__(apodize, pad, fft, interp, scale, name="NUFFT")
        self._finish(locs, grid_size, oversamp, width, pad, fft, interp)
        if (tuple(weight.shape) == tuple(grid_size) and not weight.is_complex()
                and options.get("fused", True)):
            # Apodize . Pad . centred FFT . Interp . Scalar as one weighted
            # wrapped pad, an uncentred FFT and interpolation at locations
            # relabelled by -N_pad/2 (exact under circular padding)
            # apodization, 1/scale and the orthonormal FFT's 1/sqrt(prod
            # N_pad) in one real weight; the FFTs then run unscaled
            fw = weight.float() / (scale_factor * prod(padded_size) ** 0.5)
            self.register_buffer("_fused_weight", fw.contiguous(), persistent=False)
            self.register_buffer("_fused_locs", self._relabel(self.interp.locs),
                                 persistent=False)
            self._fused_src = self._locs_key()
            self._fused_on = True
        self._fused_dir = 1

    _fused_on = False
    _fused_dir = 1

    def _relabel(self, locs):
        """Prepared locations moved by -N_pad/2 (mod N_pad): the k-space grid
        in natural FFT order instead of centred."""
        padded = tuple(self.pad.pad_im_size)
        shift = torch.tensor([n // 2 for n in padded], dtype=torch.float32, device=locs.device)
        size = torch.tensor(padded, dtype=torch.float32, device=locs.device)
        return torch.remainder(locs.detach().float() - shift, size).contiguous()

    def _locs_key(self):
        if isinstance(self.interp, Sampling):
            return tuple((i.data_ptr(), i._version, str(i.device)) for i in self.interp.idx)
        l = self.interp.locs
        return (l.data_ptr(), l._version, str(l.device))

    def _sampling_map(self):
        """Index map of the sampled cells relabelled by -N_pad/2 (mod N_pad)."""
        from ..functional import _index as FI

        padded = tuple(self.pad.pad_im_size)
        idx = tuple(torch.remainder(i - n // 2, n) for i, n in zip(self.interp.idx, padded))
        return FI.index_map(idx, padded, idx[0].device)

    def _fused_ok(self, x) -> bool:
        if not (self._fused_on and x.is_cuda and x.dtype == torch.complex64):
            return False
        if not torch.is_grad_enabled():
            return True
        # the fused route relabels a detached copy of the trajectory: with a
        # trainable trajectory (or input) the member chain keeps autograd
        locs = getattr(self.interp, "locs", None)
        return not (x.requires_grad or (locs is not None and locs.requires_grad))

    def _fused(self, x, forward: bool):
        from .toeplitz import crop_wrapped, pad_wrapped
        from .. import functional as F

        grid, padded = self.grid_size, tuple(self.pad.pad_im_size)
        dims = tuple(range(-len(grid), 0))
        if isinstance(self.interp, Sampling):
            from ..functional import _index as FI

            if self._locs_key() != self._fused_src:  # built, edited or moved
                self._fused_map = self._sampling_map()
                self._fused_src = self._locs_key()
            if forward:
                y = pad_wrapped(x, grid, padded, weight=self._fused_weight)
                y = torch.fft.fftn(y, dim=dims, norm="backward")  # unscaled
                return FI.index_with_map(y, self._fused_map)
            g = FI.index_adjoint_with_map(x, self._fused_map)
            g = torch.fft.ifftn(g, dim=dims, norm="forward")  # unscaled
            return crop_wrapped(g, grid, padded, weight=self._fused_weight)
        p = self.interp.interp_params
        if self._locs_key() != self._fused_src:  # the trajectory was edited or moved
            self._fused_locs = self._relabel(self.interp.locs)
            self._fused_src = self._locs_key()
        if forward:
            y = pad_wrapped(x, grid, padded, weight=self._fused_weight)
            y = torch.fft.fftn(y, dim=dims, norm="backward")  # unscaled
            return F.interpolate(y, self._fused_locs, **p)
        g = F.interpolate_adjoint(x, self._fused_locs, padded, **p)
        g = torch.fft.ifftn(g, dim=dims, norm="forward")  # unscaled
        return crop_wrapped(g, grid, padded, weight=self._fused_weight)

    @staticmethod
    def fn(nufft, x, /):
        if nufft._fused_ok(x):
            return nufft._fused(x, nufft._fused_dir > 0)
        return Chain.fn(nufft, x)

    @staticmethod
    def adj_fn(nufft, x, /):
        if nufft._fused_ok(x):
            return nufft._fused(x, nufft._fused_dir < 0)
        return Chain.adj_fn(nufft, x)

    @staticmethod
    def split(nufft, tile):
        out = Chain.split(nufft, tile)
        out._fused_on = False
        return out

    def _finish(self, locs, grid_size, oversamp, width, pad, fft, interp):
        self.locs = locs
        self.grid_size = tuple(grid_size)
        self.oversamp = oversamp
        self.width = width
        self.pad = pad
        self.fft = fft
        self.interp = interp

    def adjoint(self):
        adj = copy(self)
        adj._shape = adj._shape.H
        adj.linops = nn.ModuleList([l.adjoint() for l in reversed(self.linops)])
        adj._fused_dir = -self._fused_dir
        return adj

    def normal(self, inner=None):
        if self.options.get("toeplitz", False):
            os_ = self.options.get("toeplitz_oversamp", 2.0)
            kernel = toeplitz_psf(self, inner, dtype=self.options.get("toeplitz_dtype"),
                                  oversamp=os_)
            pad = Pad(scale_int(self.grid_size, os_), self.grid_size,
                      in_shape=self.input_shape, batch_shape=self.batch_shape)
            if not self.options.get("toeplitz_fused", True):
                return pad.normal(self.fft.normal(kernel))
            padded = scale_int(self.grid_size, os_)
            if isinstance(kernel, Dense):
                # batch-coupling inner: image dims after the kernel's batch dims
                nd = len(self.grid_size)
                im_dims = tuple(pad.ishape[len(pad.ishape) - nd:])
                bi = tuple(kernel.ishape[: len(kernel.ishape) - nd])
                bo = tuple(kernel.oshape[: len(kernel.oshape) - nd])
                return ToeplitzNormal(kernel, self.grid_size, padded, bi + im_dims,
                                      oshape=bo + im_dims)
            return ToeplitzNormal(kernel.weight, self.grid_size, padded, pad.ishape)
        normal = super().normal(inner)
        if inner is None and self._fused_on and self._fused_dir > 0:
            nufft = self

            def fused_normal(chain, x, /):
                if nufft._fused_ok(x):
                    return nufft._fused(nufft._fused(x, True), False)
                return Chain.fn(chain, x)

            normal.fn = fused_normal
            normal.adj_fn = fused_normal
        return normal

    @staticmethod
    def prep_locs(locs, grid_size, padded_size, pad_mode="circular", nufft_mode="interpolate"):
        """Centred coordinates in [-N/2, N/2] -> padded-grid units in
        [0, N_pad) (scale by N_pad/N, shift by N_pad//2, wrap or clamp).
        ``nufft_mode="sampling"`` rounds to the nearest cell before the
        wrap and returns int64."""
        out = locs.clone()
        for i in range(-len(grid_size), 0):
            out[..., i] *= padded_size[i] / grid_size[i]
            out[..., i] += padded_size[i] // 2
            if pad_mode == "zero":
                out[..., i] = torch.clamp(out[..., i], 0, padded_size[i] - 1)
            elif pad_mode == "circular":
                if nufft_mode == "sampling":
                    out[..., i] = torch.remainder(torch.round(out[..., i]), padded_size[i])
                else:
                    out[..., i] = torch.remainder(out[..., i], torch.tensor(padded_size[i]))
            else:
                raise ValueError(f"Unrecognized padding mode during prep: {pad_mode}")
        if nufft_mode == "sampling":
            out = out.to(torch.int64)
        return out

    @staticmethod
    def beta(width, oversamp):
        """Beatty et al. 2005: pi * sqrt(((w / os) (os - 1/2))^2 - 0.8)."""
        return torch.pi * (((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8) ** 0.5

    @staticmethod
    def apodize_weights(grid_size, padded_size, oversamp, width, beta):
        """prod_d a_d / sinh(a_d), a = sqrt(beta^2 - (pi w (n - N//2) / N_pad)^2)."""
        gs = torch.tensor(grid_size)
        ps = torch.tensor(padded_size)
        g = torch.stack(torch.meshgrid(*(torch.arange(s) for s in grid_size), indexing="ij"), -1)
        apod = (beta**2 - (torch.pi * width * (g - gs // 2) / ps) ** 2) ** 0.5
        apod = apod / torch.sinh(apod)
        return torch.prod(apod, dim=-1)

    def flatten(self):
        return [self]

    @property
    def device(self):
        if isinstance(self.interp, Sampling):
            return self.interp.idx[0].device
        return self.interp.locs.device


def scale_int(t, s):
    return tuple(int(s * x) for x in t)


def _rescale_locs(locs, c0, w0, c1, w1):
    cols = [(locs[..., d] - c0[d]) * w1[d] / w0[d] + c1[d] for d in range(locs.shape[-1])]
    return torch.stack(cols, dim=-1)


def psf_sizing(nufft: NUFFT, inner, toeplitz_oversamp: float = 2.0):
    """Shapes and sizes of a Toeplitz kernel around ``inner`` (nufft.py:501-536).

    Returns ``(weightshape, ishape, oshape, kernel_size, input_size,
    batch_sizes)``: the inner's batch dims in and out (its shape minus the
    NUFFT's output dims) followed by the k-space grid dims; the kernel holds
    one PSF per (input batch index, output batch index) pair."""
    nout = len(nufft.output_shape)
    bi = tuple(inner.ishape[: len(inner.ishape) - nout])
    bo = tuple(inner.oshape[: len(inner.oshape) - nout])
    kshape = tuple(nufft.fft._shape.output_grid_shape)
    if tuple(str(d) for d in bi) == ("...",):
        weightshape = bi + kshape
    elif "..." in [str(d) for d in bi] and "..." in [str(d) for d in bo]:
        raise ValueError(
            f"Underspecified kernel shape for toeplitz embedding with inner.shape = "
            f"{inner.shape}. Specify more dimensions of inner to avoid this.")
    else:
        weightshape = bi + bo + kshape
    sizes = tuple(inner.size(d) for d in bi)
    sizes = tuple(1 if n is None else int(n) for n in sizes)
    kernel_size = sizes + sizes + scale_int(nufft.pad.im_size, toeplitz_oversamp)
    input_size = sizes + tuple(nufft.interp.locs.shape[:-1])
    return weightshape, bi + kshape, bo + kshape, kernel_size, input_size, sizes


def all_indices(sizes):
    """Every index tuple of an array of shape ``sizes``, lexicographic."""
    return product(*(range(n) for n in sizes))


def toeplitz_psf(nufft: NUFFT, inner=None, dtype=None, oversamp: float = 2.0):
    """Fourier-domain point spread kernel K with A^H inner A =
    Crop F^H (K *) F Pad on an ``oversamp``-times larger grid
    (nufft.py:395-494).

    Each PSF is one adjoint NUFFT of ``inner`` applied to a unit batch
    entry, on the enlarged grid with rescaled locations.  An inner that
    leaves batch entries apart (None, Identity, a Diagonal broadcast over
    the batch) needs one PSF, shared by every batch entry: the kernel is a
    Diagonal.  An inner that couples batch entries (a Dense over a batch
    axis) needs one PSF per input batch index, each holding the response
    at every output batch index: the kernel is a Dense over (input batch,
    output batch, k-space) whose einsum does the coupling.
    """
    if isinstance(nufft.interp, Sampling):
        raise NotImplementedError(
            "Toeplitz embedding not yet implemented for Sampling-type NUFFT"
        )
    dtype = dtype or torch.complex64
    device = nufft.device
    grid_size = nufft.pad.im_size
    new_grid = scale_int(grid_size, oversamp)
    width0 = nufft.pad.pad_im_size
    width1 = scale_int(width0, oversamp)
    new_locs = _rescale_locs(nufft.interp.locs.detach().clone(),
                             tuple(w // 2 for w in width0), width0,
                             tuple(w // 2 for w in width1), width1)
    nufft_os = NUFFT(new_locs, grid_size=new_grid, output_shape=nufft.output_shape,
                     input_shape=nufft.input_shape, input_kshape=nufft.input_kshape,
                     batch_shape=("...",), oversamp=nufft.oversamp, width=nufft.width,
                     do_prep_locs=False)
    dims = tuple(range(-len(new_grid), 0))
    scale = oversamp ** len(grid_size) / prod(new_grid) ** 0.5
    kshape = tuple(nufft.fft._shape.output_grid_shape)

    def psf(v):
        return cfftn(nufft_os.H(v), dim=dims, norm=None) * scale

    if inner is None or isinstance(inner, Identity):
        kernel = psf(torch.ones(tuple(nufft.interp.locs.shape[:-1]), dtype=dtype,
                                device=device)).to(dtype)
        return Diagonal(kernel, ioshape=tuple(nufft.batch_shape) + kshape)
    if len(inner.ishape) != len(inner.oshape):
        raise ValueError(
            f"Inner linop must have identical input and output shape lengths but got "
            f"ishape={inner.ishape} and oshape={inner.oshape}")
    weightshape, ishape, oshape, kernel_size, input_size, sizes = psf_sizing(
        nufft, inner, oversamp)
    if prod(sizes) == 1:
        # no batch coupling: one PSF of inner(ones), shared by every entry
        ones = torch.ones(tuple(nufft.interp.locs.shape[:-1]), dtype=dtype, device=device)
        kernel = psf(inner(ones)).to(dtype)
        return Diagonal(kernel, ioshape=tuple(nufft.batch_shape) + kshape)
    kernel = torch.zeros(kernel_size, dtype=dtype, device=device)
    unit = torch.zeros(input_size, dtype=dtype, device=device)
    for b in all_indices(sizes):
        unit[b] = 1.0
        kernel[b] = psf(inner(unit)).to(dtype)
        unit[b] = 0.0
    return Dense(kernel, weightshape=weightshape, ishape=ishape, oshape=oshape)
