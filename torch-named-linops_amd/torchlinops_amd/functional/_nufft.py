"""Functional NUFFT and centred pad / crop (src/torchlinops/functional/
_nufft.py:16-183, _pad.py:10-82).

The reference rebuilds the apodization weights on every call; here they are
cached per (grid size, oversampling, width, device).  Without autograd,
complex64 calls run Apodize . Pad . centred FFT as one weighted wrapped pad
(apodization and 1/scale folded in) and an uncentred FFT, and interpolate
at locations relabelled by -N_pad/2 (exact under circular padding); the
adjoint mirrors it.  Results agree with the chain to float32 rounding.
"""

from __future__ import annotations

import threading
import weakref
from collections import OrderedDict
from math import prod
from types import SimpleNamespace

import torch

from ._interp import interpolate, interpolate_adjoint

__all__ = ["nufft", "nufft_adjoint", "center_pad", "center_crop", "init_nufft"]

_apod_cache: dict = {}
_apod_lock = threading.Lock()


def _pad_ns(im_size, pad_im_size):
    from ..linops.basic import crop_slice_from_pad, pad_to_size

    pad = pad_to_size(im_size, pad_im_size)
    return SimpleNamespace(im_size=tuple(im_size), pad_im_size=tuple(pad_im_size),
                           D=len(im_size), pad=pad, crop_slice=crop_slice_from_pad(pad))


def center_pad(x, im_size, pad_im_size):
    from ..linops.basic import Pad

    return Pad.fn(_pad_ns(im_size, pad_im_size), x)


def center_crop(x, im_size, crop_im_size):
    from ..linops.basic import Pad

    return Pad.adj_fn(_pad_ns(crop_im_size, im_size), x)


def init_nufft(grid_size, locs, oversamp, width, device):
    from ..linops.nufft import NUFFT

    ndim = locs.shape[-1]
    grid_size = tuple(int(g) for g in grid_size)
    padded = tuple(int(s * oversamp) for s in grid_size)
    beta = NUFFT.beta(width, oversamp)
    key = (grid_size, float(oversamp), float(width), str(device))
    with _apod_lock:
        apod = _apod_cache.get(key)
    if apod is None:
        apod = NUFFT.apodize_weights(grid_size, padded, oversamp, width, beta).to(device)
        with _apod_lock:
            _apod_cache[key] = apod
    return SimpleNamespace(
        ndim=ndim, dim=tuple(range(-ndim, 0)), grid_size=grid_size, padded_size=padded,
        locs=NUFFT.prep_locs(locs, grid_size, padded), beta=beta, apodize=apod,
        pad_ns=_pad_ns(grid_size, padded),
        scale_factor=width**ndim * (prod(grid_size) / prod(padded)) ** 0.5,
    )


def _fusable(x) -> bool:
    return x.is_cuda and x.dtype == torch.complex64 and not (
        torch.is_grad_enabled() and x.requires_grad)


_fused_cache: OrderedDict = OrderedDict()


def _fused_parts(locs, grid_size, oversamp, width, device):
    """Folded weight apod / scale and the relabelled locations, cached per
    trajectory (identity and version of ``locs``), so repeated calls reuse
    them and the interpolation plan built for them."""
    key = (id(locs), locs._version, locs.data_ptr(), tuple(locs.shape),
           tuple(int(g) for g in grid_size), float(oversamp), float(width), str(device))
    with _apod_lock:
        hit = _fused_cache.get(key)
        if hit is not None and hit[0]() is locs:
            _fused_cache.move_to_end(key)
            return hit[1]
    p = init_nufft(grid_size, locs, oversamp, width, device)
    shift = torch.tensor([n // 2 for n in p.padded_size], dtype=torch.float32, device=device)
    size = torch.tensor(p.padded_size, dtype=torch.float32, device=device)
    # apodization, 1/scale and the orthonormal 1/sqrt(prod N_pad) in one
    # weight: the FFTs below run unscaled
    w = (p.apodize.float() / (p.scale_factor * prod(p.padded_size) ** 0.5)).contiguous()
    lw = torch.remainder(p.locs.to(device).float() - shift, size).contiguous()
    parts = (p, w, lw)
    with _apod_lock:
        _fused_cache[key] = (weakref.ref(locs), parts)
        while len(_fused_cache) > 8:
            _fused_cache.popitem(last=False)
    return parts


def nufft(x, locs, oversamp: float = 1.25, width: float = 4.0):
    """Forward NUFFT of ``x[..., *grid]`` at centred ``locs`` in [-N/2, N/2]."""
    from ..linops.basic import Pad
    from ..linops.nufft import cfftn
    from ..linops.toeplitz import pad_wrapped

    grid_size = x.shape[-locs.shape[-1]:]
    if _fusable(x):
        p, w, lw = _fused_parts(locs, grid_size, oversamp, width, x.device)
        y = torch.fft.fftn(pad_wrapped(x, p.grid_size, p.padded_size, weight=w),
                           dim=p.dim, norm="backward")
        return interpolate(y, lw, width, kernel="kaiser_bessel",
                           kernel_params=dict(beta=p.beta))
    p = init_nufft(grid_size, locs, oversamp, width, x.device)
    x = Pad.fn(p.pad_ns, x * p.apodize)
    x = cfftn(x, dim=p.dim, norm="ortho")
    x = interpolate(x, p.locs, width, kernel="kaiser_bessel", kernel_params=dict(beta=p.beta))
    return x / p.scale_factor


def nufft_adjoint(x, locs, grid_size, oversamp: float = 1.25, width: float = 4.0):
    from ..linops.basic import Pad
    from ..linops.nufft import cifftn
    from ..linops.toeplitz import crop_wrapped

    if _fusable(x):
        p, w, lw = _fused_parts(locs, grid_size, oversamp, width, x.device)
        g = interpolate_adjoint(x, lw, p.padded_size, width, kernel="kaiser_bessel",
                                kernel_params=dict(beta=p.beta))
        g = torch.fft.ifftn(g, dim=p.dim, norm="forward")
        return crop_wrapped(g, p.grid_size, p.padded_size, weight=w)
    p = init_nufft(grid_size, locs, oversamp, width, x.device)
    x = interpolate_adjoint(x / p.scale_factor, p.locs, p.padded_size, width,
                            kernel="kaiser_bessel", kernel_params=dict(beta=p.beta))
    x = cifftn(x, dim=p.dim, norm="ortho")
    return Pad.adj_fn(p.pad_ns, x) * p.apodize
