"""Functional NUFFT and centred pad / crop (src/torchlinops/functional/
_nufft.py:16-183, _pad.py:10-82).

The reference rebuilds the apodization weights on every call; here they are
cached per (grid size, oversampling, width, device), which is the only
difference (results are identical).
"""

from __future__ import annotations

import threading
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


def nufft(x, locs, oversamp: float = 1.25, width: float = 4.0):
    """Forward NUFFT of ``x[..., *grid]`` at centred ``locs`` in [-N/2, N/2]."""
    from ..linops.basic import Pad
    from ..linops.nufft import cfftn

    grid_size = x.shape[-locs.shape[-1]:]
    p = init_nufft(grid_size, locs, oversamp, width, x.device)
    x = Pad.fn(p.pad_ns, x * p.apodize)
    x = cfftn(x, dim=p.dim, norm="ortho")
    x = interpolate(x, p.locs, width, kernel="kaiser_bessel", kernel_params=dict(beta=p.beta))
    return x / p.scale_factor


def nufft_adjoint(x, locs, grid_size, oversamp: float = 1.25, width: float = 4.0):
    from ..linops.basic import Pad
    from ..linops.nufft import cifftn

    p = init_nufft(grid_size, locs, oversamp, width, x.device)
    x = interpolate_adjoint(x / p.scale_factor, p.locs, p.padded_size, width,
                            kernel="kaiser_bessel", kernel_params=dict(beta=p.beta))
    x = cifftn(x, dim=p.dim, norm="ortho")
    return Pad.adj_fn(p.pad_ns, x) * p.apodize
