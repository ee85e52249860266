"""Toeplitz normal operator with the centring shifts folded into the pad.

The reference applies a NUFFT's Toeplitz normal as the chain
Crop . F^H . (K *) . F . Pad with centred FFTs (src/torchlinops/linops/
nufft.py:226-239; fft.py:84-131; utils/_fft.py:7-60): a zero-pad copy,
ifftshift / fftshift rolls around both FFTs and a crop copy.  Because
ifftshift(K * fftshift(v)) = ifftshift(K) * v, the same operator is

    crop_wrapped( ifftn( K_nat * fftn( pad_wrapped(x) ) ) )

with K_nat = ifftshift(K) / prod(pad_size) stored once (the orthonormal
scalings folded in, both FFTs unscaled), pad_wrapped = ifftshift . Pad and
crop_wrapped = Crop . fftshift each one HIP copy (tnl_pad_wrapped /
tnl_crop_wrapped), and the multiply done in place between the two rocFFT
calls.  Five full passes of the padded grid (pad, two rolls, multiply, two
rolls) become two partial copies and one multiply.
"""

from __future__ import annotations

from math import prod

import torch
import torch.nn as nn

from .._lib import check, i64_array, lib, require_device, stream_of
from ..nameddim import NamedShape as NS
from .namedlinop import NamedLinop

__all__ = ["pad_wrapped", "crop_wrapped", "ToeplitzNormal"]


def _wrap(pad: bool, x: torch.Tensor, im_size, pad_size, weight=None,
          out: torch.Tensor = None) -> torch.Tensor:
    require_device(x, op="pad_wrapped" if pad else "crop_wrapped")
    if weight is not None:
        if weight.dtype != torch.float32 or tuple(weight.shape) != tuple(im_size):
            raise ValueError(f"weight must be float32 of shape {tuple(im_size)}, got "
                             f"{weight.dtype} {tuple(weight.shape)}")
        if x.dtype not in (torch.float32, torch.complex64):
            raise TypeError(f"a weight needs float32 or complex64 data, got {x.dtype}")
        weight = weight.to(x.device).contiguous()
    nd = len(im_size)
    src_size = tuple(im_size) if pad else tuple(pad_size)
    if tuple(x.shape[x.ndim - nd:]) != src_size:
        raise ValueError(f"Mismatched shapes: expected {src_size} but got "
                         f"{tuple(x.shape[x.ndim - nd:])}")
    batch = tuple(x.shape[: x.ndim - nd])
    x = x.contiguous()
    shape = batch + (tuple(pad_size) if pad else tuple(im_size))
    if out is None:
        out = torch.empty(shape, dtype=x.dtype, device=x.device)
    elif (tuple(out.shape) != shape or out.dtype != x.dtype or out.device != x.device
          or not out.is_contiguous()):
        raise ValueError(f"out must be a contiguous {x.dtype} tensor of shape {shape} on "
                         f"{x.device}")
    if out.numel() == 0:
        return out
    fn = lib().tnl_pad_wrapped if pad else lib().tnl_crop_wrapped
    with torch.cuda.device(x.device):
        check(fn(x.data_ptr(), out.data_ptr(), x.element_size(), prod(batch), nd,
                 i64_array(im_size), i64_array(pad_size),
                 weight.data_ptr() if weight is not None else None, stream_of(x)))
    return out


def pad_wrapped(x: torch.Tensor, im_size, pad_size, weight=None, out=None) -> torch.Tensor:
    """``ifftshift(centred_pad(x * weight))`` over the last ``len(im_size)``
    axes (``weight``: optional real float32 ``[*im_size]``; ``out``:
    optional contiguous destination, fully written)."""
    return _wrap(True, x, im_size, pad_size, weight, out)


def crop_wrapped(y: torch.Tensor, im_size, pad_size, weight=None, out=None) -> torch.Tensor:
    """``centred_crop(fftshift(y)) * weight`` over the last ``len(im_size)``
    axes."""
    return _wrap(False, y, im_size, pad_size, weight, out)


class ToeplitzNormal(NamedLinop):
    """``Crop . F^H . (K *) . F . Pad`` applied as pad_wrapped -> fftn ->
    multiply by ifftshift(K) -> ifftn -> crop_wrapped.  ``kernel`` is the
    centred Fourier-domain PSF of ``toeplitz_psf``: a tensor
    ``[..., *pad_size]`` multiplied elementwise, or a ``Dense`` over (input
    batch, output batch, k-space) whose einsum couples the batch entries
    (then ``oshape`` names the output batch dims).  The adjoint uses the
    conjugate kernel."""

    def __init__(self, kernel, im_size, pad_size, ishape, name="ToeplitzNormal", oshape=None):
        from .basic import Dense

        oshape = tuple(ishape) if oshape is None else tuple(oshape)
        super().__init__(NS(tuple(ishape), oshape), name=name)
        self.im_size = tuple(int(n) for n in im_size)
        self.pad_size = tuple(int(n) for n in pad_size)
        dims = tuple(range(-len(self.im_size), 0))
        # the two orthonormal FFT scalings (1/sqrt(P) each) are folded into
        # K so that both transforms run unscaled (no extra pass each)
        if isinstance(kernel, Dense):
            w = (torch.fft.ifftshift(kernel.weight, dim=dims) / prod(self.pad_size)).contiguous()
            self.coupling = Dense(w, kernel.weightshape, kernel.ishape, kernel.oshape)
            self.kernel = self.coupling.weight
        else:
            self.coupling = None
            self.kernel = nn.Parameter(
                (torch.fft.ifftshift(kernel, dim=dims) / prod(self.pad_size)).contiguous(),
                requires_grad=False)

    def _multiply(self, y, conj, inplace):
        if self.coupling is not None:
            from .basic import Dense

            return Dense.adj_fn(self.coupling, y) if conj else Dense.fn(self.coupling, y)
        k = self.kernel.to(y.device)
        k = k.conj() if conj else k
        return y.mul_(k) if inplace else y * k

    def _toeplitz(self, x, conj):
        nd = len(self.im_size)
        dims = tuple(range(-nd, 0))
        if not x.is_cuda or (torch.is_grad_enabled() and x.requires_grad):
            return self._toeplitz_torch(x, conj)
        y = pad_wrapped(x, self.im_size, self.pad_size)
        y = torch.fft.fftn(y, dim=dims, norm="backward")  # unscaled
        y = self._multiply(y, conj, inplace=True)
        y = torch.fft.ifftn(y, dim=dims, norm="forward")  # unscaled
        return crop_wrapped(y, self.im_size, self.pad_size)

    def _toeplitz_torch(self, x, conj):
        """The same operator in differentiable torch ops (autograd, host
        tensors): centred zero pad, ifftshift, unscaled FFTs around the
        multiply, fftshift, centred crop."""
        from .basic import crop_slice_from_pad, pad_to_size

        nd = len(self.im_size)
        dims = tuple(range(-nd, 0))
        pad = pad_to_size(self.im_size, self.pad_size)
        y = torch.nn.functional.pad(x, pad)
        y = torch.fft.ifftshift(y, dim=dims)
        y = torch.fft.fftn(y, dim=dims, norm="backward")
        y = self._multiply(y, conj, inplace=False)
        y = torch.fft.ifftn(y, dim=dims, norm="forward")
        y = torch.fft.fftshift(y, dim=dims)
        return y[(Ellipsis, *crop_slice_from_pad(pad))]

    @staticmethod
    def fn(t, x, /):
        return t._toeplitz(x, False)

    @staticmethod
    def adj_fn(t, x, /):
        return t._toeplitz(x, True)

    def size(self, dim):
        shape = list(self.ishape)
        if dim in shape[len(shape) - len(self.im_size):]:
            return self.im_size[shape.index(dim) - len(shape) + len(self.im_size)]
        return None
