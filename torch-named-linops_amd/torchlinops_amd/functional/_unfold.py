"""array_to_blocks (unfold) / blocks_to_array (fold) and their autograd pair.

Signatures, defaults, shapes and errors follow the reference:
  * ``unfold``  src/torchlinops/functional/_unfold/unfold.py:31-131, 564-614
  * ``fold``    src/torchlinops/functional/_unfold/fold.py:32-118, 520-575
  * autograd / norm weights  array_to_blocks.py:16-140
  * ``get_nblocks``          nblocks.py:6-35
The copy / gather runs in libtnl (tnl_unfold / tnl_fold) for tensors on a
HIP device; host tensors take the package's torch restatement (``_host.py``),
the reference's own dispatch on ``is_cuda`` (unfold.py:94, fold.py:89).
"""

from __future__ import annotations

from math import prod
from typing import Optional

import torch
from torch.autograd import Function

from .._lib import TNL_F32, TNL_F64, check, i64_array, lib, stream_of
from ._host import fold_host, unfold_host

__all__ = [
    "get_nblocks",
    "unfold",
    "fold",
    "array_to_blocks",
    "blocks_to_array",
    "get_norm_weights",
]


def get_nblocks(im_size, block_size, block_stride=None) -> tuple:
    """Number of whole blocks per axis: ``(im - block) // stride + 1``.

    >>> get_nblocks((5, 5), (3, 3), (2, 2))
    (2, 2)
    >>> get_nblocks((10, 10), (8, 8), (4, 4))
    (1, 1)
    """
    assert len(im_size) == len(block_size), (
        f"im_size {im_size} and block_size {block_size} don't match"
    )
    block_stride = block_stride if block_stride is not None else (1,) * len(block_size)
    return tuple((im - bl) // st + 1 for im, bl, st in zip(im_size, block_size, block_stride))


def _interleaved(shape: tuple, is_complex: bool) -> list:
    shape = [int(s) for s in shape]
    if is_complex:
        shape[-1] *= 2
    return shape


def _elem_view(x: torch.Tensor) -> torch.Tensor:
    """Real view with complex (re, im) interleaved into the last axis."""
    if torch.is_complex(x):
        return torch.view_as_real(x).flatten(-2, -1)
    return x


def unfold(
    x: torch.Tensor,
    block_size: tuple,
    stride: Optional[tuple] = None,
    mask: Optional[torch.Tensor] = None,
    output: Optional[torch.Tensor] = None,
) -> torch.Tensor:
    """Extract (possibly overlapping) blocks: ``[B..., *im] -> [B..., *nblocks, *block]``.

    With ``mask``, the block axes are replaced by ``mask``'s True entries.
    ``output`` (same dtype as ``x``) is filled in place and returned.
    """
    ndim = len(block_size)
    block_size = tuple(int(b) for b in block_size)
    im_size = tuple(x.shape[-ndim:])
    stride = tuple(int(s) for s in stride) if stride is not None else (1,) * ndim
    nblocks = get_nblocks(im_size, block_size, stride)
    if any(b <= 0 for b in nblocks):
        raise ValueError(
            f"Found 0 in nblocks: {nblocks} with im_size {im_size}, block_size "
            f"{block_size} and stride {stride} - make sure there is at least one "
            "block in each direction."
        )
    batch_shape = x.shape[:-ndim]
    nbatch = prod(batch_shape)
    if output is not None and output.dtype != x.dtype:
        raise ValueError(
            f"Output and input dtypes must match but got output {output.dtype} != "
            f"input {x.dtype}"
        )
    if mask is not None:
        mask = mask.to(x.device)
    full_shape = (*batch_shape, *nblocks, *block_size)
    if output is None or mask is not None:
        y = torch.empty(full_shape, dtype=x.dtype, device=x.device)
    else:
        if output.numel() != prod(full_shape):
            raise RuntimeError(
                f"Output buffer of {output.numel()} elements cannot hold {full_shape}"
            )
        y = output.view(full_shape)
    is_complex = torch.is_complex(x)
    xr = _elem_view(x.reshape(nbatch, *im_size).contiguous())
    yr = _elem_view(y)
    if not x.is_cuda:
        # host tensors: the package's torch restatement (strided copy)
        if nbatch > 0:
            y.view(nbatch, *nblocks, *block_size).copy_(
                unfold_host(x.reshape(nbatch, *im_size), block_size, stride))
    elif nbatch > 0:
        with torch.cuda.device(x.device):
            check(
                lib().tnl_unfold(
                    xr.data_ptr(),
                    yr.data_ptr(),
                    xr.element_size(),
                    nbatch,
                    ndim,
                    i64_array(nblocks),
                    i64_array(_interleaved(block_size, is_complex)),
                    i64_array(_interleaved(im_size, is_complex)),
                    i64_array(_interleaved(stride, is_complex)),
                    stream_of(x),
                )
            )
    if mask is not None:
        y = y[..., mask]
        if output is not None:
            output.copy_(y.reshape(output.shape))
            return output
        return y
    return y


def fold(
    x: torch.Tensor,
    im_size: tuple,
    block_size: tuple,
    stride: Optional[tuple] = None,
    mask: Optional[torch.Tensor] = None,
    output: Optional[torch.Tensor] = None,
) -> torch.Tensor:
    """Sum (possibly overlapping) blocks back into an array:
    ``[B..., *nblocks, *block] -> [B..., *im_size]``.

    The sum at each cell runs over covering blocks in lexicographic block
    order, matching the reference's sequential accumulation bit for bit.
    """
    ndim = len(block_size)
    block_size = tuple(int(b) for b in block_size)
    im_size = tuple(int(i) for i in im_size)
    stride = tuple(int(s) for s in stride) if stride is not None else (1,) * ndim
    nblocks = get_nblocks(im_size, block_size, stride)
    if mask is not None:
        if x.ndim > len(nblocks) + 1:
            x_batch_shape = x.shape[: -(len(nblocks) + 1)]
        else:
            x_batch_shape = ()
        tmp = torch.zeros(*x_batch_shape, *nblocks, *block_size, dtype=x.dtype,
                          device=x.device)
        tmp[..., mask.to(x.device)] = x
        x = tmp
    if any(b <= 0 for b in nblocks):
        raise ValueError(
            f"Found 0 in nblocks: {nblocks} with im_size {im_size}, block_size "
            f"{block_size} and stride {stride} - make sure there is at least one "
            "block in each direction."
        )
    if output is not None and output.dtype != x.dtype:
        raise ValueError(
            f"Output and input dtypes must match but got output {output.dtype} != "
            f"input {x.dtype}"
        )
    if tuple(x.shape[-2 * ndim:]) != (*nblocks, *block_size):
        raise RuntimeError(
            f"Fold expected input with full size {(*nblocks, *block_size)} but got {x.shape}"
        )
    is_complex = torch.is_complex(x)
    real_dtype = x.real.dtype if is_complex else x.dtype
    if real_dtype == torch.float32:
        code = TNL_F32
    elif real_dtype == torch.float64:
        code = TNL_F64
    else:
        raise TypeError(f"fold supports float32/float64 (and complex); got {x.dtype}")
    batch_shape = x.shape[: x.ndim - 2 * ndim]
    nbatch = prod(batch_shape)
    full_shape = (*batch_shape, *im_size)
    if output is None:
        y = torch.empty(full_shape, dtype=x.dtype, device=x.device)
    else:
        if output.numel() != prod(full_shape):
            raise RuntimeError(
                f"Output buffer of {output.numel()} elements cannot hold {full_shape}"
            )
        y = output.view(full_shape)
    if not x.is_cuda:
        # host tensors: the package's torch restatement (same summation order)
        if nbatch > 0:
            y.view(nbatch, *im_size).copy_(
                fold_host(x.reshape(nbatch, *nblocks, *block_size), im_size, block_size,
                          stride, nblocks))
        return y
    xr = _elem_view(x.reshape(nbatch, *nblocks, *block_size).contiguous())
    yr = _elem_view(y)
    if nbatch > 0:
        with torch.cuda.device(x.device):
            check(
                lib().tnl_fold(
                    xr.data_ptr(),
                    yr.data_ptr(),
                    code,
                    nbatch,
                    ndim,
                    i64_array(nblocks),
                    i64_array(_interleaved(block_size, is_complex)),
                    i64_array(_interleaved(im_size, is_complex)),
                    i64_array(_interleaved(stride, is_complex)),
                    stream_of(x),
                )
            )
    return y


class ArrayToBlocksFn(Function):
    """unfold with fold as its gradient (array_to_blocks.py:16-56)."""

    @staticmethod
    def forward(input, block_shape, stride, mask, out):
        output = unfold(input, block_shape, stride, mask, out)
        output.requires_grad_(input.requires_grad)
        return output

    @staticmethod
    def setup_context(ctx, inputs, output):
        input, block_shape, stride, mask, out = inputs
        ctx.im_size = tuple(input.shape[-len(block_shape):])
        ctx.block_shape = block_shape
        ctx.stride = stride
        ctx.save_for_backward(mask)

    @staticmethod
    def backward(ctx, grad_output):
        grad_input = None
        if ctx.needs_input_grad[0]:
            (mask,) = ctx.saved_tensors
            grad_input = fold(grad_output, ctx.im_size, ctx.block_shape, ctx.stride, mask)
        return grad_input, None, None, None, None


def array_to_blocks(input, block_shape, stride=None, mask=None, out=None):
    """Differentiable unfold (array_to_blocks.py:59-67)."""
    return ArrayToBlocksFn.apply(input, block_shape, stride, mask, out)


class BlocksToArrayFn(Function):
    """fold with unfold as its gradient (array_to_blocks.py:70-110)."""

    @staticmethod
    def forward(input, im_size, block_shape, stride, mask, out):
        output = fold(input, im_size, block_shape, stride, mask, out)
        output.requires_grad_(input.requires_grad)
        return output

    @staticmethod
    def setup_context(ctx, inputs, output):
        input, im_size, block_shape, stride, mask, out = inputs
        ctx.block_shape = block_shape
        ctx.stride = stride
        ctx.save_for_backward(mask)

    @staticmethod
    def backward(ctx, grad_output):
        grad_input = None
        if ctx.needs_input_grad[0]:
            (mask,) = ctx.saved_tensors
            grad_input = unfold(grad_output, ctx.block_shape, ctx.stride, mask)
        return grad_input, None, None, None, None, None


def blocks_to_array(input, im_size, block_shape, stride=None, mask=None, out=None):
    """Differentiable fold (array_to_blocks.py:113-122)."""
    return BlocksToArrayFn.apply(input, im_size, block_shape, stride, mask, out)


def get_norm_weights(im_size, block_shape, stride=None, mask=None, device="cpu"):
    """``fold(unfold(ones))``: how many blocks cover each cell
    (array_to_blocks.py:125-140); ``x == fold(unfold(x)) / weights``."""
    x = torch.ones(im_size, device=device)
    Ax = array_to_blocks(x, block_shape, stride, mask)
    return blocks_to_array(Ax, im_size, block_shape, stride, mask)
