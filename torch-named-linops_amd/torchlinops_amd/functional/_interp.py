"""Interpolation (ungrid) and gridding, the differentiable pair
``interpolate`` / ``interpolate_adjoint``.

Same signatures, defaults, shape rules and errors as the reference's
functional API:
  * ``ungrid`` / ``grid``            src/torchlinops/functional/_interp/ungrid.py:37-67,
                                     grid.py:37-72
  * shape preparation              ungrid.py:577-621, grid.py:522-568
  * kernel defaults                kernels.py:97-105, 245-252
  * autograd pair                  interp.py:13-236
The arithmetic runs in the HIP kernels of libtnl (tnl_ungrid / tnl_grid,
and tnl_ungrid_f64 / tnl_grid_f64 for float64 / complex128, which compute
in double like the reference's torch path) for tensors on a HIP device;
host tensors take the package's torch restatement (``_host.py``), the
reference's own dispatch on ``is_cuda`` (grid.py:91, ungrid.py:86).

Beyond the reference, the autograd pair also differentiates with respect to
``locs`` (the reference returns None there, interp.py:46-70, 157-181): the
gradient comes from the analytic window derivative in the
``tnl_interp_grad_locs`` kernel (``interp_grad_locs`` below).
"""

from __future__ import annotations

import ctypes

from math import prod
from typing import Optional

import torch
from torch.autograd import Function

from .._lib import (
    KERNELS,
    PAD_MODES,
    check,
    f32_array,
    i64_array,
    lib,
    stream_of,
)
from ._host import grid_host, ungrid_host
from .. import config
from ._plan import get_plan

__all__ = [
    "ungrid",
    "grid",
    "interp_grad_locs",
    "interpolate",
    "interpolate_adjoint",
    "get_kernel_params",
]

# Reference limit: next_pow2(ceil(w + 1)) <= 64, doubled in the last dim for
# complex data (grid.py:145-158).
MAX_BLOCK_WIDTH = 64


def get_kernel_params(kernel: str, kernel_params: Optional[dict]) -> dict:
    """Fill in defaults: ``beta = 1.0`` for Kaiser-Bessel (kernels.py:245-252)."""
    kernel_params = {} if kernel_params is None else dict(kernel_params)
    if kernel == "kaiser_bessel":
        kernel_params["beta"] = kernel_params.get("beta", 1.0)
    elif kernel == "spline":
        pass
    else:
        raise ValueError(f"Unrecognized kernel type: {kernel}")
    return kernel_params


def _width_tuple(width, ndim: int) -> tuple:
    if isinstance(width, (int, float)) and not isinstance(width, bool):
        return (float(width),) * ndim
    width = tuple(float(w) for w in width)
    if len(width) != ndim:
        raise ValueError(
            "If width specified as tuple it must be same length as grid size but "
            f"got len(width) = {len(width)} and len(grid_size) = {ndim}"
        )
    return width


def _next_pow2(n: int) -> int:
    p = 1
    while p < n:
        p *= 2
    return p


def _check_block_width(width: tuple, is_complex: bool) -> None:
    bw = [_next_pow2(int(-(-(w + 1) // 1))) for w in width]
    if is_complex:
        bw[-1] *= 2
    if max(bw) > MAX_BLOCK_WIDTH:
        raise ValueError(
            f"Necessary block width {bw} has entry which exceeds maximum width "
            f"{MAX_BLOCK_WIDTH} (kernel width is doubled in last dim if input is complex)"
        )


_F64 = (torch.float64, torch.complex128)


def _check_dtypes(vals: torch.Tensor, locs: torch.Tensor) -> None:
    ok = (torch.float32, torch.complex64) + _F64
    if vals.dtype not in ok:
        raise TypeError(f"interpolation computes in {'/'.join(str(d)[6:] for d in ok)}; "
                        f"got {vals.dtype}")
    if not locs.is_floating_point():
        raise TypeError(f"locs must be floating point, got {locs.dtype}")


def _launch_f64(adjoint: bool, vals_flat, locs, out, nbatch, npts, ndim, grid_size, width,
                kernel, norm, pad_mode, is_complex, kernel_params):
    """Double-precision HIP kernels (tnl_ungrid_f64 / tnl_grid_f64): no plan,
    windows and sums in double."""
    _check_options(kernel, norm, pad_mode)
    locs2d = locs.reshape(npts, ndim).to(torch.float64).contiguous()
    fn = lib().tnl_grid_f64 if adjoint else lib().tnl_ungrid_f64
    with torch.cuda.device(vals_flat.device):
        check(fn(vals_flat.data_ptr(), locs2d.data_ptr(), out.data_ptr(), nbatch, npts, ndim,
                 i64_array(grid_size), f32_array(width), KERNELS[kernel], int(norm),
                 PAD_MODES[pad_mode], int(is_complex), _beta(kernel, kernel_params),
                 stream_of(vals_flat)))


def _same_device(op: str, *tensors) -> None:
    devs = {t.device for t in tensors if t is not None}
    if len(devs) > 1:
        raise RuntimeError(f"{op}: expected all tensors on one device, got {sorted(map(str, devs))}")


def _beta(kernel: str, kernel_params: dict) -> float:
    return float(kernel_params.get("beta", 1.0)) if kernel == "kaiser_bessel" else 1.0


def _check_options(kernel, norm, pad_mode) -> None:
    if kernel not in KERNELS:
        raise ValueError(f"Unrecognized kernel {kernel}")
    if norm not in (1, 2):
        raise ValueError(f"Unexpected norm: {norm}.")
    if pad_mode not in PAD_MODES:
        raise ValueError(f"Unrecognized pad_mode {pad_mode}")


def _launch(adjoint: bool, vals_flat, locs, locs2d, out, nbatch, npts, ndim,
            grid_size, width, kernel, norm, pad_mode, is_complex, kernel_params):
    _check_options(kernel, norm, pad_mode)
    beta = _beta(kernel, kernel_params)
    pad_code = PAD_MODES[pad_mode]
    plan = get_plan(locs, locs2d, ndim, grid_size, width, norm, pad_code)
    args = [
        vals_flat.data_ptr(),
        locs2d.data_ptr(),
        out.data_ptr(),
        nbatch,
        npts,
        ndim,
        i64_array(grid_size),
        f32_array(width),
        KERNELS[kernel],
        int(norm),
        pad_code,
        int(is_complex),
        beta,
        None if plan is None else plan.layout,
        None if plan is None else plan.ptr,
    ]
    with torch.cuda.device(vals_flat.device):
        if plan is not None:
            L = plan.layout
            if L.off_wrec and (L.wrec_kernel != KERNELS[kernel]
                               or L.wrec_beta != ctypes.c_float(beta).value):
                # window records kept with the plan: evaluated once per
                # (plan, kernel, beta), not on every call
                check(lib().tnl_plan_windows(L, plan.ptr, locs2d.data_ptr(), KERNELS[kernel],
                                             beta, stream_of(vals_flat)))
        if not adjoint:
            pm = config.interp_point_major
            if pm not in ("auto", "always", "never"):
                raise ValueError(
                    f"config.interp_point_major must be auto/always/never, got {pm!r}")
            floats = nbatch * (2 if is_complex else 1)
            if plan is not None and (pm == "always" or (pm == "auto" and plan.scattered
                                                        and floats >= 8)):
                # point-major values into a scratch, then one transpose into
                # out (whole lines instead of scattered 8-byte stores)
                nbytes = ctypes.c_int64(0)
                check(lib().tnl_ungrid_scratch_bytes(npts, nbatch, int(is_complex),
                                                     ctypes.byref(nbytes)))
                scratch = torch.empty(nbytes.value, dtype=torch.uint8, device=vals_flat.device)
                check(lib().tnl_ungrid_ws(*args, scratch.data_ptr(), nbytes.value,
                                          stream_of(vals_flat)))
            else:
                check(lib().tnl_ungrid(*args, stream_of(vals_flat)))
            return
        nbytes = ctypes.c_int64(0)
        if plan is not None:
            check(lib().tnl_grid_scratch_bytes(plan.layout, nbatch, int(is_complex),
                                               ctypes.byref(nbytes)))
        if nbytes.value > 0:
            # point-major records for the matrix-core gridding kernel; from
            # torch's caching allocator, so repeated calls do not hipMalloc
            scratch = torch.empty(nbytes.value, dtype=torch.uint8, device=vals_flat.device)
            check(lib().tnl_grid_ws(*args, scratch.data_ptr(), nbytes.value,
                                    stream_of(vals_flat)))
        else:
            check(lib().tnl_grid(*args, stream_of(vals_flat)))


def ungrid(
    vals: torch.Tensor,
    locs: torch.Tensor,
    width,
    kernel: str = "kaiser_bessel",
    norm: int = 1,
    pad_mode: str = "circular",
    kernel_params: Optional[dict] = None,
) -> torch.Tensor:
    """Interpolate on-grid ``vals[..., *grid]`` to off-grid ``locs[..., D]``.

    Output shape ``[*batch_shape, *locs.shape[:-1]]``; ``locs`` are in grid
    units and shared by every batch entry (ungrid.py:37-67, 577-621).
    """
    kernel_params = get_kernel_params(kernel, kernel_params)
    ndim = locs.shape[-1]
    grid_size = tuple(vals.shape[-ndim:])
    batch_shape = vals.shape[:-ndim]
    width = _width_tuple(width, ndim)
    locs_batch_shape = locs.shape[:-1]
    npts = prod(locs_batch_shape)
    is_complex = torch.is_complex(vals)
    _check_block_width(width, is_complex)
    _same_device("ungrid", vals, locs)
    _check_dtypes(vals, locs)
    nbatch = prod(batch_shape)
    vals_flat = vals.reshape(nbatch, *grid_size).contiguous()
    if not vals.is_cuda:
        _check_options(kernel, norm, pad_mode)
        out = ungrid_host(vals_flat, locs.reshape(npts, ndim), grid_size, width, kernel,
                          norm, pad_mode, _beta(kernel, kernel_params))
        return out.reshape(*batch_shape, *locs_batch_shape)
    out = torch.empty(nbatch, npts, dtype=vals.dtype, device=vals.device)
    if is_complex:
        vals_flat = torch.view_as_real(vals_flat)
        out_r = torch.view_as_real(out)
    else:
        out_r = out
    if vals.dtype in _F64:
        if nbatch > 0 and npts > 0:
            _launch_f64(False, vals_flat, locs, out_r, nbatch, npts, ndim, grid_size, width,
                        kernel, norm, pad_mode, is_complex, kernel_params)
        return out.reshape(*batch_shape, *locs_batch_shape)
    locs2d = locs.reshape(npts, ndim).to(torch.float32).contiguous()
    if nbatch > 0 and npts > 0:
        _launch(False, vals_flat, locs, locs2d, out_r, nbatch, npts, ndim,
                grid_size, width, kernel, norm, pad_mode, is_complex, kernel_params)
    return out.reshape(*batch_shape, *locs_batch_shape)


def grid(
    vals: torch.Tensor,
    locs: torch.Tensor,
    grid_size: tuple,
    width,
    kernel: str = "kaiser_bessel",
    norm: int = 1,
    pad_mode: str = "circular",
    kernel_params: Optional[dict] = None,
) -> torch.Tensor:
    """Grid off-grid ``vals[..., *locs.shape[:-1]]`` onto ``grid_size``.

    Output shape ``[*batch_shape, *grid_size]`` (grid.py:37-72, 522-568).
    """
    kernel_params = get_kernel_params(kernel, kernel_params)
    grid_size = tuple(int(g) for g in grid_size)
    ndim = len(grid_size)
    if locs.shape[-1] != ndim:
        raise ValueError(
            f"Expected last dim of locs: to equal ndim but got locs.shape[-1] "
            f"{locs.shape[-1]} != ndim {ndim}"
        )
    locs_batch_shape = locs.shape[:-1]
    npts = prod(locs_batch_shape)
    nlb = len(locs_batch_shape)
    batch_shape = vals.shape[: vals.ndim - nlb]
    width = _width_tuple(width, ndim)
    is_complex = torch.is_complex(vals)
    _check_block_width(width, is_complex)
    _same_device("grid", vals, locs)
    _check_dtypes(vals, locs)
    if tuple(vals.shape[vals.ndim - nlb:]) != tuple(locs_batch_shape):
        raise RuntimeError(
            f"grid expected vals with trailing shape {tuple(locs_batch_shape)} "
            f"but got {tuple(vals.shape)}"
        )
    nbatch = prod(batch_shape)
    vals_flat = vals.reshape(nbatch, npts).contiguous()
    if not vals.is_cuda:
        _check_options(kernel, norm, pad_mode)
        out = grid_host(vals_flat, locs.reshape(npts, ndim), grid_size, width, kernel,
                        norm, pad_mode, _beta(kernel, kernel_params))
        return out.reshape(*batch_shape, *grid_size)
    out = torch.empty(nbatch, *grid_size, dtype=vals.dtype, device=vals.device)
    if is_complex:
        vals_flat = torch.view_as_real(vals_flat)
        out_r = torch.view_as_real(out)
    else:
        out_r = out
    if vals.dtype in _F64:
        if nbatch > 0:
            _launch_f64(True, vals_flat, locs, out_r, nbatch, npts, ndim, grid_size, width,
                        kernel, norm, pad_mode, is_complex, kernel_params)
        return out.reshape(*batch_shape, *grid_size)
    locs2d = locs.reshape(npts, ndim).to(torch.float32).contiguous()
    if nbatch > 0:
        _launch(True, vals_flat, locs, locs2d, out_r, nbatch, npts, ndim,
                grid_size, width, kernel, norm, pad_mode, is_complex, kernel_params)
    return out.reshape(*batch_shape, *grid_size)


def interp_grad_locs(
    grid_vals: torch.Tensor,
    pt_vals: torch.Tensor,
    locs: torch.Tensor,
    width,
    kernel: str = "kaiser_bessel",
    norm: int = 1,
    pad_mode: str = "circular",
    kernel_params: Optional[dict] = None,
) -> torch.Tensor:
    """Trajectory gradient shared by ungrid and grid.

    ``grid_vals[*batch, *grid]`` and ``pt_vals[*batch, *locs.shape[:-1]]``;
    returns ``sum_b Re(conj(pt_vals) * dW/dlocs applied to grid_vals)`` with
    the shape and dtype of ``locs``.  For ``y = ungrid(x, locs)`` and upstream
    gradient ``gy`` it is ``dL/dlocs = interp_grad_locs(x, gy, locs)``; for
    ``X = grid(v, locs)`` with upstream ``gX`` it is
    ``interp_grad_locs(gX, v, locs)``.
    """
    kernel_params = get_kernel_params(kernel, kernel_params)
    ndim = locs.shape[-1]
    grid_size = tuple(grid_vals.shape[-ndim:])
    batch_shape = grid_vals.shape[:-ndim]
    width = _width_tuple(width, ndim)
    npts = prod(locs.shape[:-1])
    nbatch = prod(batch_shape)
    if tuple(pt_vals.shape) != tuple(batch_shape) + tuple(locs.shape[:-1]):
        raise RuntimeError(
            f"interp_grad_locs: pt_vals shape {tuple(pt_vals.shape)} does not match "
            f"batch {tuple(batch_shape)} + points {tuple(locs.shape[:-1])}"
        )
    is_complex = torch.is_complex(grid_vals)
    if torch.is_complex(pt_vals) != is_complex:
        raise TypeError("grid_vals and pt_vals must both be real or both complex")
    _check_options(kernel, norm, pad_mode)
    _same_device("interp_grad_locs", grid_vals, pt_vals, locs)
    _check_dtypes(grid_vals, locs)
    _check_dtypes(pt_vals, locs)
    if grid_vals.is_cuda and (grid_vals.dtype in _F64 or pt_vals.dtype in _F64):
        raise TypeError("the trajectory gradient's HIP kernel computes in float32/complex64; "
                        "use float32/complex64 device tensors or float64 host tensors")
    g = grid_vals.reshape(nbatch, *grid_size).contiguous()
    h = pt_vals.reshape(nbatch, npts).contiguous()
    if not grid_vals.is_cuda:
        # host: differentiate the host interpolation with respect to locs
        with torch.enable_grad():
            l2 = locs.detach().reshape(npts, ndim).to(g.real.dtype).requires_grad_(True)
            y = ungrid_host(g.detach(), l2, grid_size, width, kernel, norm, pad_mode,
                            _beta(kernel, kernel_params))
            loss = (h.detach().conj() * y).real.sum()
            (gl,) = torch.autograd.grad(loss, l2)
        return gl.reshape(locs.shape).to(locs.dtype)
    if is_complex:
        g = torch.view_as_real(g)
        h = torch.view_as_real(h)
    locs2d = locs.detach().reshape(npts, ndim).to(torch.float32).contiguous()
    out = torch.empty(npts, ndim, dtype=torch.float32, device=locs.device)
    beta = _beta(kernel, kernel_params)
    if npts > 0:
        with torch.cuda.device(locs.device):
            check(
                lib().tnl_interp_grad_locs(
                    g.data_ptr(), h.data_ptr(), locs2d.data_ptr(), out.data_ptr(),
                    nbatch, npts, ndim, i64_array(grid_size), f32_array(width),
                    KERNELS[kernel], int(norm), PAD_MODES[pad_mode], int(is_complex),
                    beta, stream_of(locs2d),
                )
            )
    return out.reshape(locs.shape).to(locs.dtype)


class InterpolateFn(Function):
    """``ungrid`` with ``grid`` as its vals-gradient (interp.py:13-70)."""

    @staticmethod
    def forward(vals, locs, width, kernel, norm, pad_mode, kernel_params):
        out = ungrid(vals, locs, width, kernel, norm, pad_mode, kernel_params)
        out.requires_grad_(vals.requires_grad or locs.requires_grad)
        return out

    @staticmethod
    def setup_context(ctx, inputs, output):
        vals, locs, width, kernel, norm, pad_mode, kernel_params = inputs
        ndim = locs.shape[-1]
        ctx.grid_size = tuple(vals.shape[-ndim:])
        ctx.params = (width, kernel, norm, pad_mode, kernel_params)
        if ctx.needs_input_grad[1]:
            ctx.save_for_backward(locs, vals)
        else:
            ctx.save_for_backward(locs)

    @staticmethod
    def backward(ctx, grad_output):
        grad_vals = grad_locs = None
        locs = ctx.saved_tensors[0]
        width, kernel, norm, pad_mode, kernel_params = ctx.params
        if ctx.needs_input_grad[0]:
            grad_vals = grid(grad_output, locs, ctx.grid_size, width, kernel,
                             norm, pad_mode, kernel_params)
        if ctx.needs_input_grad[1]:
            vals = ctx.saved_tensors[1]
            grad_locs = interp_grad_locs(vals, grad_output, locs, width, kernel,
                                         norm, pad_mode, kernel_params)
        return grad_vals, grad_locs, None, None, None, None, None


def interpolate(
    vals: torch.Tensor,
    locs: torch.Tensor,
    width,
    kernel: str = "kaiser_bessel",
    norm: int = 1,
    pad_mode: str = "circular",
    kernel_params: Optional[dict] = None,
) -> torch.Tensor:
    """Differentiable ungrid: grid values -> nonuniform locations
    (src/torchlinops/functional/_interp/interp.py:73-114)."""
    return InterpolateFn.apply(vals, locs, width, kernel, norm, pad_mode, kernel_params)


class InterpolateAdjointFn(Function):
    """``grid`` with ``ungrid`` as its vals-gradient (interp.py:117-181)."""

    @staticmethod
    def forward(vals, locs, grid_size, width, kernel, norm, pad_mode, kernel_params):
        out = grid(vals, locs, grid_size, width, kernel, norm, pad_mode, kernel_params)
        out.requires_grad_(vals.requires_grad or locs.requires_grad)
        return out

    @staticmethod
    def setup_context(ctx, inputs, output):
        vals, locs, grid_size, width, kernel, norm, pad_mode, kernel_params = inputs
        ctx.params = (width, kernel, norm, pad_mode, kernel_params)
        if ctx.needs_input_grad[1]:
            ctx.save_for_backward(locs, vals)
        else:
            ctx.save_for_backward(locs)

    @staticmethod
    def backward(ctx, grad_output):
        grad_vals = grad_locs = None
        locs = ctx.saved_tensors[0]
        width, kernel, norm, pad_mode, kernel_params = ctx.params
        if ctx.needs_input_grad[0]:
            grad_vals = ungrid(grad_output, locs, width, kernel, norm, pad_mode,
                               kernel_params)
        if ctx.needs_input_grad[1]:
            vals = ctx.saved_tensors[1]
            grad_locs = interp_grad_locs(grad_output, vals, locs, width, kernel,
                                         norm, pad_mode, kernel_params)
        return grad_vals, grad_locs, None, None, None, None, None, None


def interpolate_adjoint(
    vals: torch.Tensor,
    locs: torch.Tensor,
    grid_size: tuple,
    width,
    kernel: str = "kaiser_bessel",
    norm: int = 1,
    pad_mode: str = "circular",
    kernel_params: Optional[dict] = None,
) -> torch.Tensor:
    """Differentiable grid: nonuniform values -> grid
    (src/torchlinops/functional/_interp/interp.py:184-236)."""
    return InterpolateAdjointFn.apply(
        vals, locs, grid_size, width, kernel, norm, pad_mode, kernel_params
    )
