"""ctypes binding of libtnl.so (the C ABI declared in include/tnl.h).

Only plain pointers, sizes and a hipStream_t cross the boundary; torch is
used here solely to hand over device pointers and the current stream.
Every entry point takes tensors on a HIP device, and loading raises if the
library is missing: device tensors never run anywhere but libtnl.  Host
tensors are dispatched before this module, at the reference's own
``is_cuda`` branch, to the package's torch restatement (functional/_host.py).
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch

__all__ = [
    "lib",
    "library_path",
    "TNLError",
    "PlanLayout",
    "Tuning",
    "KERNELS",
    "PAD_MODES",
]

_HERE = Path(__file__).resolve().parent
_LIB_NAME = "libtnl.so"

TNL_OK = 0
TNL_EINVAL = 1
TNL_ESHAPE = 2
TNL_EHIP = 3
TNL_EWORKSPACE = 4

KERNELS = {"kaiser_bessel": 0, "spline": 1}
PAD_MODES = {"circular": 0, "zero": 1}
TNL_F32, TNL_F64 = 0, 1
TNL_I32, TNL_I64, TNL_I16, TNL_I8, TNL_U8 = 2, 3, 4, 5, 6

_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f32 = ctypes.c_float
_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)


class TNLError(RuntimeError):
    """A HIP-level failure reported by libtnl."""


class Tuning(ctypes.Structure):
    """Mirror of tnl_tuning (include/tnl.h); 0 means library default."""

    _fields_ = [
        ("u_tile", ctypes.c_int32 * 3),
        ("g_tile", ctypes.c_int32 * 3),
        ("max_sub", ctypes.c_int32),
        ("g_cap", ctypes.c_int32),
        ("lds_budget", ctypes.c_int32),
        ("grid_accum", ctypes.c_int32),
    ]


class PlanLayout(ctypes.Structure):
    """Mirror of tnl_plan_layout (include/tnl.h)."""

    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("ntaps", ctypes.c_int32 * 3),
        ("u_tile", ctypes.c_int32 * 3),
        ("u_ntiles", ctypes.c_int32 * 3),
        ("u_nbins", ctypes.c_int64),
        ("g_tile", ctypes.c_int32 * 3),
        ("g_ntiles", ctypes.c_int32 * 3),
        ("g_ntiles_total", ctypes.c_int64),
        ("grid_size", ctypes.c_int64 * 3),
        ("width", ctypes.c_float * 3),
        ("pad_mode", ctypes.c_int32),
        ("npts", ctypes.c_int64),
        ("max_sub", ctypes.c_int64),
        ("sub_capacity", ctypes.c_int64),
        ("g_cap", ctypes.c_int64),
        ("touch_capacity", ctypes.c_int64),
        ("item_capacity", ctypes.c_int64),
        ("lds_budget", ctypes.c_int64),
        ("grid_accum", ctypes.c_int64),
        ("off_counts", ctypes.c_int64),
        ("off_starts", ctypes.c_int64),
        ("off_cursor", ctypes.c_int64),
        ("off_perm", ctypes.c_int64),
        ("off_slocs", ctypes.c_int64),
        ("off_suboff", ctypes.c_int64),
        ("off_sub", ctypes.c_int64),
        ("off_gcounts", ctypes.c_int64),
        ("off_gstarts", ctypes.c_int64),
        ("off_gcursor", ctypes.c_int64),
        ("off_tlist", ctypes.c_int64),
        ("off_itemoff", ctypes.c_int64),
        ("off_items", ctypes.c_int64),
        ("off_items_tmp", ctypes.c_int64),
        ("off_tpm", ctypes.c_int64),
        ("off_wrec", ctypes.c_int64),
        ("wrec_kernel", ctypes.c_int32),
        ("wrec_beta", ctypes.c_float),
        ("bytes", ctypes.c_int64),
    ]


_layout_p = ctypes.POINTER(PlanLayout)
_tuning_p = ctypes.POINTER(Tuning)


class BatchProblem(ctypes.Structure):
    """Mirror of tnl_batch_problem (include/tnl.h)."""

    _fields_ = [
        ("nbatch", ctypes.c_int64),
        ("npts", ctypes.c_int64),
        ("grid_size", ctypes.c_int64 * 3),
        ("width", ctypes.c_float * 3),
        ("beta", ctypes.c_float),
        ("grid_offset", ctypes.c_int64),
        ("pts_offset", ctypes.c_int64),
        ("layout", _layout_p),
        ("plan", ctypes.c_void_p),
    ]


class BatchLayout(ctypes.Structure):
    """Mirror of tnl_batch_layout (include/tnl.h)."""

    _fields_ = [
        ("nprob", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("kernel", ctypes.c_int32),
        ("norm", ctypes.c_int32),
        ("pad_mode", ctypes.c_int32),
        ("is_complex", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("ungrid_cc", ctypes.c_int32),
        ("grid_cc", ctypes.c_int32),
        ("ungrid_chunks", ctypes.c_int32),
        ("grid_mfma", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("total_sub", ctypes.c_int64),
        ("total_items", ctypes.c_int64),
        ("ungrid_lds", ctypes.c_int64),
        ("grid_lds", ctypes.c_int64),
        ("bytes", ctypes.c_int64),
    ]


_blayout_p = ctypes.POINTER(BatchLayout)


class SegLayout(ctypes.Structure):
    """Mirror of tnl_seg_layout (include/tnl.h)."""

    _fields_ = [
        ("nseg", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("numel", ctypes.c_int64),
        ("block_floats", ctypes.c_int64),
        ("nblocks", ctypes.c_int64),
        ("off_blocks", ctypes.c_int64),
        ("off_blkseg", ctypes.c_int64),
        ("off_segblk", ctypes.c_int64),
        ("off_partial", ctypes.c_int64),
        ("bytes", ctypes.c_int64),
    ]


_slayout_p = ctypes.POINTER(SegLayout)

# (name, restype, argtypes) for every symbol in include/tnl.h
SIGNATURES = {
    "tnl_last_error": (ctypes.c_char_p, []),
    "tnl_abi_version": (_i32, []),
    "tnl_plan_layout_query": (
        _i32,
        [_i64, _i32, _i64p, _f32p, _i32, _tuning_p, _layout_p],
    ),
    "tnl_plan_build": (_i32, [_vp, _layout_p, _vp, _i64, _vp]),
    "tnl_plan_windows": (_i32, [_layout_p, _vp, _vp, _i32, _f32, _vp]),
    "tnl_ungrid": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32,
         _i32, _f32, _layout_p, _vp, _vp],
    ),
    "tnl_grid": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32,
         _i32, _f32, _layout_p, _vp, _vp],
    ),
    "tnl_ungrid_scratch_bytes": (_i32, [_i64, _i64, _i32, _i64p]),
    "tnl_ungrid_ws": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32,
         _i32, _f32, _layout_p, _vp, _vp, _i64, _vp],
    ),
    "tnl_grid_scratch_bytes": (_i32, [_layout_p, _i64, _i32, _i64p]),
    "tnl_grid_ws": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32,
         _i32, _f32, _layout_p, _vp, _vp, _i64, _vp],
    ),
    "tnl_ungrid_f64": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32, _i32,
         ctypes.c_double, _vp],
    ),
    "tnl_grid_f64": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32, _i32,
         ctypes.c_double, _vp],
    ),
    "tnl_grid_static": (
        _i32,
        [_vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32,
         _i32, _f32, _layout_p, _vp, _vp],
    ),
    "tnl_batch_layout_query": (
        _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _blayout_p]),
    "tnl_batch_build": (_i32, [_vp, _blayout_p, _vp, _vp]),
    "tnl_ungrid_batched": (_i32, [_vp, _vp, _blayout_p, _vp, _vp]),
    "tnl_grid_batched": (_i32, [_vp, _vp, _blayout_p, _vp, _vp]),
    "tnl_seg_layout_query": (_i32, [_i64p, _i32, _i64, _slayout_p]),
    "tnl_seg_build": (_i32, [_i64p, _slayout_p, _vp, _vp]),
    "tnl_seg_dot": (_i32, [_vp, _vp, _vp, _slayout_p, _vp, _vp]),
    "tnl_seg_cg_update": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _slayout_p, _vp, _vp]),
    "tnl_seg_cg_direction": (_i32, [_vp, _vp, _vp, _vp, _slayout_p, _vp, _vp]),
    "tnl_conv3d_fwd": (
        _i32, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp]),
    "tnl_conv3d_wgrad_workspace": (_i64, [_i64, _i32, _i32, _i32]),
    "tnl_conv3d_wgrad": (
        _i32, [_vp, _vp, _vp, _vp, _i64, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _vp]),
    "tnl_interp_grad_locs": (
        _i32,
        [_vp, _vp, _vp, _vp, _i64, _i64, _i32, _i64p, _f32p, _i32, _i32, _i32,
         _i32, _f32, _vp],
    ),
    "tnl_unfold": (
        _i32,
        [_vp, _vp, _i32, _i64, _i32, _i64p, _i64p, _i64p, _i64p, _vp],
    ),
    "tnl_fold": (
        _i32,
        [_vp, _vp, _i32, _i64, _i32, _i64p, _i64p, _i64p, _i64p, _vp],
    ),
    "tnl_index": (_i32, [_vp, _vp, _vp, _i32, _i64, _i64, _i64, _vp]),
    "tnl_pad_wrapped": (_i32, [_vp, _vp, _i32, _i64, _i32, _i64p, _i64p, _vp, _vp]),
    "tnl_crop_wrapped": (_i32, [_vp, _vp, _i32, _i64, _i32, _i64p, _i64p, _vp, _vp]),
    "tnl_index_adjoint": (
        _i32, [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _i32, _i32, _i64, _i64, _i64, _vp]),
}

_lock = threading.Lock()
_handle = None


def library_path() -> Path:
    return _HERE / _LIB_NAME


def lib():
    """Load libtnl.so once (after torch, so both share one HIP runtime)."""
    global _handle
    if _handle is not None:
        return _handle
    with _lock:
        if _handle is None:
            path = library_path()
            if not path.exists():
                raise RuntimeError(
                    f"{path} is not built; run "
                    "`python torch-named-linops_amd/build_lib.py` (or "
                    "__graft_entry__.build()) first"
                )
            h = ctypes.CDLL(str(path))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
            _handle = h
    return _handle


def check(rc: int) -> None:
    """Map a libtnl status code to the reference's exception types."""
    if rc == TNL_OK:
        return
    msg = lib().tnl_last_error().decode(errors="replace")
    if rc == TNL_EINVAL:
        raise ValueError(msg)
    if rc == TNL_ESHAPE:
        raise RuntimeError(msg)
    raise TNLError(f"libtnl error {rc}: {msg}")


def i64_array(vals) -> ctypes.Array:
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * max(len(vals), 1))(*vals)


def f32_array(vals) -> ctypes.Array:
    vals = [float(v) for v in vals]
    return (ctypes.c_float * max(len(vals), 1))(*vals)


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def require_device(*tensors: torch.Tensor, op: str) -> None:
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError(
                f"torchlinops_amd.{op} runs on HIP devices only; got a tensor on "
                f"{t.device}. Move inputs to the GPU (this op has no host path)."
            )
