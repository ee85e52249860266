"""Functional API of the gridding path (mirrors torchlinops.functional)."""

from ._batch import BatchedInterp
from ._index import (
    canonicalize_idx,
    ensure_tensor_indexing,
    index,
    index_adjoint,
    mask2idx,
    multi_grid,
    multi_index,
    ravel,
    slice2range,
)
from ._interp import grid, interp_grad_locs, interpolate, interpolate_adjoint, ungrid
from ._nufft import center_crop, center_pad, nufft, nufft_adjoint
from ._plan import clear_plan_cache, record_plans, release_plan_pins
from ._unfold import (
    array_to_blocks,
    blocks_to_array,
    fold,
    get_nblocks,
    get_norm_weights,
    unfold,
)

__all__ = [
    "BatchedInterp",
    "grid",
    "ungrid",
    "interpolate",
    "interpolate_adjoint",
    "interp_grad_locs",
    "array_to_blocks",
    "blocks_to_array",
    "fold",
    "unfold",
    "get_nblocks",
    "get_norm_weights",
    "clear_plan_cache",
    "record_plans",
    "release_plan_pins",
    "nufft",
    "nufft_adjoint",
    "center_pad",
    "center_crop",
    "index",
    "index_adjoint",
    "multi_index",
    "multi_grid",
    "ravel",
    "mask2idx",
    "canonicalize_idx",
    "slice2range",
    "ensure_tensor_indexing",
]
