"""Point-binning plans for the 2-D/3-D interpolation kernels, cached per
trajectory tensor.

A plan (see tnl_plan_build in include/tnl.h) depends only on the locations,
the grid size, the kernel width and the padding mode.  It is keyed on the
identity and version counter of the caller's ``locs`` tensor, so in-place
edits of ``locs`` invalidate it, and a weak reference guards against a new
tensor reusing a freed tensor's id.
"""

from __future__ import annotations

import threading
import weakref
from collections import OrderedDict

import torch

from .. import config
from .._lib import PlanLayout, check, f32_array, i64_array, lib, stream_of

__all__ = ["get_plan", "plan_layout", "clear_plan_cache"]

_MAX_PLANS = 4
_AUTO_MIN_POINTS = 1 << 16


def plan_layout(npts, ndim, grid_size, width, pad_mode_code) -> PlanLayout:
    L = PlanLayout()
    check(
        lib().tnl_plan_layout_query(
            npts, ndim, i64_array(grid_size), f32_array(width), pad_mode_code, L
        )
    )
    return L


class _PlanCache:
    def __init__(self):
        self._plans: OrderedDict = OrderedDict()
        self._seen: OrderedDict = OrderedDict()
        self._lock = threading.Lock()

    @staticmethod
    def _key(locs, ndim, grid_size, width, pad_code):
        return (
            id(locs),
            locs._version,
            locs.data_ptr(),
            tuple(locs.shape),
            str(locs.device),
            int(ndim),
            tuple(int(g) for g in grid_size),
            tuple(float(w) for w in width),
            int(pad_code),
        )

    def lookup(self, locs, key):
        with self._lock:
            hit = self._plans.get(key)
            if hit is not None and hit[0]() is locs:
                self._plans.move_to_end(key)
                return hit[1]
            return None

    def seen_before(self, locs, key) -> bool:
        with self._lock:
            ref = self._seen.get(key)
            seen = ref is not None and ref() is locs
            self._seen[key] = weakref.ref(locs)
            self._seen.move_to_end(key)
            while len(self._seen) > 64:
                self._seen.popitem(last=False)
            return seen

    def store(self, locs, key, plan):
        with self._lock:
            self._plans[key] = (weakref.ref(locs), plan)
            self._plans.move_to_end(key)
            while len(self._plans) > _MAX_PLANS:
                self._plans.popitem(last=False)

    def clear(self):
        with self._lock:
            self._plans.clear()
            self._seen.clear()


_cache = _PlanCache()


def clear_plan_cache() -> None:
    _cache.clear()


def build_plan(locs2d: torch.Tensor, ndim, grid_size, width, pad_code) -> torch.Tensor:
    npts = locs2d.shape[0]
    L = plan_layout(npts, ndim, grid_size, width, pad_code)
    ws = torch.empty(int(L.bytes), dtype=torch.uint8, device=locs2d.device)
    with torch.cuda.device(locs2d.device):
        check(
            lib().tnl_plan_build(
                locs2d.data_ptr(),
                npts,
                ndim,
                i64_array(grid_size),
                f32_array(width),
                pad_code,
                ws.data_ptr(),
                int(L.bytes),
                stream_of(locs2d),
            )
        )
    return ws


def get_plan(locs, locs2d, ndim, grid_size, width, norm, pad_code):
    """Plan workspace for this trajectory, or None to use direct kernels."""
    mode = config.interp_plan
    if mode not in ("auto", "always", "never"):
        raise ValueError(f"config.interp_plan must be auto/always/never, got {mode!r}")
    if mode == "never" or ndim < 2 or norm != 1:
        return None
    npts = locs2d.shape[0]
    if npts == 0 or npts > (1 << 28):
        return None
    key = _cache._key(locs, ndim, grid_size, width, pad_code)
    plan = _cache.lookup(locs, key)
    if plan is not None:
        return plan
    if mode == "auto" and npts < _AUTO_MIN_POINTS and not _cache.seen_before(locs, key):
        return None
    plan = build_plan(locs2d, ndim, grid_size, width, pad_code)
    _cache.store(locs, key, plan)
    return plan
