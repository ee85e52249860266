"""Point-binning plans for the 2-D/3-D interpolation kernels, cached per
trajectory tensor.

A plan (tnl_plan_layout_query / tnl_plan_build in include/tnl.h) depends
only on the locations, grid size, kernel width, padding mode and tiling.  It
is keyed on the identity and version counter of the caller's ``locs``
tensor, so in-place edits of ``locs`` invalidate it, and a weak reference
guards against a new tensor reusing a freed tensor's id.

HIP graph capture: a plan build only records its kernels while a stream is
being captured, so a plan stored then would hold an unfilled workspace until
the first replay.  Under capture the cache therefore never builds: a cached
plan is used (and pinned, so neither eviction nor ``clear_plan_cache`` frees a
workspace the graph points to), and a miss takes the plan-free kernels.
``record_plans`` lets a capturer keep its own references to the plans its
graph uses; ``release_plan_pins`` drops the cache's pins.
"""

from __future__ import annotations

import threading
import weakref
from collections import OrderedDict
from contextlib import contextmanager
from dataclasses import dataclass
from typing import Optional

import torch

from .. import config
from .._lib import PlanLayout, Tuning, check, f32_array, i64_array, lib, stream_of

__all__ = ["Plan", "get_plan", "build_plan", "plan_layout", "clear_plan_cache",
           "set_tuning", "get_tuning", "record_plans", "release_plan_pins"]

_MAX_PLANS = 4
_AUTO_MIN_POINTS = 1 << 16

# gridding accumulators (TNL_GRID_ACCUM_* in tnl.h)
GRID_ACCUM = {"auto": 0, "lds": 1, "mfma": 2}

# per-ndim tuning overrides (see autotune.py); {} = library defaults
_tuning: dict = {}
_tuning_lock = threading.Lock()


def set_tuning(ndim: int, **knobs) -> None:
    """Set tiling knobs for ndim-D plans: u_tile, g_tile (3-tuples over the
    left-padded axes), max_sub, g_cap, lds_budget, grid_accum ("auto",
    "lds" or "mfma": the 3-D gridding accumulator).  Cached plans are
    dropped so the next call builds with the new tiling."""
    allowed = {"u_tile", "g_tile", "max_sub", "g_cap", "lds_budget", "grid_accum"}
    bad = set(knobs) - allowed
    if bad:
        raise ValueError(f"unknown tuning knobs {bad}")
    with _tuning_lock:
        cur = dict(_tuning.get(ndim, {}))
        cur.update({k: v for k, v in knobs.items() if v is not None})
        _tuning[ndim] = cur
    clear_plan_cache()


def get_tuning(ndim: int) -> dict:
    with _tuning_lock:
        return dict(_tuning.get(ndim, {}))


def _tuning_struct(ndim: int, overrides: Optional[dict] = None) -> Tuning:
    knobs = get_tuning(ndim)
    if overrides:
        knobs.update(overrides)
    t = Tuning()
    for k in ("u_tile", "g_tile"):
        if k in knobs:
            v = tuple(int(x) for x in knobs[k])
            if len(v) != 3:
                raise ValueError(f"{k} must have 3 entries (left-padded axes)")
            getattr(t, k)[:] = v
    for k in ("max_sub", "g_cap", "lds_budget"):
        if k in knobs:
            setattr(t, k, int(knobs[k]))
    if "grid_accum" in knobs:
        mode = knobs["grid_accum"]
        if mode not in GRID_ACCUM:
            raise ValueError(f"grid_accum must be one of {sorted(GRID_ACCUM)}, got {mode!r}")
        t.grid_accum = GRID_ACCUM[mode]
    return t


def plan_layout(npts, ndim, grid_size, width, pad_mode_code, tuning=None) -> PlanLayout:
    L = PlanLayout()
    t = _tuning_struct(ndim, tuning)
    check(
        lib().tnl_plan_layout_query(
            npts, ndim, i64_array(grid_size), f32_array(width), pad_mode_code, t, L
        )
    )
    return L


@dataclass
class Plan:
    layout: PlanLayout
    workspace: torch.Tensor
    # trajectory listed in scattered order (order_coherence below
    # SCATTERED_BELOW): interpolation then writes through a point-major scratch
    scattered: bool = False

    @property
    def ptr(self) -> int:
        return self.workspace.data_ptr()


# Below this fraction of caller-order neighbours sharing an interpolation bin
# a trajectory counts as listed in scattered order: spokes and spirals in
# acquisition order give 0.9 and more, points in random order about 0.
SCATTERED_BELOW = 0.25


def order_coherence(locs2d: torch.Tensor, tile, runs: int = 256, run_len: int = 4096) -> float:
    """Fraction of consecutive points (p, p + 1) of the caller's order whose
    coordinates fall in the same interpolation bin (tile of `tile` cells per
    axis).  Lists longer than runs * run_len points are sampled as `runs`
    evenly spaced runs of `run_len` consecutive points.  One host read."""
    n = locs2d.shape[0]
    if n < 2:
        return 1.0
    t = torch.tensor([float(v) for v in tile], dtype=locs2d.dtype, device=locs2d.device)
    if n > runs * run_len:
        starts = torch.linspace(0, n - run_len, runs, device=locs2d.device).long()
        idx = starts[:, None] + torch.arange(run_len, device=locs2d.device)[None, :]
        key = torch.floor(locs2d[idx] / t)  # [runs, run_len, ndim]
        return float((key[:, 1:] == key[:, :-1]).all(dim=-1).float().mean())
    key = torch.floor(locs2d / t)
    return float((key[1:] == key[:-1]).all(dim=1).float().mean())


def build_plan(locs2d: torch.Tensor, ndim, grid_size, width, pad_code, tuning=None) -> Plan:
    npts = locs2d.shape[0]
    L = plan_layout(npts, ndim, grid_size, width, pad_code, tuning)
    ws = torch.empty(int(L.bytes), dtype=torch.uint8, device=locs2d.device)
    with torch.cuda.device(locs2d.device):
        check(lib().tnl_plan_build(locs2d.data_ptr(), L, ws.data_ptr(), int(L.bytes),
                                   stream_of(locs2d)))
        tile = [L.u_tile[d] for d in range(3 - ndim, 3)]
        scattered = order_coherence(locs2d, tile) < SCATTERED_BELOW
    return Plan(L, ws, scattered)


class _PlanCache:
    def __init__(self):
        self._plans: OrderedDict = OrderedDict()
        self._seen: OrderedDict = OrderedDict()
        self._pinned: dict = {}
        self._lock = threading.Lock()

    @staticmethod
    def key(locs, ndim, grid_size, width, pad_code):
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

    def pin(self, plan):
        with self._lock:
            self._pinned[id(plan)] = plan

    def release_pins(self):
        with self._lock:
            self._pinned.clear()


_cache = _PlanCache()
_recorders = threading.local()


def clear_plan_cache() -> None:
    """Drop the cached plans (plans pinned by a graph capture stay alive)."""
    _cache.clear()


def release_plan_pins() -> None:
    """Drop the cache's references to plans used under graph capture; call
    once the graphs that recorded them are gone."""
    _cache.release_pins()


@contextmanager
def record_plans(sink: list):
    """Append every plan used by this thread while active to ``sink``, so a
    graph capturer can hold the workspaces its graph reads."""
    stack = getattr(_recorders, "stack", None)
    if stack is None:
        stack = _recorders.stack = []
    stack.append(sink)
    try:
        yield sink
    finally:
        stack.pop()


def _note(plan: "Plan", capturing: bool) -> "Plan":
    for sink in getattr(_recorders, "stack", ()):
        if not any(p is plan for p in sink):
            sink.append(plan)
    if capturing:
        _cache.pin(plan)
    return plan


def _capturing(device) -> bool:
    return device.type == "cuda" and torch.cuda.is_current_stream_capturing()


def get_plan(locs, locs2d, ndim, grid_size, width, norm, pad_code) -> Optional[Plan]:
    """Plan for this trajectory, or None to use the direct kernels."""
    mode = config.interp_plan
    if mode not in ("auto", "always", "never"):
        raise ValueError(f"config.interp_plan must be auto/always/never, got {mode!r}")
    if mode == "never" or ndim < 2 or norm != 1:
        return None
    npts = locs2d.shape[0]
    if npts == 0 or npts > (1 << 28):
        return None
    capturing = _capturing(locs2d.device)
    key = _cache.key(locs, ndim, grid_size, width, pad_code)
    plan = _cache.lookup(locs, key)
    if plan is not None:
        return _note(plan, capturing)
    if capturing:
        # a build would only be recorded, not run: take the plan-free kernels
        return None
    if mode == "auto" and npts < _AUTO_MIN_POINTS and not _cache.seen_before(locs, key):
        return None
    plan = build_plan(locs2d, ndim, grid_size, width, pad_code)
    _cache.store(locs, key, plan)
    return _note(plan, False)
