"""Autotuner for the plan tiling of the interpolation kernels.

The plan knobs (bin tile ``u_tile`` of the interpolation kernel, gridding
tile ``g_tile``, LDS budget; ``set_tuning``) trade LDS reuse against
occupancy and boundary waste, and the best choice depends on the grid
size, kernel width, channel count and point density.  ``autotune_interp``
(one problem) and ``autotune_batch`` (a ``BatchedInterp`` workload) time
one interpolation plus one gridding for each candidate tiling on the
caller's own inputs and return the fastest; results are memoised per
workload signature and can be saved to / loaded from a JSON file.
"""

from __future__ import annotations

import json
import threading
from math import prod
from pathlib import Path
from typing import Optional, Sequence

import torch

from .. import config
from ._plan import clear_plan_cache, get_tuning, set_tuning

__all__ = ["CANDIDATES", "autotune_interp", "autotune_batch", "save_cache", "load_cache",
           "cached"]

# tilings over the left-padded axes (x, y, z)
CANDIDATES = {
    2: [
        dict(u_tile=(1, 32, 32), g_tile=(1, 16, 64)),
        dict(u_tile=(1, 32, 32), g_tile=(1, 16, 32)),
        dict(u_tile=(1, 32, 32), g_tile=(1, 32, 32)),
        dict(u_tile=(1, 16, 16), g_tile=(1, 16, 32)),
        dict(u_tile=(1, 64, 64), g_tile=(1, 32, 64), lds_budget=81920),
    ],
    3: [
        dict(u_tile=(8, 8, 8), g_tile=(8, 8, 8)),
        dict(u_tile=(8, 8, 8), g_tile=(4, 8, 16)),
        dict(u_tile=(8, 8, 8), g_tile=(8, 8, 16), lds_budget=81920),
        dict(u_tile=(16, 16, 16), g_tile=(8, 8, 8)),
    ],
}

_memo: dict = {}
_lock = threading.Lock()


def _bucket(n: int) -> int:
    """Power-of-two bucket, so near-identical workloads share a result."""
    return max(0, int(n).bit_length() - 1)


def _time_ms(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def cached(key) -> Optional[dict]:
    with _lock:
        hit = _memo.get(key)
    return None if hit is None else dict(hit)


def _store(key, best: dict) -> None:
    with _lock:
        _memo[key] = dict(best)


def autotune_interp(locs: torch.Tensor, grid_size: Sequence[int], width, nbatch: int,
                    kernel_params: Optional[dict] = None, dtype=torch.complex64,
                    candidates: Optional[list] = None, reps: int = 3, apply: bool = True):
    """Pick the tiling for one interpolation problem; with ``apply`` it
    becomes the ndim-D tuning (``set_tuning``).  Returns (best, timings)."""
    from ._interp import interpolate, interpolate_adjoint

    ndim = locs.shape[-1]
    npts = prod(locs.shape[:-1])
    w = width if isinstance(width, (int, float)) else tuple(width)
    key = ("interp", ndim, tuple(int(g) for g in grid_size), w, int(nbatch), _bucket(npts),
           str(dtype))
    hit = cached(key)
    timings = []
    if hit is None:
        cands = candidates if candidates is not None else CANDIDATES[ndim]
        saved = get_tuning(ndim)
        x = torch.randn(nbatch, *grid_size, dtype=dtype, device=locs.device)
        y = torch.randn(nbatch, npts, dtype=dtype, device=locs.device)
        l2 = locs.reshape(npts, ndim)
        try:
            for c in cands:
                set_tuning(ndim, **c)

                def run():
                    with config.using(interp_plan="always"):
                        interpolate(x, l2, width, kernel_params=kernel_params)
                        interpolate_adjoint(y, l2, grid_size, width, kernel_params=kernel_params)

                timings.append((dict(c), _time_ms(run, reps)))
        finally:
            from . import _plan

            with _plan._tuning_lock:
                _plan._tuning[ndim] = saved
            clear_plan_cache()
        hit = min(timings, key=lambda t: t[1])[0]
        _store(key, hit)
    if apply:
        set_tuning(ndim, **hit)
    return hit, timings


def autotune_batch(locs: Sequence[torch.Tensor], grid_sizes, widths, nbatches,
                   kernel_params=None, dtype=torch.complex64,
                   candidates: Optional[list] = None, reps: int = 3):
    """Pick the tiling of a ``BatchedInterp`` workload by timing one batched
    interpolation plus one batched gridding per candidate.  Returns
    (best, timings); pass ``best`` as ``tuning=`` to BatchedInterp (or to
    BatchedNUFFTNormal)."""
    from ._batch import BatchedInterp

    ndim = locs[0].shape[-1]
    key = ("batch", ndim, len(locs),
           tuple(sorted((tuple(int(s) for s in g), int(c), _bucket(prod(l.shape[:-1])))
                        for g, c, l in zip(grid_sizes, nbatches, locs))), str(dtype))
    hit = cached(key)
    if hit is not None:
        return hit, []
    cands = candidates if candidates is not None else CANDIDATES[ndim]
    timings = []
    for c in cands:
        try:
            B = BatchedInterp(locs, grid_sizes, widths, nbatches, kernel_params=kernel_params,
                              dtype=dtype, tuning=c)
        except ValueError:
            continue  # tiling outside what every problem supports
        g = B.new_grid().normal_()
        p = B.new_points()

        def run(B=B, g=g, p=p):
            B.ungrid(g, out=p)
            B.grid(p, out=g)

        timings.append((dict(c), _time_ms(run, reps)))
        del B, g, p
    if not timings:
        raise ValueError("no candidate tiling fits this batch")
    best = min(timings, key=lambda t: t[1])[0]
    _store(key, best)
    return best, timings


def save_cache(path) -> None:
    with _lock:
        items = [[list(map(_jsonable, k)), v] for k, v in _memo.items()]
    Path(path).write_text(json.dumps(items))


def load_cache(path) -> None:
    items = json.loads(Path(path).read_text())
    with _lock:
        for k, v in items:
            _memo[_tupled(k)] = {kk: (tuple(vv) if isinstance(vv, list) else vv)
                                 for kk, vv in v.items()}


def _jsonable(x):
    if isinstance(x, tuple):
        return [_jsonable(v) for v in x]
    return x


def _tupled(x):
    if isinstance(x, list):
        return tuple(_tupled(v) for v in x)
    return x
