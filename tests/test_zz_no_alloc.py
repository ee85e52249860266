"""The allocation-free gridding entry point tnl_grid_static, through the C
ABI, against tnl_grid_ws with a caller workspace: two calls in a row, the
second with a smaller input, each equal to the workspace entry point's
result (the same kernels on the same plan: bit-identical without split
tiles) and within 1e-5 relative L2 of the float64 oracle."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests._util import crandn, rel_l2

pytestmark = pytest.mark.gpu


def _grid_both(dev, npts, nbatch, n, seed):
    from torchlinops_amd import _lib
    from torchlinops_amd.functional._plan import build_plan

    rng = np.random.default_rng(seed)
    locs = rng.uniform(0, n, (npts, 3)).astype(np.float32)
    vals = crandn(rng, (nbatch, npts))
    beta = float(O.nufft_beta(4.0, 2.0))
    lt = torch.from_numpy(locs).to(dev)
    vt = torch.view_as_real(torch.from_numpy(vals).to(dev)).contiguous()
    plan = build_plan(lt, 3, (n,) * 3, (4.0,) * 3, 0)
    L = plan.layout
    assert L.grid_accum == 2  # matrix-core gridding: the kernels use the records
    h = _lib.lib()
    need = ctypes.c_int64(0)
    _lib.check(h.tnl_grid_scratch_bytes(L, nbatch, 1, ctypes.byref(need)))
    assert 0 < need.value <= (64 << 20)
    args = lambda out: [vt.data_ptr(), lt.data_ptr(), out.data_ptr(), nbatch, npts, 3,  # noqa: E731
                        _lib.i64_array((n,) * 3), _lib.f32_array((4.0,) * 3), 0, 1, 0, 1,
                        beta, L, plan.ptr]
    stream = torch.cuda.current_stream(dev).cuda_stream
    out_ws = torch.full((nbatch, n, n, n, 2), float("nan"), device=dev)
    out_st = torch.full((nbatch, n, n, n, 2), float("nan"), device=dev)
    ws = torch.empty(need.value, dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        _lib.check(h.tnl_grid_ws(*args(out_ws), ws.data_ptr(), need.value, stream))
        _lib.check(h.tnl_grid_static(*args(out_st), stream))
    torch.cuda.synchronize()
    got_ws = torch.view_as_complex(out_ws).cpu().numpy()
    got_st = torch.view_as_complex(out_st).cpu().numpy()
    ref = O.grid(vals, locs, (n,) * 3, 4.0, beta=beta, precision="f64", nthreads=1)
    return got_ws, got_st, ref


def test_grid_static_matches_grid_ws_twice(dev):
    # first call, then a smaller one on the same device: the shared static
    # records of the first call must not leak into the second
    for npts, nbatch, n, seed in ((120_000, 8, 48, 1), (30_000, 3, 40, 2)):
        got_ws, got_st, ref = _grid_both(dev, npts, nbatch, n, seed)
        assert np.isfinite(got_st).all()
        assert rel_l2(got_st, got_ws) < 1e-6, (npts, nbatch)
        assert np.array_equal(got_st, got_ws), (npts, nbatch)
        assert rel_l2(got_st, ref) < 1e-5, (npts, nbatch)
