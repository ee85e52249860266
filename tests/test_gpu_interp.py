"""GPU parity of interpolate / interpolate_adjoint against the oracle.

Bar: relative L2 error <= 1e-5 against the oracle's reference-order float32
restatement, and the error against the float64 oracle is asserted (and
printed) under the same bound.  Both the binned (plan) kernels and the
direct kernels are exercised.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests._util import crandn, radial_3d, rel_l2, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5
BETA = float(O.nufft_beta(4.0, 2.0))


def _run(dev, grid_size, locs, nbatch, width, kernel, pad, norm, beta, plan, cplx, seed):
    import torchlinops_amd.functional as F
    from torchlinops_amd import config

    rng = np.random.default_rng(seed)
    shape = (nbatch, *grid_size)
    x = crandn(rng, shape) if cplx else rng.standard_normal(shape).astype(np.float32)
    npts = locs.shape[0]
    y = crandn(rng, (nbatch, npts)) if cplx else rng.standard_normal((nbatch, npts)).astype(np.float32)
    kp = {"beta": beta}
    with config.using(interp_plan=plan):
        got_u = F.interpolate(to_dev(x, dev), to_dev(locs, dev), width, kernel, norm, pad, kp)
        got_g = F.interpolate_adjoint(to_dev(y, dev), to_dev(locs, dev), grid_size, width,
                                      kernel, norm, pad, kp)
    torch.cuda.synchronize()
    got_u, got_g = got_u.cpu().numpy(), got_g.cpu().numpy()
    ref_u = O.ungrid(x, locs, width, kernel, norm, pad, beta, "f32")
    ref_g = O.grid(y, locs, grid_size, width, kernel, norm, pad, beta, "f32")
    ref_u64 = O.ungrid(x, locs, width, kernel, norm, pad, beta, "f64")
    ref_g64 = O.grid(y, locs, grid_size, width, kernel, norm, pad, beta, "f64")
    e = dict(u32=rel_l2(got_u, ref_u), g32=rel_l2(got_g, ref_g),
             u64=rel_l2(got_u, ref_u64), g64=rel_l2(got_g, ref_g64))
    print(f"grid={grid_size} npts={npts} nb={nbatch} w={width} {kernel} {pad} norm={norm} "
          f"plan={plan} cplx={cplx}: {e}")
    for k, v in e.items():
        assert v <= TOL, (k, v)
    return got_u, got_g


CASES = [
    # grid_size, npts, nbatch, width, kernel, pad, norm
    ((64,), 300, 2, 4.0, "kaiser_bessel", "circular", 1),
    ((50,), 200, 1, 5.5, "spline", "zero", 1),
    ((32, 40), 2000, 3, 4.0, "kaiser_bessel", "circular", 1),
    ((33, 31), 1500, 2, (3.0, 6.0), "kaiser_bessel", "zero", 1),
    ((24, 24), 800, 1, 2.0, "spline", "circular", 1),
    ((96, 80), 3000, 11, 4.0, "kaiser_bessel", "circular", 1),  # 2-D matrix-core tiles, 2 chunks
    ((20, 24), 500, 2, 4.0, "spline", "circular", 2),
    ((16, 18, 20), 1500, 2, 4.0, "kaiser_bessel", "circular", 1),
    ((16, 16, 16), 1000, 5, (4.0, 3.0, 7.0), "kaiser_bessel", "zero", 1),
    ((12, 10, 14), 400, 1, 4.0, "spline", "zero", 2),
    ((9, 10, 11), 300, 2, 9.5, "kaiser_bessel", "circular", 1),
    ((6, 5, 7), 300, 3, 4.0, "kaiser_bessel", "circular", 1),  # grid smaller than taps*2
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[4]}-{c[5]}-n{c[6]}")
@pytest.mark.parametrize("plan", ["never", "always"])
@pytest.mark.parametrize("cplx", [True, False])
def test_interp_parity(dev, case, plan, cplx):
    grid_size, npts, nbatch, width, kernel, pad, norm = case
    rng = np.random.default_rng(11)
    ndim = len(grid_size)
    locs = rng.uniform(-3, np.array(grid_size) + 3, (npts, ndim)).astype(np.float32)
    locs[:7] = np.round(locs[:7])  # integer coordinates: w+1 taps
    _run(dev, grid_size, locs, nbatch, width, kernel, pad, norm, 1.7, plan, cplx, seed=12)


def test_c1_shape_full_size(dev):
    """2-D 128^2 complex64, 16384 uniform points, KB w=4, beta=8.996."""
    rng = np.random.default_rng(1)
    locs = rng.uniform(0, 128, (16384, 2)).astype(np.float32)
    for plan in ("never", "always"):
        _run(dev, (128, 128), locs, 1, 4.0, "kaiser_bessel", "circular", 1, BETA, plan, True, 2)


def test_radial_3d_reduced(dev):
    """C2's trajectory family at reduced size: 3-D radial, 2 coils, 48^3."""
    rng = np.random.default_rng(2)
    locs = radial_3d(300, 48, 48, rng).reshape(-1, 3)
    for plan in ("never", "always"):
        _run(dev, (48, 48, 48), locs, 2, 4.0, "kaiser_bessel", "circular", 1, BETA, plan, True, 3)


def test_dense_tiles_split(dev):
    """Many points in one gridding tile: the tile is split across workgroups
    (shared tiles combined with global atomics)."""
    rng = np.random.default_rng(4)
    locs = (rng.standard_normal((60000, 2)) * 1.5 + 30).astype(np.float32)
    _run(dev, (64, 64), locs, 2, 4.0, "kaiser_bessel", "circular", 1, BETA, "always", True, 5)
    locs3 = (rng.standard_normal((40000, 3)) * 1.0 + 10).astype(np.float32)
    _run(dev, (24, 24, 32), locs3, 3, 4.0, "kaiser_bessel", "zero", 1, BETA, "always", True, 6)


def test_empty_and_edge_inputs(dev):
    import torchlinops_amd.functional as F

    x = torch.randn(2, 16, 16, dtype=torch.complex64, device=dev)
    locs = torch.zeros(0, 2, device=dev)
    assert F.interpolate(x, locs, 4.0).shape == (2, 0)
    g = F.interpolate_adjoint(torch.zeros(2, 0, dtype=torch.complex64, device=dev), locs, (16, 16), 4.0)
    assert g.shape == (2, 16, 16) and bool((g == 0).all())
    # non-finite coordinates contribute nothing
    bad = torch.tensor([[float("nan"), 3.0], [2.0, float("inf")]], device=dev)
    assert bool((F.interpolate(x, bad, 4.0) == 0).all())
    assert bool((F.interpolate_adjoint(torch.ones(2, 2, dtype=torch.complex64, device=dev), bad,
                                       (16, 16), 4.0) == 0).all())
    # batch shapes: [..B, *grid] with locs [..L, D]
    x5 = torch.randn(2, 3, 16, 16, dtype=torch.complex64, device=dev)
    l5 = torch.rand(4, 5, 2, device=dev) * 16
    out = F.interpolate(x5, l5, 4.0)
    assert out.shape == (2, 3, 4, 5)
    back = F.interpolate_adjoint(out, l5, (16, 16), 4.0)
    assert back.shape == (2, 3, 16, 16)


def test_autograd_pair(dev):
    """vals.grad of interpolate is interpolate_adjoint(ones) and vice versa
    (the reference's test_interp_diff.py)."""
    import torchlinops_amd.functional as F

    locs = torch.rand(700, 2, device=dev) * 32
    x = torch.randn(2, 32, 32, dtype=torch.complex64, device=dev, requires_grad=True)
    y = F.interpolate(x, locs, 4.0)
    y.real.sum().backward()
    want = F.interpolate_adjoint(torch.ones_like(y).real.to(y.dtype), locs, (32, 32), 4.0)
    assert torch.allclose(x.grad, want, rtol=1e-5, atol=1e-5)
    v = torch.randn(2, 700, dtype=torch.complex64, device=dev, requires_grad=True)
    g = F.interpolate_adjoint(v, locs, (32, 32), 4.0)
    g.real.sum().backward()
    want2 = F.interpolate(torch.ones(2, 32, 32, dtype=torch.complex64, device=dev), locs, 4.0)
    assert torch.allclose(v.grad, want2, rtol=1e-5, atol=1e-5)


def test_plan_cache_reuse_and_invalidation(dev):
    import torchlinops_amd.functional as F
    from torchlinops_amd import config

    rng = np.random.default_rng(8)
    locs_np = rng.uniform(0, 40, (3000, 2)).astype(np.float32)
    locs = to_dev(locs_np, dev)
    x = to_dev(crandn(rng, (1, 40, 40)), dev)
    with config.using(interp_plan="always"):
        a = F.interpolate(x, locs, 4.0).cpu().numpy()
        locs.add_(1.25)  # in-place edit bumps the version counter: new plan
        b = F.interpolate(x, locs, 4.0).cpu().numpy()
    ref_a = O.ungrid(x.cpu().numpy(), locs_np, 4.0, beta=1.0)
    ref_b = O.ungrid(x.cpu().numpy(), locs_np + 1.25, 4.0, beta=1.0)
    assert rel_l2(a, ref_a) < TOL and rel_l2(b, ref_b) < TOL


ACC_CASES = [
    # grid_size, npts, nbatch, width, kernel, pad, cplx, g_cap
    ((24, 40, 33), 3000, 3, 4.0, "kaiser_bessel", "circular", True, 0),
    ((24, 40, 33), 3000, 8, 4.0, "kaiser_bessel", "zero", True, 0),
    ((32, 32, 32), 2500, 11, 4.0, "kaiser_bessel", "circular", True, 64),
    ((32, 32, 32), 2500, 17, 4.0, "spline", "zero", False, 0),
    ((32, 24, 40), 2000, 2, 6.0, "kaiser_bessel", "circular", True, 0),
    ((32, 24, 40), 2000, 1, (4.0, 3.0, 7.0), "kaiser_bessel", "zero", False, 128),
    ((17, 19, 23), 1500, 5, 2.5, "spline", "circular", True, 0),
    # 2-D, 16x32 tiles
    ((40, 70), 3000, 3, 4.0, "kaiser_bessel", "circular", True, 0),
    ((64, 96), 4000, 11, 6.0, "kaiser_bessel", "zero", True, 64),
    ((48, 80), 3000, 17, (3.0, 5.0), "spline", "circular", False, 0),
    ((40, 70), 3000, 2, 5.0, "kaiser_bessel", "zero", False, 0),
]


@pytest.mark.parametrize("case", ACC_CASES, ids=lambda c: f"{c[0]}-w{c[3]}-{c[5]}-nb{c[2]}")
@pytest.mark.parametrize("accum", ["mfma", "lds"])
def test_grid_accumulators(dev, case, accum):
    """Gridding on 8x8x8 (3-D) / 16x32 (2-D) tiles through either
    accumulator (matrix-core blocks or the LDS read-modify-write) against
    the oracle: channel chunks (nbatch above 8 complex / 16 real), widths
    past 5 taps, anisotropic widths, both pads, points on integer
    coordinates and at the edges, and split tiles (small g_cap)."""
    from torchlinops_amd.functional import _plan

    grid_size, npts, nbatch, width, kernel, pad, cplx, g_cap = case
    ndim = len(grid_size)
    ntap = int(np.ceil(np.max(width))) + 1
    if accum == "lds" and ndim == 3 and ntap > 5:
        pytest.skip("the LDS accumulator covers at most 5 taps per axis in 3-D")
    rng = np.random.default_rng(31)
    locs = rng.uniform(-2, np.array(grid_size) + 2, (npts, ndim)).astype(np.float32)
    locs[:9] = np.round(locs[:9])
    locs[9:12] = np.array(grid_size, dtype=np.float32) - 0.25
    knobs = {"grid_accum": accum}
    if g_cap:
        knobs["g_cap"] = g_cap
    try:
        _plan.set_tuning(ndim, **knobs)
        _run(dev, grid_size, locs, nbatch, width, kernel, pad, 1, 1.7, "always", cplx, seed=32)
    finally:
        with _plan._tuning_lock:
            _plan._tuning.pop(ndim, None)
        _plan.clear_plan_cache()


# The window is unnormalised, so its D-fold product reaches I0(beta)^D:
# beta is kept where that stays well inside float32 (3-D up to 25, 2-D up
# to 30, 1-D beyond).
BETA_CASES = [(3, 0.5), (3, 9.0), (3, 10.1), (3, 19.9), (3, 20.1), (3, 25.0), (2, 29.9), (2, 30.1),
              (1, 45.0)]


@pytest.mark.parametrize("ndim,beta", BETA_CASES, ids=lambda v: str(v))
def test_kb_beta_regimes(dev, ndim, beta):
    """Kaiser-Bessel weights across the window evaluation regimes (power
    series of I0 with 16 terms up to beta 10, 24 up to 20, 32 up to 30, the
    Abramowitz & Stegun polynomials above) against the oracle's
    reference-formula weights, both directions; 3-D on 8x8x8 tiles
    (matrix-core gridding)."""
    rng = np.random.default_rng(41)
    n = {3: 24, 2: 40, 1: 64}[ndim]
    locs = rng.uniform(0, n, (1500, ndim)).astype(np.float32)
    locs[:5] = np.round(locs[:5])
    width = 4.0 if ndim == 3 else 6.0
    _run(dev, (n,) * ndim, locs, 2, width, "kaiser_bessel", "circular", 1, beta, "always", True, 42)


def test_grid_mfma_order_is_deterministic(dev):
    """With matrix-core gridding the plan orders each work item's touches
    by (block mask, point), so the summation order no longer depends on
    the order of the plan's atomics: two plans built for the same
    trajectory grid bit-identically (no split tiles here)."""
    import torchlinops_amd.functional as F
    from torchlinops_amd import config
    from torchlinops_amd.functional import _plan

    rng = np.random.default_rng(51)
    locs = to_dev(rng.uniform(0, 32, (6000, 3)).astype(np.float32), dev)
    y = to_dev(crandn(rng, (3, 6000)), dev)
    outs = []
    for _ in range(2):
        _plan.clear_plan_cache()
        with config.using(interp_plan="always"):
            outs.append(F.interpolate_adjoint(y, locs, (32, 32, 32), 4.0).clone())
    assert torch.equal(outs[0], outs[1])
    _plan.clear_plan_cache()


@pytest.mark.parametrize("n", [(32, 32, 32), (96, 80)])
def test_plan_sort_forms_agree(dev, n):
    """The plan's two ordering passes (gridding-item touches by block mask,
    subproblem points by LDS bank) size their LDS for the plan's caps:
    g_cap <= 2048 and max_sub <= 512 take the small forms.  With 2000
    uniform points no item or subproblem reaches those caps (about 30
    points per 8^3 bin and 250 per 32x32 bin), so g_cap 2048 / max_sub 512
    (the small forms, set explicitly: 3-D plans default to g_cap 4096) and
    g_cap 4096 / max_sub 2048 (the large forms) must give bit-identical
    interpolation and gridding."""
    import torchlinops_amd.functional as F
    from torchlinops_amd import config
    from torchlinops_amd.functional import _plan

    print(f"plan sort forms n={n}", flush=True)
    nd = len(n)
    rng = np.random.default_rng(71)
    locs = to_dev(rng.uniform(0, np.array(n), (2000, nd)).astype(np.float32), dev)
    x = to_dev(crandn(rng, (3, *n)), dev)
    y = to_dev(crandn(rng, (3, 2000)), dev)
    outs = []
    try:
        for knobs in ({"g_cap": 2048, "max_sub": 512}, {"g_cap": 4096, "max_sub": 2048}):
            with _plan._tuning_lock:
                _plan._tuning.pop(nd, None)
            _plan.set_tuning(nd, **knobs)
            with config.using(interp_plan="always"):
                u = F.interpolate(x, locs, 4.0)
                g = F.interpolate_adjoint(y, locs, n, 4.0)
            torch.cuda.synchronize()
            outs.append((u, g))
    finally:
        with _plan._tuning_lock:
            _plan._tuning.pop(nd, None)
        _plan.clear_plan_cache()
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
    ref = O.grid(y.cpu().numpy(), locs.cpu().numpy(), n, 4.0, "kaiser_bessel", 1, "circular",
                 1.0, "f64")
    assert rel_l2(outs[0][1].cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("n,cplx,nbatch", [
    ((24, 32, 40), True, 3), ((24, 32, 40), True, 11), ((24, 32, 40), False, 17),
    # 2-D matrix-core tiles (16 x 32); 80 leaves a half tile at the edge
    ((96, 80), True, 11), ((96, 80), False, 17)])
def test_grid_ws_records_match_direct_reads(dev, n, cplx, nbatch):
    """tnl_grid_ws (values re-packed point-major into the caller's scratch)
    against tnl_grid (values read from the coil-major input) through the C
    ABI, also with a partial last channel chunk.  2-D: the same kernel and
    summation order (no split tiles here), bit-identical.  3-D plans carry
    touch pairs, so tnl_grid_ws takes the one-wave kernel with its window
    records (one accumulator per tile instead of two wave pairs'): equal
    to float rounding and both within the bar of the float64 oracle.  A
    scratch below tnl_grid_scratch_bytes is rejected."""
    import ctypes
    from torchlinops_amd import _lib
    from torchlinops_amd.functional import _plan

    print(f"grid_ws n={n} cplx={cplx} nbatch={nbatch}", flush=True)
    rng = np.random.default_rng(61)
    nd = len(n)
    locs = to_dev(rng.uniform(-2, np.array(n) + 2, (5000, nd)).astype(np.float32), dev)
    y = crandn(rng, (nbatch, 5000)) if cplx else rng.standard_normal((nbatch, 5000)).astype(np.float32)
    yv = to_dev(y, dev)
    yr = torch.view_as_real(yv) if cplx else yv
    plan = _plan.build_plan(locs, nd, n, (4.0,) * nd, 0)
    assert plan.layout.grid_accum == 2
    h = _lib.lib()
    nbytes = ctypes.c_int64(0)
    _lib.check(h.tnl_grid_scratch_bytes(plan.layout, nbatch, int(cplx), ctypes.byref(nbytes)))
    ch = 8 if cplx else 16
    wave = plan.layout.off_tpm != 0
    assert wave == (nd == 3)
    assert nbytes.value == (-(-nbatch // ch) + (1 if wave else 0)) * 5000 * 64
    outs = []
    for scratch in (None, torch.empty(nbytes.value, dtype=torch.uint8, device=dev)):
        out = torch.empty(nbatch, *n, dtype=yv.dtype, device=dev)
        outr = torch.view_as_real(out) if cplx else out
        args = [yr.data_ptr(), locs.data_ptr(), outr.data_ptr(), nbatch, 5000, nd,
                _lib.i64_array(n), _lib.f32_array((4.0,) * nd), 0, 1, 0, int(cplx), 1.7,
                plan.layout, plan.ptr]
        if scratch is None:
            _lib.check(h.tnl_grid(*args, _lib.stream_of(yv)))
        else:
            _lib.check(h.tnl_grid_ws(*args, scratch.data_ptr(), nbytes.value, _lib.stream_of(yv)))
        torch.cuda.synchronize()
        outs.append(out)
    if wave:
        assert rel_l2(outs[1].cpu().numpy(), outs[0].cpu().numpy()) < 1e-6
    else:
        assert torch.equal(outs[0], outs[1])
    ref = O.grid(y, locs.cpu().numpy(), n, 4.0, "kaiser_bessel", 1, "circular", 1.7, "f64")
    assert rel_l2(outs[0].cpu().numpy(), ref) < TOL
    assert rel_l2(outs[1].cpu().numpy(), ref) < TOL
    small = torch.empty(nbytes.value - 16, dtype=torch.uint8, device=dev)
    with pytest.raises(ValueError, match="scratch"):
        _lib.check(h.tnl_grid_ws(*args, small.data_ptr(), small.numel(), _lib.stream_of(yv)))


UNGRID_FORM_CASES = [
    # grid_size, nbatch, pad, misaligned: 8^3 bins with 5 taps per axis stage
    # their 12^3 windows two channels interleaved when every axis has >= 12
    # cells, the last axis is even and the values are 16-byte aligned
    ((24, 32, 40), 5, "circular", False),  # interleaved, chunks of 2, 2, 1
    ((24, 32, 40), 3, "zero", False),  # interleaved, zero padding past the grid
    ((24, 32, 41), 4, "circular", False),  # odd last axis: channel planes
    ((24, 32, 40), 4, "circular", True),  # 8-byte aligned values: channel planes
]


@pytest.mark.parametrize("case", UNGRID_FORM_CASES, ids=lambda c: f"{c[0]}-nb{c[1]}-{c[2]}-mis{int(c[3])}")
def test_ungrid_staging_forms(dev, case):
    """Each staging form of the 3-D complex interpolation against the
    oracle, with points on integer coordinates, at the grid edges and
    outside it (both pads)."""
    import torchlinops_amd.functional as F
    from torchlinops_amd import config

    grid_size, nbatch, pad, misaligned = case
    print(f"ungrid staging form {case}", flush=True)
    rng = np.random.default_rng(91)
    locs = rng.uniform(-2, np.array(grid_size) + 2, (4000, 3)).astype(np.float32)
    locs[:9] = np.round(locs[:9])
    locs[9:12] = np.array(grid_size, dtype=np.float32) - 0.25
    x = crandn(rng, (nbatch, *grid_size))
    xd = to_dev(x, dev)
    if misaligned:
        buf = torch.zeros(xd.numel() + 1, dtype=xd.dtype, device=dev)
        buf[1:].copy_(xd.reshape(-1))
        xd = buf[1:].view(xd.shape)
        assert xd.data_ptr() % 16 == 8
    with config.using(interp_plan="always"):
        got = F.interpolate(xd, to_dev(locs, dev), 4.0, "kaiser_bessel", 1, pad, {"beta": BETA})
    got = got.cpu().numpy()
    ref32 = O.ungrid(x, locs, 4.0, "kaiser_bessel", 1, pad, BETA, "f32")
    ref64 = O.ungrid(x, locs, 4.0, "kaiser_bessel", 1, pad, BETA, "f64")
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print(f"  rel err f32 {e32:.2e} f64 {e64:.2e}", flush=True)
    assert e32 <= TOL and e64 <= TOL


@pytest.mark.parametrize("grid_size,nbatch,cplx", [
    ((32, 40, 36), 8, True), ((24, 28, 32), 5, True), ((32, 32, 32), 9, False),
    ((96, 80), 4, True), ((64, 72), 12, False)],
    ids=lambda v: str(v))
def test_point_major_interpolation(dev, grid_size, nbatch, cplx):
    """Planned interpolation through the point-major scratch and one
    transpose (tnl_ungrid_ws) writes the same bits as the channel-major
    stores and matches the oracle; points in random order make the plan
    mark the trajectory scattered, so "auto" takes that form too.  Points
    without taps (non-finite coordinates) come out zero."""
    import torchlinops_amd.functional as F
    from torchlinops_amd import config
    from torchlinops_amd.functional import _plan

    print(f"point-major interpolation {grid_size} nb={nbatch} cplx={cplx}", flush=True)
    nd = len(grid_size)
    rng = np.random.default_rng(97)
    locs = rng.uniform(-2, np.array(grid_size) + 2, (5000, nd)).astype(np.float32)
    locs[:9] = np.round(locs[:9])
    bad = [11, 4000]
    locs_bad = locs.copy()
    locs_bad[bad] = np.nan
    shape = (nbatch, *grid_size)
    x = crandn(rng, shape) if cplx else rng.standard_normal(shape).astype(np.float32)
    xd, ld = to_dev(x, dev), to_dev(locs_bad, dev)
    outs = {}
    for pm in ("never", "always", "auto"):
        with config.using(interp_plan="always", interp_point_major=pm):
            outs[pm] = F.interpolate(xd, ld, 4.0, "kaiser_bessel", 1, "circular", {"beta": BETA})
    torch.cuda.synchronize()
    _plan.clear_plan_cache()
    assert torch.equal(outs["never"], outs["always"])
    assert torch.equal(outs["never"], outs["auto"])
    got = outs["always"].cpu().numpy()
    assert (got[:, bad] == 0).all()
    keep = np.setdiff1d(np.arange(locs.shape[0]), bad)
    ref = O.ungrid(x, locs[keep], 4.0, "kaiser_bessel", 1, "circular", BETA, "f64")
    e = rel_l2(got[:, keep], ref)
    print(f"  rel err f64 {e:.2e}", flush=True)
    assert e <= TOL


def test_c2_full_size_properties(dev):
    """The headline configuration at full size: 8 coils, 256^3 complex64,
    20000 x 512 radial points, KB w=4.  The planned kernels (LDS-staged
    interpolation, matrix-core gridding with point records) against the
    plan-free kernels, adjointness in float64, linearity, the oracle on
    4096 sampled points of the interpolation and on 124 sampled cells of
    the gridding."""
    import torchlinops_amd.functional as F
    from torchlinops_amd import config

    print("c2 full size: 8 x 256^3, 20000 x 512 radial", flush=True)
    n, C, spokes, nread = 256, 8, 20000, 512
    g = torch.Generator().manual_seed(21)
    u = torch.randn(spokes, 3, generator=g)
    u = u / u.norm(dim=1, keepdim=True)
    r = -n / 2 + torch.arange(nread, dtype=torch.float32) * (n / nread)
    locs = torch.remainder(r[None, :, None] * u[:, None, :] + n / 2, n).reshape(-1, 3)
    locs = locs.to(dev).contiguous()
    gd = torch.Generator(device=dev).manual_seed(22)
    x = torch.randn(C, n, n, n, dtype=torch.complex64, device=dev, generator=gd)
    y = torch.randn(C, locs.shape[0], dtype=torch.complex64, device=dev, generator=gd)
    kp = {"beta": BETA}
    with config.using(interp_plan="always"):
        ax = F.interpolate(x, locs, 4.0, kernel_params=kp)
        ahy = F.interpolate_adjoint(y, locs, (n, n, n), 4.0, kernel_params=kp)
        ax2 = F.interpolate(2 * x, locs, 4.0, kernel_params=kp)
    with config.using(interp_plan="never"):
        ax_d = F.interpolate(x, locs, 4.0, kernel_params=kp)
        ahy_d = F.interpolate_adjoint(y, locs, (n, n, n), 4.0, kernel_params=kp)

    def rel(a, b):
        return float(torch.linalg.vector_norm((a - b).to(torch.complex128))
                     / torch.linalg.vector_norm(b.to(torch.complex128)))

    e_u, e_g = rel(ax, ax_d), rel(ahy, ahy_d)
    lhs = torch.vdot(y.reshape(-1).to(torch.complex128), ax.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(ahy.reshape(-1).to(torch.complex128), x.reshape(-1).to(torch.complex128))
    e_adj = float(abs(lhs - rhs) / abs(lhs))
    e_lin = rel(ax2, 2 * ax)
    print(f"plan vs direct: ungrid {e_u:.2e} grid {e_g:.2e}; adjointness {e_adj:.2e}; "
          f"linearity {e_lin:.2e}", flush=True)
    assert e_u <= TOL and e_g <= TOL and e_adj <= TOL and e_lin <= TOL
    idx = torch.randperm(locs.shape[0], generator=g)[:4096]
    want = O.ungrid(x.cpu().numpy(), locs[idx.to(dev)].cpu().numpy(), 4.0, beta=BETA,
                    precision="f64", nthreads=8)
    got = ax[:, idx.to(dev)].cpu().numpy()
    e_o = rel_l2(got, want)
    print(f"oracle on 4096 sampled points: {e_o:.2e}", flush=True)
    assert e_o <= TOL
    # the gridding output against the float64 oracle on sampled cells: the
    # densest 4^3 block (the centre every spoke crosses, ~10^5 touches per
    # cell) and 60 random cells.  Gridding sums over points, so the points
    # whose footprints reach those cells (circular distance <= w/2 on every
    # axis) give the whole value there; coils 0 and 7.
    rs = np.random.default_rng(23)
    centre = np.stack(np.meshgrid(*[np.arange(126, 130)] * 3, indexing="ij"), -1).reshape(-1, 3)
    cells = np.concatenate([centre, rs.integers(0, n, (60, 3))])
    near = torch.zeros(locs.shape[0], dtype=torch.bool, device=dev)
    for c in torch.from_numpy(cells).to(dev, torch.float32):
        d = (locs - c).abs()
        near |= (torch.minimum(d, n - d) <= 2.0).all(dim=1)
    sel = near.nonzero().squeeze(1)
    print(f"oracle on {len(cells)} grid cells from {sel.numel()} points", flush=True)
    want = O.grid(y[[0, 7]][:, sel].cpu().numpy(), locs[sel].cpu().numpy(), (n, n, n), 4.0,
                  "kaiser_bessel", 1, "circular", BETA, "f64")
    at = (slice(None),) + tuple(cells.T)
    e_c = rel_l2(ahy[[0, 7]].cpu().numpy()[at], want[at])
    print(f"oracle on sampled grid cells: {e_c:.2e}", flush=True)
    assert e_c <= TOL


F64_CASES = [
    ((40,), 4.0, "kaiser_bessel", 1, "circular"),
    ((20, 24), 4.0, "kaiser_bessel", 1, "zero"),
    ((20, 24), (3.0, 5.0), "spline", 1, "circular"),
    ((10, 12, 14), 4.0, "kaiser_bessel", 1, "circular"),
    ((10, 12, 14), 3.0, "spline", 2, "zero"),
]


@pytest.mark.parametrize("grid_size,width,kernel,norm,pad", F64_CASES)
def test_float64_interpolation_on_device(dev, grid_size, width, kernel, norm, pad):
    """float64 / complex128 device tensors compute in double (tnl_ungrid_f64
    / tnl_grid_f64): within 1e-12 of the float64 oracle for the spline and
    within the oracle's A&S window accuracy (5e-7) for Kaiser-Bessel, and
    within 1e-12 of the package's host path (exact I0 on both); adjoint to
    1e-12."""
    import torchlinops_amd.functional as F

    rng = np.random.default_rng(len(grid_size) * 11 + norm)
    nd = len(grid_size)
    x = (rng.standard_normal((3, *grid_size)) + 1j * rng.standard_normal((3, *grid_size)))
    locs = rng.uniform(-3, 3 + max(grid_size), (700, nd))
    locs[:60] = np.round(locs[:60])
    y = rng.standard_normal((3, 700)) + 1j * rng.standard_normal((3, 700))
    beta = float(O.nufft_beta(4.0, 2.0)) if kernel == "kaiser_bessel" else 1.0
    kw = dict(kernel=kernel, norm=norm, pad_mode=pad, kernel_params={"beta": beta})
    xt, lt, yt = (torch.from_numpy(a).to(dev) for a in (x, locs, y))
    u = F.interpolate(xt, lt, width, **kw)
    g = F.interpolate_adjoint(yt, lt, grid_size, width, **kw)
    assert u.dtype == torch.complex128 and g.dtype == torch.complex128
    ok = dict(kernel=kernel, norm=norm, pad_mode=pad, beta=beta, precision="f64")
    tol = 5e-7 if kernel == "kaiser_bessel" else 1e-12
    assert rel_l2(u.cpu().numpy(), O.ungrid(x, locs, width, **ok)) < tol
    assert rel_l2(g.cpu().numpy(), O.grid(y, locs, grid_size, width, **ok)) < tol
    uh = F.interpolate(torch.from_numpy(x), torch.from_numpy(locs), width, **kw)
    gh = F.interpolate_adjoint(torch.from_numpy(y), torch.from_numpy(locs), grid_size, width, **kw)
    assert rel_l2(u.cpu().numpy(), uh.numpy()) < 1e-12
    assert rel_l2(g.cpu().numpy(), gh.numpy()) < 1e-12
    lhs = np.vdot(y, u.cpu().numpy())
    rhs = np.vdot(g.cpu().numpy(), x)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
