"""Many problems in one persistent launch: BatchedInterp against the
per-problem functional calls and the oracle; SegmentedCG over
BatchedNUFFTNormal against the float64 oracle CG per problem; graph
replays of the captured CG against the eager iterations."""

import numpy as np
import pytest
import torch

import torchlinops_amd as tl
import torchlinops_amd.functional as F
from torchlinops_amd import config
from oracle import oracle as O
from _util import crandn, rel_l2, to_dev, vd_spiral_2d

pytestmark = pytest.mark.gpu


def _problems(rng, ndim, n, grids, mrange, cmax, widths):
    out = []
    for _ in range(n):
        N = int(rng.choice(grids))
        M = int(np.exp(rng.uniform(np.log(mrange[0]), np.log(mrange[1]))))
        C = int(rng.integers(1, cmax + 1))
        w = float(rng.choice(widths))
        locs = rng.uniform(0, N, (M, ndim)).astype(np.float32)
        out.append(dict(grid=(N,) * ndim, M=M, C=C, w=w, beta=float(rng.uniform(5, 12)),
                        locs=locs))
    return out


CFG = {
    2: dict(n=7, grids=[32, 48, 64, 96], mrange=(300, 20000), cmax=6, widths=[4.0, 5.0, 6.0]),
    3: dict(n=3, grids=[16, 24, 32], mrange=(300, 6000), cmax=3, widths=[3.0, 4.0]),
}


@pytest.mark.parametrize("ndim", [2, 3])
def test_batched_interp_matches_single_and_oracle(dev, ndim):
    rng = np.random.default_rng(500 + ndim)
    P = _problems(rng, ndim, **CFG[ndim])
    locs = [to_dev(p["locs"], dev) for p in P]
    B = F.BatchedInterp(locs, [p["grid"] for p in P], [p["w"] for p in P], [p["C"] for p in P],
                        kernel_params=[{"beta": p["beta"]} for p in P])
    xs = [crandn(rng, (p["C"], *p["grid"])) for p in P]
    ys = [crandn(rng, (p["C"], p["M"])) for p in P]
    gflat = B.new_grid()
    for v, x in zip(B.grid_views(gflat), xs):
        v.copy_(to_dev(x, dev))
    pflat = B.new_points()
    for v, y in zip(B.pts_views(pflat), ys):
        v.copy_(to_dev(y, dev))
    u = B.ungrid(gflat)
    g = B.grid(pflat)
    with config.using(interp_plan="always"):
        for i, p in enumerate(P):
            kp = {"beta": p["beta"]}
            su = F.interpolate(to_dev(xs[i], dev), locs[i], p["w"], kernel_params=kp)
            sg = F.interpolate_adjoint(to_dev(ys[i], dev), locs[i], p["grid"], p["w"],
                                       kernel_params=kp)
            bu = B.pts_views(u)[i]
            bg = B.grid_views(g)[i]
            # same per-point tap order; split tiles combine with atomics
            assert rel_l2(bu.cpu().numpy(), su.cpu().numpy()) < 1e-6, i
            assert rel_l2(bg.cpu().numpy(), sg.cpu().numpy()) < 1e-6, i
            ou = O.ungrid(xs[i], p["locs"], p["w"], beta=p["beta"], precision="f64")
            og = O.grid(ys[i], p["locs"], p["grid"], p["w"], beta=p["beta"], precision="f64")
            assert rel_l2(bu.cpu().numpy(), ou) < 1e-5, i
            assert rel_l2(bg.cpu().numpy(), og) < 1e-5, i


def test_batched_rejects_mixed_tiling_and_wide_3d(dev):
    rng = np.random.default_rng(9)
    l = [to_dev(rng.uniform(0, 16, (100, 3)).astype(np.float32), dev)]
    with pytest.raises(ValueError):
        F.BatchedInterp(l, [(16, 16, 16)], [6.0], [1])  # 7 taps per axis in 3-D
    with pytest.raises(ValueError):
        F.BatchedInterp(l * 2, [(16, 16), (16, 16, 16)], [4.0, 4.0], [1, 1])


def test_batched_wide_3d_on_matrix_cores(dev):
    """3-D batches wider than 5 taps per axis (widths 5 and 6: 6 and 7
    taps) whose plans grid on the matrix cores (grids above 16 per axis):
    batched interpolation and gridding equal the per-problem calls and the
    float64 oracle.  (With LDS-accumulator plans they are refused, above.)"""
    rng = np.random.default_rng(710)
    P = _problems(rng, 3, n=3, grids=[24, 32], mrange=(2000, 8000), cmax=3,
                  widths=[5.0, 6.0])
    locs = [to_dev(p["locs"], dev) for p in P]
    B = F.BatchedInterp(locs, [p["grid"] for p in P], [p["w"] for p in P], [p["C"] for p in P],
                        kernel_params=[{"beta": p["beta"]} for p in P])
    assert B.layout.grid_mfma == 1 and B.layout.nt == 8
    xs = [crandn(rng, (p["C"], *p["grid"])) for p in P]
    ys = [crandn(rng, (p["C"], p["M"])) for p in P]
    gflat = B.new_grid()
    for v, x in zip(B.grid_views(gflat), xs):
        v.copy_(to_dev(x, dev))
    pflat = B.new_points()
    for v, y in zip(B.pts_views(pflat), ys):
        v.copy_(to_dev(y, dev))
    u = B.ungrid(gflat)
    g = B.grid(pflat)
    with config.using(interp_plan="always"):
        for i, p in enumerate(P):
            kp = {"beta": p["beta"]}
            su = F.interpolate(to_dev(xs[i], dev), locs[i], p["w"], kernel_params=kp)
            sg = F.interpolate_adjoint(to_dev(ys[i], dev), locs[i], p["grid"], p["w"],
                                       kernel_params=kp)
            bu = B.pts_views(u)[i]
            bg = B.grid_views(g)[i]
            assert rel_l2(bu.cpu().numpy(), su.cpu().numpy()) < 1e-6, i
            assert rel_l2(bg.cpu().numpy(), sg.cpu().numpy()) < 1e-6, i
            ou = O.ungrid(xs[i], p["locs"], p["w"], beta=p["beta"], precision="f64")
            og = O.grid(ys[i], p["locs"], p["grid"], p["w"], beta=p["beta"], precision="f64")
            assert rel_l2(bu.cpu().numpy(), ou) < 1e-5, i
            assert rel_l2(bg.cpu().numpy(), og) < 1e-5, i


def _cg_problems(rng, n=3):
    P = []
    for i in range(n):
        N = int(rng.choice([24, 32, 40]))
        M = int(rng.integers(1500, 4000))
        P.append(dict(N=N, C=int(rng.integers(1, 4)), w=float(rng.choice([4.0, 5.0, 6.0])),
                      k=vd_spiral_2d(M, N, rng)))
    return P


def _batched_cg(dev, P, lam=0.0):
    A = tl.alg.BatchedNUFFTNormal([to_dev(p["k"], dev) for p in P], [(p["N"],) * 2 for p in P],
                                  [p["C"] for p in P], oversamp=2.0,
                                  widths=[p["w"] for p in P])
    cg = tl.alg.SegmentedCG(A, A.seg_elems, dev, lam=lam)
    return A, cg


def _lam(A, y, frac=0.05, iters=30):
    """frac x the largest eigenvalue over the problems (power iteration):
    keeps (A + lam I) conditioned near 1/frac, so float32 and float64 CG
    iterates stay close.  Unregularised spiral normal equations here have
    lambda_max / Rayleigh quotient ~ 60, enough that 1e-6 operator noise
    moves the 8th iterate by percent."""
    v = y / torch.linalg.vector_norm(y)
    lmax = 0.0
    for _ in range(iters):
        w = A(v)
        lmax = float(torch.linalg.vector_norm(w))
        v = w / lmax
    return frac * lmax


def test_segmented_cg_kernels_on_diagonal_systems(dev):
    """CG kernels alone: per-segment diagonal SPD operators with spectra in
    [1, 10]; float32 device CG against float64 numpy CG per segment."""
    rng = np.random.default_rng(79)
    sizes = [1, 700, 16384 * 3 + 5, 40000, 3]
    d = [rng.uniform(1, 10, n) for n in sizes]
    y = [crandn(rng, (n,)) for n in sizes]
    dflat = to_dev(np.concatenate(d).astype(np.float32), dev)
    normal = lambda x, out: torch.mul(x, dflat, out=out)  # noqa: E731
    cg = tl.alg.SegmentedCG(normal, sizes, dev)
    yflat = to_dev(np.concatenate(y), dev)
    iters = 12
    x = cg.solve(yflat, iters=iters, use_graph=False).cpu().numpy()
    off = 0
    for di, yi, n in zip(d, y, sizes):
        want = O.conjugate_gradients(lambda v, di=di: di * v, yi.astype(np.complex128),
                                     max_num_iters=iters)
        assert rel_l2(x[off:off + n], want) < 1e-5, n
        off += n
    # the one-element segment is solved exactly in one step
    assert rel_l2(x[:1], y[0] / d[0]) < 1e-6


def test_segmented_cg_matches_oracle_cg(dev):
    rng = np.random.default_rng(77)
    P = _cg_problems(rng)
    A, _ = _batched_cg(dev, P)
    ys = [crandn(rng, (p["C"], p["N"], p["N"])) for p in P]
    y = A.new_image()
    for v, yy in zip(A.image_views(y), ys):
        v.copy_(to_dev(yy, dev))
    # one application against the oracle normal operator
    Ay = A(y)
    for i, p in enumerate(P):
        want = O.nufft_adjoint(O.nufft(ys[i].astype(np.complex128), p["k"], 2.0, p["w"]),
                               p["k"], (p["N"],) * 2, 2.0, p["w"])
        assert rel_l2(A.image_views(Ay)[i].cpu().numpy(), want) < 1e-5, i
    lam = _lam(A, y)
    cg = tl.alg.SegmentedCG(A, A.seg_elems, dev, lam=lam)
    iters = 8
    x = cg.solve(y, iters=iters, use_graph=False)
    res = cg.relative_residuals()
    for i, p in enumerate(P):
        An = lambda v, p=p: (O.nufft_adjoint(O.nufft(v, p["k"], 2.0, p["w"]), p["k"],  # noqa: E731
                                             (p["N"],) * 2, 2.0, p["w"]) + lam * v)
        want = O.conjugate_gradients(An, ys[i].astype(np.complex128), max_num_iters=iters)
        err = rel_l2(A.image_views(x)[i].cpu().numpy(), want)
        print(f"problem {i}: rel-L2 vs float64 CG after {iters} iterations = {err:.2e}, "
              f"relative residual {float(res[A.position[i]]):.2e}")
        # the float32 operator is within 1e-5 of float64 (checked above);
        # CG carries that error times the condition number (~20 here)
        assert err < 2e-3, i


def test_segmented_cg_graph_replays_match_eager(dev):
    rng = np.random.default_rng(78)
    P = _cg_problems(rng, 4)
    A, _ = _batched_cg(dev, P)
    y = A.new_image()
    y.copy_(to_dev(crandn(rng, (A.numel,)), dev))
    cg = tl.alg.SegmentedCG(A, A.seg_elems, dev, lam=_lam(A, y))
    eager = cg.solve(y, iters=5, use_graph=False).clone()
    rs_eager = cg.rs.clone()
    for rep in range(3):
        got = cg.solve(y, iters=5, use_graph=True).clone()
        torch.cuda.synchronize()
        # eager work between replays, as a caller would do
        _ = (got.abs() ** 2).sum().item()
        assert rel_l2(got.cpu().numpy(), eager.cpu().numpy()) < 1e-5, rep
        assert torch.allclose(cg.rs, rs_eager, rtol=1e-4), rep
    assert (cg.relative_residuals() < 1).all()


def test_autotuner_picks_a_candidate_and_keeps_results(dev, tmp_path):
    from torchlinops_amd.functional import autotune as AT

    rng = np.random.default_rng(90)
    P = _problems(rng, 2, n=4, grids=[64, 96], mrange=(2000, 8000), cmax=3, widths=[4.0, 6.0])
    locs = [to_dev(p["locs"], dev) for p in P]
    args = ([p["grid"] for p in P], [p["w"] for p in P], [p["C"] for p in P])
    kp = [{"beta": p["beta"]} for p in P]
    best, timings = AT.autotune_batch(locs, *args, kernel_params=kp, reps=2)
    assert best in [t[0] for t in timings] and len(timings) >= 3
    assert all(ms > 0 for _, ms in timings)
    again, t2 = AT.autotune_batch(locs, *args, kernel_params=kp)
    assert again == best and t2 == []  # memoised
    AT.save_cache(tmp_path / "tune.json")
    AT.load_cache(tmp_path / "tune.json")
    B0 = F.BatchedInterp(locs, *args, kernel_params=kp)
    B1 = F.BatchedInterp(locs, *args, kernel_params=kp, tuning=best)
    g = B0.new_grid().normal_()
    assert rel_l2(B1.ungrid(g).cpu().numpy(), B0.ungrid(g).cpu().numpy()) < 1e-6
    pts = B0.new_points().normal_()
    assert rel_l2(B1.grid(pts).cpu().numpy(), B0.grid(pts).cpu().numpy()) < 1e-6
    # single problem: applies the winner as the 2-D tuning
    from torchlinops_amd.functional import _plan

    saved = _plan.get_tuning(2)
    try:
        b1, t1 = AT.autotune_interp(locs[0], P[0]["grid"], P[0]["w"], P[0]["C"],
                                    kernel_params=kp[0], reps=2)
        assert b1 in [t[0] for t in t1]
        assert all(_plan.get_tuning(2)[k] == v for k, v in b1.items())
    finally:
        with _plan._tuning_lock:
            _plan._tuning[2] = saved
        _plan.clear_plan_cache()


def test_batched_equals_per_problem_launches(dev):
    rng = np.random.default_rng(91)
    P = _problems(rng, 2, n=5, grids=[48, 64], mrange=(1000, 6000), cmax=4, widths=[4.0, 6.0])
    B = F.BatchedInterp([to_dev(p["locs"], dev) for p in P], [p["grid"] for p in P],
                        [p["w"] for p in P], [p["C"] for p in P],
                        kernel_params=[{"beta": p["beta"]} for p in P])
    g = B.new_grid().normal_()
    pts = B.new_points().normal_()
    a, b = B.ungrid(g), B.ungrid_each(g, B.new_points())
    assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6
    a, b = B.grid(pts), B.grid_each(pts, B.new_grid())
    assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("ndim", [2, 3])
def test_batched_matrix_core_gridding(dev, ndim):
    """A batch whose plans all grid on the matrix cores (2-D grids above
    32 x 64, 3-D above 16 per axis) runs the batched matrix-core kernel:
    equal to the per-problem calls and to the float64 oracle, split tiles
    included (one dense problem), channel chunks included (9 coils)."""
    rng = np.random.default_rng(700 + ndim)
    grids = [80, 96, 128] if ndim == 2 else [24, 32]
    P = _problems(rng, ndim, n=5, grids=grids, mrange=(2000, 20000), cmax=9,
                  widths=[4.0, 5.0, 6.0] if ndim == 2 else [3.0, 4.0])
    dense = P[0]
    dense["locs"] = (rng.standard_normal((30000, ndim)) * 1.5 + 10).astype(np.float32)
    dense["M"] = 30000
    locs = [to_dev(p["locs"], dev) for p in P]
    B = F.BatchedInterp(locs, [p["grid"] for p in P], [p["w"] for p in P], [p["C"] for p in P],
                        kernel_params=[{"beta": p["beta"]} for p in P])
    assert B.layout.grid_mfma == 1
    ys = [crandn(rng, (p["C"], p["M"])) for p in P]
    pflat = B.new_points()
    for v, y in zip(B.pts_views(pflat), ys):
        v.copy_(to_dev(y, dev))
    g = B.grid(pflat)
    with config.using(interp_plan="always"):
        for i, p in enumerate(P):
            kp = {"beta": p["beta"]}
            sg = F.interpolate_adjoint(to_dev(ys[i], dev), locs[i], p["grid"], p["w"],
                                       kernel_params=kp)
            bg = B.grid_views(g)[i]
            assert rel_l2(bg.cpu().numpy(), sg.cpu().numpy()) < 1e-6, i
            og = O.grid(ys[i], p["locs"], p["grid"], p["w"], beta=p["beta"], precision="f64")
            assert rel_l2(bg.cpu().numpy(), og) < 1e-5, i
