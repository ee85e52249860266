"""Pins the oracle (oracle/) before any kernel is compared against it.

Every oracle function is checked against something that does not share its
code: a closed form, an independent library routine (scipy.special.i0,
numpy.interp, numpy sliding_window_view, numpy.fft), brute force over every
grid cell, or an algebraic invariant.  A few tests plant a mistake in a
copy of the oracle's inputs and check the comparison notices.
"""

from itertools import product
from math import prod

import numpy as np
import pytest
from scipy.special import i0

from oracle import oracle as O

BETA_NUFFT = O.nufft_beta(4.0, 2.0)  # pi * sqrt(8.2) = 8.996...


def kb_exact(x, beta):
    x = np.asarray(x, dtype=np.float64)
    return np.where(np.abs(x) <= 1, i0(beta * np.sqrt(np.clip(1 - x * x, 0, None))), 0.0)


# ------------------------------------------------------------------ KB
@pytest.mark.parametrize("beta", [1.0, 2.5, BETA_NUFFT, 14.0])
def test_kb_matches_scipy_i0(beta):
    x = np.linspace(-1, 1, 4001)
    ref = kb_exact(x, beta)
    got64 = O.kb(x, beta, "f64")
    x32 = x.astype(np.float32)
    got32 = O.kb(x32, beta, "f32")
    ref32 = kb_exact(x32.astype(np.float64), beta)
    # Abramowitz & Stegun 9.8.1/9.8.2: |rel err| < 1.9e-7
    assert np.max(np.abs(got64 - ref) / ref) < 2.5e-7
    assert np.max(np.abs(got32 - ref32) / ref32) < 2e-6


def test_kb_edges_and_centre():
    for beta in (1.0, BETA_NUFFT):
        assert O.kb([1.0, -1.0], beta, "f64").tolist() == [1.0, 1.0]
        assert O.kb([1.0, -1.0], beta, "f32").tolist() == [1.0, 1.0]
        assert abs(O.kb([0.0], beta, "f64")[0] / i0(beta) - 1) < 2.5e-7


def test_nufft_beta_closed_form():
    assert abs(O.nufft_beta(4.0, 2.0) - np.pi * np.sqrt(8.2)) < 1e-12
    assert abs(O.nufft_beta(4.0, 1.25) - np.pi * np.sqrt((4 / 1.25 * 0.75) ** 2 - 0.8)) < 1e-12


# ------------------------------------------------------- interpolation
def brute_weights(t, grid_size, width, beta, kernel="kaiser_bessel", pad="circular"):
    """Weight of every grid cell for a point at t: sum over unwrapped integer
    taps g with |t - g| <= w/2 that land on the cell (brute force, exact I0)."""
    W = np.zeros(grid_size)
    axes = []
    for d, n in enumerate(grid_size):
        r = width[d] / 2
        gs = [g for g in range(int(np.floor(t[d] - r)) - 1, int(np.ceil(t[d] + r)) + 2)
              if abs(t[d] - g) <= r]
        axes.append(gs)
    for gvec in product(*axes):
        w = 1.0
        cell = []
        ok = True
        for d, g in enumerate(gvec):
            x = (t[d] - g) / (width[d] / 2)
            w *= kb_exact(x, beta) if kernel == "kaiser_bessel" else max(1 - abs(x), 0)
            if pad == "zero" and not (0 <= g < grid_size[d]):
                ok = False
            cell.append(g % grid_size[d])
        if ok:
            W[tuple(cell)] += w
    return W


@pytest.mark.parametrize("grid_size", [(7,), (6, 5), (5, 4, 6), (3, 3)])
@pytest.mark.parametrize("pad", ["circular", "zero"])
@pytest.mark.parametrize("kernel", ["kaiser_bessel", "spline"])
def test_ungrid_grid_vs_brute_force(grid_size, pad, kernel):
    rng = np.random.default_rng(1)
    ndim = len(grid_size)
    width = tuple(rng.choice([2.0, 3.0, 4.0, 3.5]) for _ in range(ndim))
    npts = 9
    locs = rng.uniform(-2, np.array(grid_size) + 2, size=(npts, ndim))
    locs[0] = np.round(locs[0])  # exact integer: w+1 taps with edge weight 1
    vals = rng.standard_normal((2, *grid_size)) + 1j * rng.standard_normal((2, *grid_size))
    beta = 2.3
    got = O.ungrid(vals, locs, width, kernel=kernel, pad_mode=pad, beta=beta, precision="f64")
    ys = rng.standard_normal((2, npts)) + 1j * rng.standard_normal((2, npts))
    got_g = O.grid(ys, locs, grid_size, width, kernel=kernel, pad_mode=pad, beta=beta,
                   precision="f64")
    want = np.zeros((2, npts), complex)
    want_g = np.zeros((2, *grid_size), complex)
    for p in range(npts):
        W = brute_weights(locs[p], grid_size, width, beta, kernel, pad)
        want[:, p] = (vals * W).reshape(2, -1).sum(-1)
        want_g += ys[:, p].reshape(2, *([1] * ndim)) * W
    assert np.allclose(got, want, rtol=1e-6, atol=1e-9)
    assert np.allclose(got_g, want_g, rtol=1e-6, atol=1e-9)


def test_brute_force_notices_swapped_axes():
    """Planted mistake: feeding the oracle transposed coordinates must fail."""
    rng = np.random.default_rng(2)
    grid_size = (6, 5)
    locs = rng.uniform(0, 5, size=(5, 2))
    vals = rng.standard_normal(grid_size)
    wrong = O.ungrid(vals, locs[:, ::-1].copy(), 4.0, beta=2.0, precision="f64")
    want = np.array([(vals * brute_weights(l, grid_size, (4.0, 4.0), 2.0)).sum() for l in locs])
    assert not np.allclose(wrong, want, rtol=1e-3)


def test_integer_location_known_answer():
    """t = 4.0, w = 4: five taps with weights 1, I0(b sqrt(3)/2), I0(b),
    I0(b sqrt(3)/2), 1 (the case of the reference's test_edgecases.py)."""
    beta = 1.0
    vals = np.arange(1, 9, dtype=np.float64)
    locs = np.array([[4.0]])
    w = [1.0, i0(beta * np.sqrt(0.75)), i0(beta), i0(beta * np.sqrt(0.75)), 1.0]
    want = sum(wk * vals[2 + k] for k, wk in enumerate(w))
    got64 = O.ungrid(vals, locs, 4.0, beta=beta, precision="f64")
    got32 = O.ungrid(vals.astype(np.float32), locs, 4.0, beta=beta, precision="f32")
    assert abs(got64[0] / want - 1) < 5e-7
    assert abs(got32[0] / want - 1) < 2e-6
    g = O.grid(np.array([1.0]), locs, (8,), 4.0, beta=beta, precision="f64")
    assert np.allclose(g[2:7], w, rtol=5e-7) and np.all(g[[0, 1, 7]] == 0)


def test_spline_width2_is_linear_interpolation():
    rng = np.random.default_rng(3)
    n = 16
    vals = rng.standard_normal(n)
    t = rng.uniform(0, n, 50)
    got = O.ungrid(vals, t[:, None], 2.0, kernel="spline", precision="f64")
    want = np.interp(t, np.arange(n + 1), np.append(vals, vals[0]))
    assert np.allclose(got, want, atol=1e-12)
    # 2-D: bilinear
    v2 = rng.standard_normal((8, 9))
    pts = rng.uniform(0, 7, (20, 2))
    got2 = O.ungrid(v2, pts, 2.0, kernel="spline", precision="f64")
    x0, y0 = np.floor(pts[:, 0]).astype(int), np.floor(pts[:, 1]).astype(int)
    fx, fy = pts[:, 0] - x0, pts[:, 1] - y0
    want2 = ((1 - fx) * (1 - fy) * v2[x0 % 8, y0 % 9] + fx * (1 - fy) * v2[(x0 + 1) % 8, y0 % 9]
             + (1 - fx) * fy * v2[x0 % 8, (y0 + 1) % 9] + fx * fy * v2[(x0 + 1) % 8, (y0 + 1) % 9])
    assert np.allclose(got2, want2, atol=1e-12)


def test_delta_grid_gives_separable_weights():
    grid_size = (10, 9, 8)
    vals = np.zeros(grid_size)
    vals[4, 4, 4] = 1.0
    t = np.array([[4.3, 3.1, 5.7]])
    beta = BETA_NUFFT
    got = O.ungrid(vals, t, 4.0, beta=beta, precision="f64")
    want = np.prod([kb_exact((t[0, d] - 4) / 2, beta) for d in range(3)])
    assert abs(got[0] / want - 1) < 1e-6


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_adjointness_f64(ndim):
    rng = np.random.default_rng(4 + ndim)
    grid_size = (9, 8, 7)[:ndim]
    locs = rng.uniform(-3, 12, (40, ndim))
    x = rng.standard_normal((3, *grid_size)) + 1j * rng.standard_normal((3, *grid_size))
    y = rng.standard_normal((3, 40)) + 1j * rng.standard_normal((3, 40))
    Ux = O.ungrid(x, locs, 3.7, beta=5.0, precision="f64")
    Gy = O.grid(y, locs, grid_size, 3.7, beta=5.0, precision="f64")
    lhs = np.vdot(y, Ux)
    rhs = np.vdot(Gy, x)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_f32_oracle_close_to_f64():
    rng = np.random.default_rng(5)
    grid_size = (32, 32)
    locs = rng.uniform(0, 32, (500, 2)).astype(np.float32)
    x = (rng.standard_normal((2, *grid_size)) + 1j * rng.standard_normal((2, *grid_size))).astype(np.complex64)
    a = O.ungrid(x, locs, 4.0, beta=BETA_NUFFT, precision="f32")
    b = O.ungrid(x, locs, 4.0, beta=BETA_NUFFT, precision="f64")
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6


# ------------------------------------------------------- unfold / fold
def sliding_blocks(x, block, stride):
    ndim = len(block)
    v = np.lib.stride_tricks.sliding_window_view(x, block, axis=tuple(range(x.ndim - ndim, x.ndim)))
    slc = (Ellipsis,) + tuple(slice(None, None, s) for s in stride) + (slice(None),) * ndim
    return v[slc]


@pytest.mark.parametrize(
    "shape,block,stride",
    [((2, 17), (5,), (3,)), ((1, 12, 11), (4, 3), (2, 2)), ((2, 9, 10, 8), (3, 4, 2), (2, 3, 1)),
     ((1, 128, 128), (8, 8), (4, 4))],
)
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64])
def test_unfold_vs_sliding_window(shape, block, stride, dtype):
    rng = np.random.default_rng(6)
    x = rng.standard_normal(shape).astype(dtype)
    if np.iscomplexobj(x):
        x = x + 1j * rng.standard_normal(shape).astype(np.float32)
        x = x.astype(dtype)
    got = O.unfold(x, block, stride)
    want = sliding_blocks(x, block, stride)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape,block,stride", [((2, 13), (4,), (3,)), ((1, 9, 8), (3, 3), (2, 1)),
                                                ((1, 7, 6, 8), (3, 2, 4), (2, 2, 2))])
def test_fold_brute_force_and_adjoint(shape, block, stride):
    rng = np.random.default_rng(7)
    ndim = len(block)
    im = shape[-ndim:]
    nbl = O.get_nblocks(im, block, stride)
    xb = rng.standard_normal((shape[0], *nbl, *block))
    got = O.fold(xb, im, block, stride)
    want = np.zeros(shape)
    for n in range(shape[0]):
        for b in product(*(range(k) for k in nbl)):
            slc = tuple(slice(bi * s, bi * s + bl) for bi, s, bl in zip(b, stride, block))
            want[(n, *slc)] += xb[(n, *b)]
    assert np.array_equal(got, want)  # same order of additions -> identical
    x = rng.standard_normal(shape)
    assert abs(np.vdot(xb, O.unfold(x, block, stride)) - np.vdot(got, x)) < 1e-9


def test_fold_is_lexicographic_order_in_f32():
    """Planted ordering change: summing blocks in reverse order generally
    changes float32 bits, so bit-exactness pins the order."""
    rng = np.random.default_rng(8)
    block, stride, im = (8, 8), (2, 2), (20, 20)
    nbl = O.get_nblocks(im, block, stride)
    xb = (rng.standard_normal((1, *nbl, *block)) * 1e3).astype(np.float32)
    fwd = O.fold(xb, im, block, stride)
    rev = O.fold(xb[:, ::-1, ::-1].copy(), im, block, stride)  # not the same op
    flipped = np.zeros_like(fwd)
    for b in reversed(list(product(*(range(k) for k in nbl)))):
        slc = tuple(slice(bi * s, bi * s + bl) for bi, s, bl in zip(b, stride, block))
        flipped[(0, *slc)] += xb[(0, *b)]
    assert not np.array_equal(fwd, flipped)  # order matters in float32
    assert rev.shape == fwd.shape


# ------------------------------------------------------------- NUFFT
@pytest.mark.parametrize("grid_size,oversamp", [((16, 16), 2.0), ((12, 10), 1.25), ((8, 8, 8), 2.0)])
def test_nufft_oracle_approximates_dft(grid_size, oversamp):
    rng = np.random.default_rng(9)
    ndim = len(grid_size)
    locs = rng.uniform(-0.5, 0.5, (60, ndim)) * np.array(grid_size)
    x = rng.standard_normal(grid_size) + 1j * rng.standard_normal(grid_size)
    want = O.dft(x, locs)
    got = O.nufft(x, locs, oversamp=oversamp, width=4.0)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < (2e-2 if oversamp < 2 else 5e-3), err
    y = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    # adjoint of the same operator
    lhs = np.vdot(y, got)
    rhs = np.vdot(O.nufft_adjoint(y, locs, grid_size, oversamp=oversamp), x)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_pad_rules():
    assert O.pad_amounts(4, 9) == (2, 3)  # even image, odd pad
    assert O.pad_amounts(5, 10) == (3, 2)  # odd image, odd pad
    assert O.pad_amounts(4, 8) == (2, 2)
    x = np.arange(20.0).reshape(4, 5)
    assert np.array_equal(O.center_crop(O.center_pad(x, (4, 5), (9, 10)), (9, 10), (4, 5)), x)


# ----------------------------------------------------------------- CG
def test_cg_known_answers():
    rng = np.random.default_rng(10)
    y = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    x = O.conjugate_gradients(lambda v: 2 * v, y, max_num_iters=3)
    assert np.allclose(x, y / 2)
    M = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    A = M.conj().T @ M + 0.5 * np.eye(6)
    x = O.conjugate_gradients(lambda v: A @ v, y[:6], max_num_iters=6)
    assert np.allclose(x, np.linalg.solve(A, y[:6]), rtol=1e-8, atol=1e-10)


# ------------------------------------------------- trajectory gradient
# The oracle's analytic window derivative and trajectory gradient have no
# reference counterpart; they are pinned by scipy's I1 and by float64
# central differences of the oracle's own (reference-following) ungrid.

@pytest.mark.parametrize("beta", [1.0, 2.5, BETA_NUFFT, 14.0])
def test_kb_deriv_matches_scipy_i1(beta):
    from scipy.special import i1

    u = np.linspace(0, 1, 2001)
    z = beta * np.sqrt(1 - u * u)
    ratio = np.where(z > 0, i1(z) / np.where(z > 0, z, 1.0), 0.5)  # I1(z)/z -> 1/2
    ref = -beta * beta * u * ratio
    got = O.kb_deriv(u, beta)
    # A&S 9.8.3 / 9.8.4: |rel err| of I1 below 2.2e-7
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300 + beta * beta * 1e-3)) < 3e-7


@pytest.mark.parametrize("beta", [2.5, BETA_NUFFT])
def test_kb_deriv_matches_fd_of_kb(beta):
    u = np.linspace(0.01, 0.99, 99)
    h = 1e-6
    fd = (O.kb(u + h, beta, "f64") - O.kb(u - h, beta, "f64")) / (2 * h)
    got = O.kb_deriv(u, beta)
    # I0 and I1 are separate A&S fits: their derivative relation holds to ~1e-7
    assert np.max(np.abs(got - fd) / np.abs(fd)) < 1e-6


def _away_from_tap_changes(t, width, margin):
    """True where moving t by +-margin changes no tap set: some tap sits at
    |g - t| = w/2 exactly when t - w/2 is an integer."""
    f = np.mod(t - np.asarray(width) / 2.0, 1.0)
    return (f > margin) & (f < 1 - margin)


@pytest.mark.parametrize("ndim,kernel,norm,pad,cplx", [
    (1, "kaiser_bessel", 1, "circular", True),
    (2, "kaiser_bessel", 1, "circular", True),
    (2, "kaiser_bessel", 2, "zero", False),
    (2, "spline", 1, "zero", True),
    (3, "kaiser_bessel", 1, "zero", True),
    (3, "spline", 2, "circular", False),
])
def test_oracle_grad_locs_matches_fd(ndim, kernel, norm, pad, cplx):
    rng = np.random.default_rng(40 + ndim + norm)
    grid = (12, 10, 9)[:ndim]
    width = (4.0, 3.0, 3.5)[:ndim]
    beta = 5.0
    B, M = 2, 40
    x = rng.standard_normal((B, *grid)) + (1j * rng.standard_normal((B, *grid)) if cplx else 0)
    h = rng.standard_normal((B, M)) + (1j * rng.standard_normal((B, M)) if cplx else 0)
    locs = rng.uniform(-1.0, np.array(grid) + 1.0, (M, ndim))
    kw = dict(kernel=kernel, norm=norm, pad_mode=pad, beta=beta)

    def loss(t):  # L(t) = sum Re(conj(h) * ungrid(x, t)), linear in x
        y = O.ungrid(x, t, width, precision="f64", **kw)
        return float(np.sum((np.conj(h) * y).real))

    g = O.grad_locs(x, h, locs, width, **kw)
    step = 1e-6
    checked = 0
    for p in range(M):
        for d in range(ndim):
            if not _away_from_tap_changes(locs[p, d], width[d], 10 * step):
                continue
            if norm == 2 or kernel == "spline":
                # kinks where a tap crosses |u| = 1 radially or u = 0 (hat)
                frac = np.mod(locs[p, d], 1.0)
                if min(frac, 1 - frac) < 10 * step:
                    continue
            tp, tm = locs.copy(), locs.copy()
            tp[p, d] += step
            tm[p, d] -= step
            fd = (loss(tp) - loss(tm)) / (2 * step)
            assert abs(g[p, d] - fd) <= 1e-5 * (1 + abs(fd)), (p, d, g[p, d], fd)
            checked += 1
    assert checked > M * ndim // 2
    # the grid form: X = grid(v, t), L = sum Re(conj(G) X)  ->  grad_locs(G, v, t)
    G = x
    v = h

    def loss_grid(t):
        X = O.grid(v, t, grid, width, precision="f64", **kw)
        return float(np.sum((np.conj(G) * X).real))

    p, d = 3, ndim - 1
    if _away_from_tap_changes(locs[p, d], width[d], 1e-4):
        tp, tm = locs.copy(), locs.copy()
        tp[p, d] += step
        tm[p, d] -= step
        fd = (loss_grid(tp) - loss_grid(tm)) / (2 * step)
        assert abs(O.grad_locs(G, v, locs, width, **kw)[p, d] - fd) <= 1e-5 * (1 + abs(fd))


def test_oracle_grad_locs_planted_sign_error_is_caught():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((1, 16, 16)) + 1j * rng.standard_normal((1, 16, 16))
    h = rng.standard_normal((1, 5)) + 1j * rng.standard_normal((1, 5))
    locs = rng.uniform(2, 14, (5, 2)) + 0.25
    g = O.grad_locs(x, h, locs, 4.0, beta=5.0)
    step = 1e-6
    tp, tm = locs.copy(), locs.copy()
    tp[0, 1] += step
    tm[0, 1] -= step
    f = lambda t: float(np.sum((np.conj(h) * O.ungrid(x, t, 4.0, beta=5.0, precision="f64")).real))
    fd = (f(tp) - f(tm)) / (2 * step)
    assert abs(g[0, 1] - fd) < 1e-5 * (1 + abs(fd))
    assert abs(-g[0, 1] - fd) > 1e-3 * (1 + abs(fd))  # a flipped sign would fail


# --------------------------------------- learned-preconditioner convs
def test_conv3d_oracle_matches_scipy_correlate():
    from scipy.ndimage import correlate

    rng = np.random.default_rng(21)
    x = rng.standard_normal((2, 3, 5, 6, 7))
    w = rng.standard_normal((4, 3, 3, 3, 3))
    b = rng.standard_normal(4)
    out = O.conv3d(x, w, b)
    for bi in range(2):
        for o in range(4):
            want = sum(correlate(x[bi, c], w[o, c], mode="constant", cval=0.0)
                       for c in range(3)) + b[o]
            assert np.allclose(out[bi, o], want, atol=1e-12)
    assert np.array_equal(O.conv3d(x, w, b, relu=True), np.maximum(out, 0))


def test_conv3d_oracle_backward_matches_fd():
    rng = np.random.default_rng(22)
    x = rng.standard_normal((1, 2, 4, 5, 3))
    w = rng.standard_normal((3, 2, 3, 3, 3))
    b = rng.standard_normal(3)
    g = rng.standard_normal((1, 3, 4, 5, 3))
    dx, dw, db = O.conv3d_backward(x, w, g)
    L = lambda xx, ww, bb: float(np.sum(g * O.conv3d(xx, ww, bb)))  # noqa: E731
    h = 1e-6
    for idx in [(0, 0, 0, 0, 0), (0, 1, 3, 4, 2), (0, 1, 2, 2, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        assert abs((L(xp, w, b) - L(xm, w, b)) / (2 * h) - dx[idx]) < 1e-6
    for idx in [(0, 0, 0, 0, 0), (2, 1, 1, 2, 0), (1, 0, 2, 1, 2)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        assert abs((L(x, wp, b) - L(x, wm, b)) / (2 * h) - dw[idx]) < 1e-6
    bp = b.copy()
    bp[1] += h
    assert abs((L(x, w, bp) - L(x, w, b)) / h - db[1]) < 1e-5


# ------------------------------------------------------------ sampling
def test_index_mask_known_answer():
    """The reference's from_mask case (tests/test_sampling.py:57-68): the
    True cells (2, 3) and (5, 6) of an 8x8 mask read those two values."""
    x = np.random.default_rng(0).standard_normal((8, 8))
    mask = np.zeros((8, 8), bool)
    mask[2, 3] = mask[5, 6] = True
    got = O.index(x, tuple(np.nonzero(mask)))
    assert got.shape == (2,)
    assert got[0] == x[2, 3] and got[1] == x[5, 6]


def test_index_brute_force_and_wrap():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 5, 4))
    i0 = rng.integers(-5, 10, (6, 7))  # negative and >= size wrap per axis
    i1 = rng.integers(-4, 8, (6, 7))
    got = O.index(x, (i0, i1))
    assert got.shape == (2, 3, 6, 7)
    for a, b in product(range(6), range(7)):
        assert np.array_equal(got[:, :, a, b], x[:, :, i0[a, b] % 5, i1[a, b] % 4])
    # a transposed index tuple reads other cells
    assert not np.array_equal(O.index(x[..., :4, :4], (i1 % 4, i0 % 4)),
                              O.index(x[..., :4, :4], (i0 % 4, i1 % 4)))


def test_index_adjoint_counts_adjointness_and_add_at():
    rng = np.random.default_rng(2)
    grid = (6, 5, 4)
    idx = tuple(rng.integers(0, n, (300,)) for n in grid)  # many repeats
    flat = O.ravel(idx, grid)
    # ones scatter to the hit count of each cell
    counts = O.index_adjoint(np.ones(300), idx, grid)
    assert np.array_equal(counts.reshape(-1), np.bincount(flat, minlength=prod(grid)))
    # adjoint of index in float64
    x = rng.standard_normal((3,) + grid) + 1j * rng.standard_normal((3,) + grid)
    y = rng.standard_normal((3, 300)) + 1j * rng.standard_normal((3, 300))
    lhs = np.vdot(y, O.index(x, idx))
    rhs = np.vdot(O.index_adjoint(y, idx, grid), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # float32: the sequential sum equals numpy's unbuffered add.at, bit for bit
    y32 = rng.standard_normal((2, 300)).astype(np.float32)
    want = np.zeros((2, prod(grid)), np.float32)
    np.add.at(want, (slice(None), flat), y32)
    got = O.index_adjoint(y32, idx, grid)
    assert got.dtype == np.float32
    assert np.array_equal(got.reshape(2, -1), want)


def test_prep_locs_sampling_doctest():
    """The reference's prep_locs doctest (linops/nufft.py:295-305): locs on
    the 1.25x lattice of a 64^3 grid map to integer cells 0..79."""
    import torch

    torch.manual_seed(0)
    locs = torch.rand(1000, 3) * 64 - 32
    locs = (torch.round(locs * 1.25) / 1.25).numpy()
    lp = O.prep_locs(locs, (64, 64, 64), (80, 80, 80), "sampling")
    assert lp.dtype == np.int64
    assert lp.min() == 0 and lp.max() == 79


@pytest.mark.parametrize("grid_size,oversamp", [((16, 16), 1.25), ((12, 10), 1.5), ((5, 7, 6), 2.0)])
def test_nufft_sampling_is_exact_dft_on_the_padded_lattice(grid_size, oversamp):
    """At k = m N / N_pad the sampled, zero-padded centred FFT is the exact
    DFT scaled by sqrt(prod N / prod N_pad) (odd and even sizes)."""
    rng = np.random.default_rng(3)
    ndim = len(grid_size)
    padded = tuple(int(n * oversamp) for n in grid_size)
    m = np.stack([rng.integers(-(p // 2), p - p // 2, 40) for p in padded], -1)
    locs = m * np.array(grid_size) / np.array(padded)
    x = rng.standard_normal(grid_size) + 1j * rng.standard_normal(grid_size)
    want = O.dft(x, locs) * np.sqrt(prod(grid_size) / prod(padded))
    got = O.nufft_sampling(x, locs, oversamp)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    y = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    lhs = np.vdot(y, got)
    rhs = np.vdot(O.nufft_sampling_adjoint(y, locs, grid_size, oversamp), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


# ------------------------------------------------------ Toeplitz wraps
@pytest.mark.parametrize("im,pad", [((4,), (9,)), ((5,), (10,)), ((5, 6), (10, 12)),
                                    ((3, 4, 5), (6, 9, 10)), ((7,), (7,))])
def test_wrapped_pad_crop_brute_force(im, pad):
    """pad_wrapped puts image index i at (i - N//2) mod P, zeros elsewhere;
    crop_wrapped reads it back (odd and even sizes)."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2,) + im) + 1j * rng.standard_normal((2,) + im)
    y = O.pad_wrapped(x, im, pad)
    want = np.zeros((2,) + pad, complex)
    for i in product(*(range(n) for n in im)):
        q = tuple((a - n // 2) % p for a, n, p in zip(i, im, pad))
        want[(slice(None),) + q] = x[(slice(None),) + i]
    assert np.array_equal(y, want)
    assert np.array_equal(O.crop_wrapped(y, im, pad), x)
    # crop_wrapped is the adjoint of pad_wrapped
    z = rng.standard_normal((2,) + pad) + 1j * rng.standard_normal((2,) + pad)
    assert abs(np.vdot(z, y) - np.vdot(O.crop_wrapped(z, im, pad), x)) < 1e-12 * abs(np.vdot(z, y))


@pytest.mark.parametrize("im,pad", [((6, 5), (12, 10)), ((5, 4, 3), (10, 8, 6)), ((9,), (18,))])
def test_shift_free_toeplitz_equals_centred_chain(im, pad):
    """crop_wrapped(ifftn(ifftshift(K) * fftn(pad_wrapped(x)))) equals the
    reference's Crop . cifftn . (K *) . cfftn . Pad."""
    rng = np.random.default_rng(6)
    nd = len(im)
    axes = tuple(range(-nd, 0))
    x = rng.standard_normal((3,) + im) + 1j * rng.standard_normal((3,) + im)
    K = rng.standard_normal(pad) + 1j * rng.standard_normal(pad)
    chain = O.center_crop(O.cfftn(K * O.cfftn(O.center_pad(x, im, pad), nd), nd, inverse=True),
                          pad, im)
    y = np.fft.fftn(O.pad_wrapped(x, im, pad), axes=axes, norm="ortho")
    y = np.fft.ifftn(np.fft.ifftshift(K, axes=axes) * y, axes=axes, norm="ortho")
    fused = O.crop_wrapped(y, im, pad)
    assert np.abs(fused - chain).max() < 1e-12 * np.abs(chain).max()
