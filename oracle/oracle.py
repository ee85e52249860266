"""TEST INFRASTRUCTURE ONLY -- the oracle for the gridding path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The product
(torch-named-linops_amd/) never imports it.

Numpy front end over gridding_oracle.c (a C restatement of the reference's
interpolation / unfold / fold, see that file's header) plus numpy
restatements of the NUFFT chain and of CG:

  nufft / nufft_adjoint   src/torchlinops/linops/nufft.py:123-184 (chain),
                          241-334 (prep_locs), 336-346 (beta),
                          348-365 (apodize_weights); pad_last.py:139-199;
                          src/torchlinops/utils/_fft.py:7-60 (centred FFT)
  index / index_adjoint   src/torchlinops/functional/_index/index.py:17-99
                          (gather; sequential scatter-add of index_add_,
                          140-198 multi_grid / ravel)
  nufft_sampling(_adjoint) src/torchlinops/linops/nufft.py:185-203
                          (Pad -> centred FFT -> integer sampling),
                          prep_locs sampling branch 326-333
  pad_wrapped / crop_wrapped  centred pad + ifftshift, fftshift + crop
                          (pad_last.py:139-199, utils/_fft.py:7-31)
  conjugate_gradients     src/torchlinops/alg/pcg.py:13-88
  dft / dft_adjoint       brute-force nonuniform DFT (the tutorial's exact
                          operator, tutorials/mri_reconstruction.py:224-258)

Every function here is
pinned by a test in tests/test_oracle.py against a closed form, an
independent library routine (scipy / numpy) or brute force.
"""

from __future__ import annotations

import ctypes
from math import prod
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None

KERNELS = {"kaiser_bessel": 0, "spline": 1}
PADS = {"circular": 0, "zero": 1}

_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_vp = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            from . import build as _b  # type: ignore

            _b.build()
        h = ctypes.CDLL(str(path))
        h.orc_kb_f32.restype = ctypes.c_float
        h.orc_kb_f32.argtypes = [ctypes.c_float, ctypes.c_float]
        h.orc_kb_f64.restype = ctypes.c_double
        h.orc_kb_f64.argtypes = [ctypes.c_double, ctypes.c_double]
        h.orc_kb_vec_f32.argtypes = [_vp, _vp, _i64, ctypes.c_float]
        h.orc_kb_vec_f64.argtypes = [_vp, _vp, _i64, ctypes.c_double]
        common = [_vp, _vp, _vp, _i64, _i64, _i32, _vp, _vp, _i32, _i32, _i32, _i32]
        h.orc_ungrid_f32.argtypes = common + [ctypes.c_float, _i32]
        h.orc_ungrid_f64.argtypes = common + [ctypes.c_double, _i32]
        h.orc_grid_f32.argtypes = common + [ctypes.c_float, _i32]
        h.orc_grid_f64.argtypes = common + [ctypes.c_double]
        for f in (h.orc_ungrid_f32, h.orc_ungrid_f64, h.orc_grid_f32, h.orc_grid_f64):
            f.restype = _i32
        h.orc_kb_deriv_vec_f64.argtypes = [_vp, _vp, _i64, ctypes.c_double]
        h.orc_grad_locs_f64.argtypes = [_vp] + common + [ctypes.c_double, _i32]
        h.orc_grad_locs_f64.restype = _i32
        h.orc_unfold.argtypes = [_vp, _vp, _i32, _i64, _i32, _vp, _vp, _vp, _vp]
        h.orc_fold_f32.argtypes = [_vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp]
        h.orc_fold_f64.argtypes = [_vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp]
        _LIB = h
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i64arr(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int64))


def _f32arr(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32))


# ------------------------------------------------------------------ kernels
def kb(x, beta: float, precision: str = "f32") -> np.ndarray:
    """Kaiser-Bessel window I0~(beta sqrt(1 - x^2)) (kernels.py:28-89)."""
    if precision == "f32":
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
        out = np.empty_like(x)
        lib().orc_kb_vec_f32(_ptr(x), _ptr(out), x.size, float(beta))
    else:
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        out = np.empty_like(x)
        lib().orc_kb_vec_f64(_ptr(x), _ptr(out), x.size, float(beta))
    return out


def kb_deriv(u, beta: float) -> np.ndarray:
    """d/du of the Kaiser-Bessel window, float64 (A&S 9.8.3 / 9.8.4)."""
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    out = np.empty_like(u)
    lib().orc_kb_deriv_vec_f64(_ptr(u), _ptr(out), u.size, float(beta))
    return out


def _width(width, ndim):
    if np.isscalar(width):
        return (float(width),) * ndim
    width = tuple(float(w) for w in width)
    assert len(width) == ndim
    return width


def _as_interleaved(vals: np.ndarray, precision: str):
    is_complex = np.iscomplexobj(vals)
    if precision == "f32":
        dt, cdt = np.float32, np.complex64
    else:
        dt, cdt = np.float64, np.complex128
    if is_complex:
        v = np.ascontiguousarray(vals.astype(cdt)).view(dt)
    else:
        v = np.ascontiguousarray(vals.astype(dt))
    return v, is_complex, dt, cdt


# -------------------------------------------------------- interpolation
def ungrid(vals, locs, width, kernel="kaiser_bessel", norm=1,
           pad_mode="circular", beta=1.0, precision="f32", nthreads=1):
    """vals [B..., *grid], locs [L..., D] -> [B..., L...]
    (ungrid.py:37-67, 624-684)."""
    locs = np.asarray(locs)
    ndim = locs.shape[-1]
    grid_size = vals.shape[vals.ndim - ndim:]
    batch = vals.shape[: vals.ndim - ndim]
    lb = locs.shape[:-1]
    nb, npts = prod(batch), prod(lb)
    width = _width(width, ndim)
    v, is_complex, dt, cdt = _as_interleaved(vals.reshape(nb, *grid_size), precision)
    l = np.ascontiguousarray(locs.reshape(npts, ndim).astype(dt))
    out = np.empty((nb, npts, 2 if is_complex else 1), dtype=dt)
    fn = lib().orc_ungrid_f32 if precision == "f32" else lib().orc_ungrid_f64
    rc = fn(_ptr(v), _ptr(l), _ptr(out), nb, npts, ndim, _ptr(_i64arr(grid_size)),
            _ptr(_f32arr(width)), KERNELS[kernel], norm, PADS[pad_mode],
            int(is_complex), float(beta), int(nthreads))
    assert rc == 0
    out = out.view(cdt)[..., 0] if is_complex else out[..., 0]
    return out.reshape(*batch, *lb)


def grid(vals, locs, grid_size, width, kernel="kaiser_bessel", norm=1,
         pad_mode="circular", beta=1.0, precision="f32", nthreads=1):
    """vals [B..., L...], locs [L..., D] -> [B..., *grid_size]
    (grid.py:37-72, 571-639)."""
    locs = np.asarray(locs)
    grid_size = tuple(int(g) for g in grid_size)
    ndim = len(grid_size)
    lb = locs.shape[:-1]
    npts = prod(lb)
    batch = vals.shape[: vals.ndim - len(lb)]
    nb = prod(batch)
    width = _width(width, ndim)
    v, is_complex, dt, cdt = _as_interleaved(vals.reshape(nb, npts), precision)
    l = np.ascontiguousarray(locs.reshape(npts, ndim).astype(dt))
    out = np.empty((nb, prod(grid_size), 2 if is_complex else 1), dtype=dt)
    args = (_ptr(v), _ptr(l), _ptr(out), nb, npts, ndim, _ptr(_i64arr(grid_size)),
            _ptr(_f32arr(width)), KERNELS[kernel], norm, PADS[pad_mode],
            int(is_complex), float(beta))
    if precision == "f32":
        rc = lib().orc_grid_f32(*args, int(nthreads))
    else:
        rc = lib().orc_grid_f64(*args)
    assert rc == 0
    out = out.view(cdt)[..., 0] if is_complex else out[..., 0]
    return out.reshape(*batch, *grid_size)


# ------------------------------------------------------- unfold / fold
def grad_locs(grid_vals, pt_vals, locs, width, kernel="kaiser_bessel", norm=1,
              pad_mode="circular", beta=1.0, nthreads=1):
    """float64 trajectory gradient: sum_b Re(conj(pt_vals) * dW/dlocs x grid_vals).

    grid_vals [B..., *grid], pt_vals [B..., L...], locs [L..., D] -> [L..., D].
    For y = ungrid(x, locs) and a real loss with upstream gradient gy this is
    dL/dlocs = grad_locs(x, gy, locs); for X = grid(v, locs) with upstream gX
    it is grad_locs(gX, v, locs).
    """
    locs = np.asarray(locs)
    ndim = locs.shape[-1]
    grid_size = grid_vals.shape[grid_vals.ndim - ndim:]
    batch = grid_vals.shape[: grid_vals.ndim - ndim]
    lb = locs.shape[:-1]
    nb, npts = prod(batch), prod(lb)
    width = _width(width, ndim)
    is_complex = np.iscomplexobj(grid_vals) or np.iscomplexobj(pt_vals)
    if is_complex:
        grid_vals = grid_vals.astype(np.complex128)
        pt_vals = pt_vals.astype(np.complex128)
    x, _, dt, _ = _as_interleaved(grid_vals.reshape(nb, *grid_size), "f64")
    h, _, _, _ = _as_interleaved(pt_vals.reshape(nb, npts), "f64")
    l = np.ascontiguousarray(locs.reshape(npts, ndim).astype(np.float64))
    out = np.empty((npts, ndim), dtype=np.float64)
    rc = lib().orc_grad_locs_f64(
        _ptr(x), _ptr(h), _ptr(l), _ptr(out), nb, npts, ndim, _ptr(_i64arr(grid_size)),
        _ptr(_f32arr(width)), KERNELS[kernel], norm, PADS[pad_mode], int(is_complex),
        float(beta), int(nthreads))
    assert rc == 0
    return out.reshape(*lb, ndim)


def get_nblocks(im_size, block_size, stride=None):
    stride = stride if stride is not None else (1,) * len(block_size)
    return tuple((i - b) // s + 1 for i, b, s in zip(im_size, block_size, stride))


def _interleave_sizes(sizes, is_complex):
    sizes = list(int(s) for s in sizes)
    if is_complex:
        sizes[-1] *= 2
    return sizes


def unfold(x: np.ndarray, block_size, stride=None) -> np.ndarray:
    """[B..., *im] -> [B..., *nblocks, *block] (unfold.py:531-561)."""
    ndim = len(block_size)
    stride = tuple(stride) if stride is not None else (1,) * ndim
    im = x.shape[x.ndim - ndim:]
    batch = x.shape[: x.ndim - ndim]
    nbl = get_nblocks(im, block_size, stride)
    is_complex = np.iscomplexobj(x)
    xr = np.ascontiguousarray(x)
    y = np.empty((*batch, *nbl, *block_size), dtype=x.dtype)
    eb = xr.itemsize // (2 if is_complex else 1)
    lib().orc_unfold(_ptr(xr), _ptr(y), eb, prod(batch), ndim,
                     _ptr(_i64arr(nbl)),
                     _ptr(_i64arr(_interleave_sizes(block_size, is_complex))),
                     _ptr(_i64arr(_interleave_sizes(im, is_complex))),
                     _ptr(_i64arr(_interleave_sizes(stride, is_complex))))
    return y


def fold(x: np.ndarray, im_size, block_size, stride=None) -> np.ndarray:
    """[B..., *nblocks, *block] -> [B..., *im] summing blocks in
    lexicographic order (fold.py:489-517)."""
    ndim = len(block_size)
    stride = tuple(stride) if stride is not None else (1,) * ndim
    nbl = get_nblocks(im_size, block_size, stride)
    batch = x.shape[: x.ndim - 2 * ndim]
    is_complex = np.iscomplexobj(x)
    real = x.real.dtype if is_complex else x.dtype
    xr = np.ascontiguousarray(x)
    y = np.empty((*batch, *im_size), dtype=x.dtype)
    args = (_ptr(xr), _ptr(y), prod(batch), ndim, _ptr(_i64arr(nbl)),
            _ptr(_i64arr(_interleave_sizes(block_size, is_complex))),
            _ptr(_i64arr(_interleave_sizes(im_size, is_complex))),
            _ptr(_i64arr(_interleave_sizes(stride, is_complex))))
    if real == np.float32:
        lib().orc_fold_f32(*args)
    elif real == np.float64:
        lib().orc_fold_f64(*args)
    else:
        raise TypeError(real)
    return y


# ------------------------------------------------------------- NUFFT
def nufft_beta(width, oversamp):
    """nufft.py:336-346."""
    return np.pi * (((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8) ** 0.5


def apodize_weights(grid_size, padded_size, width, beta):
    """nufft.py:348-365: prod_d a/sinh(a), a = sqrt(beta^2 - (pi w (n - N//2)/Np)^2)."""
    out = np.ones(tuple(grid_size), dtype=np.float64)
    for d, (n, np_) in enumerate(zip(grid_size, padded_size)):
        k = np.arange(n) - n // 2
        a = np.sqrt(beta**2 - (np.pi * width * k / np_) ** 2 + 0j)
        w = (a / np.sinh(a)).real
        shape = [1] * len(grid_size)
        shape[d] = n
        out = out * w.reshape(shape)
    return out


def pad_amounts(n, npad):
    """Centred pad rule (pad_last.py:139-185): (left, right) per axis."""
    tp = npad - n
    if tp % 2:
        return (tp // 2 + 1, tp // 2) if n % 2 else (tp // 2, tp // 2 + 1)
    return (tp // 2, tp // 2)


def center_pad(x, im_size, pad_size):
    ndim = len(im_size)
    pads = [(0, 0)] * (x.ndim - ndim) + [pad_amounts(n, p) for n, p in zip(im_size, pad_size)]
    return np.pad(x, pads)


def center_crop(x, pad_size, im_size):
    ndim = len(im_size)
    slc = [slice(None)] * (x.ndim - ndim)
    for n, p in zip(im_size, pad_size):
        lft, _ = pad_amounts(n, p)
        slc.append(slice(lft, lft + n))
    return x[tuple(slc)]


def cfftn(x, ndim, inverse=False):
    """ifftshift -> (i)fftn(norm='ortho') -> fftshift over the last ndim axes."""
    axes = tuple(range(-ndim, 0))
    x = np.fft.ifftshift(x, axes=axes)
    x = (np.fft.ifftn if inverse else np.fft.fftn)(x, axes=axes, norm="ortho")
    return np.fft.fftshift(x, axes=axes)


def pad_wrapped(x, im_size, pad_size):
    """ifftshift of the centred zero-pad (pad_last.py:139-199 then
    utils/_fft.py:7-31's ifftshift) over the last len(im_size) axes."""
    ndim = len(im_size)
    return np.fft.ifftshift(center_pad(x, im_size, pad_size), axes=tuple(range(-ndim, 0)))


def crop_wrapped(y, im_size, pad_size):
    """Centred crop of fftshift(y): the adjoint of pad_wrapped."""
    ndim = len(im_size)
    return center_crop(np.fft.fftshift(y, axes=tuple(range(-ndim, 0))), pad_size, im_size)


def prep_locs(locs, grid_size, padded_size, nufft_mode="interpolate"):
    """Scale centred coords to the padded grid and wrap (nufft.py:315-334);
    the sampling branch rounds half to even before the wrap and returns
    int64 (326-333)."""
    out = np.array(locs, dtype=np.float64, copy=True)
    for i in range(-len(grid_size), 0):
        out[..., i] = out[..., i] * (padded_size[i] / grid_size[i]) + padded_size[i] // 2
        if nufft_mode == "sampling":
            out[..., i] = np.round(out[..., i])
        out[..., i] = np.mod(out[..., i], padded_size[i])
    if nufft_mode == "sampling":
        out = out.astype(np.int64)
    return out


def ravel(idx, grid_size):
    """Row-major cell of each index vector, components taken mod the axis
    size (index.py:183-198).  idx: tuple of broadcastable int arrays."""
    idx = np.broadcast_arrays(*[np.asarray(i, dtype=np.int64) for i in idx])
    flat = np.zeros(idx[0].shape, dtype=np.int64)
    for i, n in zip(idx, grid_size):
        flat = flat * int(n) + np.mod(i, int(n))
    return flat


def index(x, idx):
    """x[..., D...] gathered at the index tuple over the last D axes
    (multi_index, index.py:52-72): [N..., D...] -> [N..., I...]."""
    x = np.asarray(x)
    nd = len(idx)
    grid_size = x.shape[x.ndim - nd:]
    flat = ravel(idx, grid_size)
    xf = x.reshape(x.shape[: x.ndim - nd] + (-1,))
    return xf[..., flat.reshape(-1)].reshape(x.shape[: x.ndim - nd] + flat.shape)


def index_adjoint(y, idx, grid_size):
    """Sequential scatter-add: out[..., cell(p)] += y[..., p] for p in
    ascending order into zeros (multi_grid's index_add_, index.py:140-180).
    Accumulates in y's dtype, one point at a time."""
    y = np.asarray(y)
    flat = ravel(idx, grid_size)
    nl = flat.ndim
    batch = y.shape[: y.ndim - nl]
    yf = y.reshape(batch + (-1,))
    out = np.zeros(batch + (prod(grid_size),), dtype=y.dtype)
    for p, c in enumerate(flat.reshape(-1)):
        out[..., c] += yf[..., p]
    return out.reshape(batch + tuple(grid_size))


def nufft_sampling(x, locs, oversamp=1.25, prep=True):
    """Sampling-mode NUFFT in float64: Pad -> centred FFT -> read at the
    integer padded-grid cells of prep_locs (nufft.py:185-203)."""
    ndim = locs.shape[-1]
    grid_size = x.shape[x.ndim - ndim:]
    padded = tuple(int(n * oversamp) for n in grid_size)
    y = cfftn(center_pad(x, grid_size, padded), ndim)
    lp = prep_locs(locs, grid_size, padded, "sampling") if prep else np.asarray(locs)
    return index(y, tuple(lp[..., d] for d in range(ndim)))


def nufft_sampling_adjoint(y, locs, grid_size, oversamp=1.25, prep=True):
    ndim = locs.shape[-1]
    padded = tuple(int(n * oversamp) for n in grid_size)
    lp = prep_locs(locs, grid_size, padded, "sampling") if prep else np.asarray(locs)
    g = index_adjoint(y, tuple(lp[..., d] for d in range(ndim)), padded)
    return center_crop(cfftn(g, ndim, inverse=True), padded, grid_size)


def nufft(x, locs, oversamp=1.25, width=4.0, nthreads=1):
    """Forward NUFFT in float64: Scale * Interp(FFT_c(Pad(apod * x)))
    (``nthreads`` threads for the interpolation)."""
    ndim = locs.shape[-1]
    grid_size = x.shape[x.ndim - ndim:]
    padded = tuple(int(n * oversamp) for n in grid_size)
    beta = nufft_beta(width, oversamp)
    apod = apodize_weights(grid_size, padded, width, beta)
    y = center_pad(x * apod, grid_size, padded)
    y = cfftn(y, ndim)
    lp = prep_locs(locs, grid_size, padded)
    y = ungrid(y, lp, width, beta=beta, precision="f64", nthreads=nthreads)
    scale = width**ndim * (prod(grid_size) / prod(padded)) ** 0.5
    return y / scale


def nufft_adjoint(y, locs, grid_size, oversamp=1.25, width=4.0):
    ndim = locs.shape[-1]
    padded = tuple(int(n * oversamp) for n in grid_size)
    beta = nufft_beta(width, oversamp)
    scale = width**ndim * (prod(grid_size) / prod(padded)) ** 0.5
    lp = prep_locs(locs, grid_size, padded)
    x = grid(y / scale, lp, padded, width, beta=beta, precision="f64")
    x = cfftn(x, ndim, inverse=True)
    x = center_crop(x, padded, grid_size)
    return x * apodize_weights(grid_size, padded, width, beta)


def dft(x, locs):
    """Exact nonuniform DFT with the centred, orthonormal convention:
    y_m = N^{-1/2} sum_n x_n exp(-2 pi i sum_d k_md (n_d - N_d//2) / N_d)."""
    ndim = locs.shape[-1]
    grid_size = x.shape[x.ndim - ndim:]
    batch = x.shape[: x.ndim - ndim]
    k = locs.reshape(-1, ndim)
    phase = np.zeros((k.shape[0],) + tuple(grid_size))
    for d, n in enumerate(grid_size):
        idx = (np.arange(n) - n // 2) / n
        shape = [1] * ndim
        shape[d] = n
        phase = phase + k[:, d].reshape(-1, *([1] * ndim)) * idx.reshape(shape)
    E = np.exp(-2j * np.pi * phase).reshape(k.shape[0], -1) / np.sqrt(prod(grid_size))
    out = x.reshape(prod(batch), -1) @ E.T
    return out.reshape(*batch, *locs.shape[:-1])


# --------------------------------------------------------------- CG
def conjugate_gradients(A, y, max_num_iters=20, x0=None):
    """Plain CG from x0 = 0 (pcg.py:67-87), no tracking."""
    x = np.zeros_like(y) if x0 is None else x0.copy()
    r = y - A(x)
    p = r.copy()
    rs = np.vdot(r, r).real
    for _ in range(max_num_iters):
        if rs == 0:  # exact solution reached
            break
        Ap = A(p)
        alpha = rs / np.vdot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        rs_old = rs
        rs = np.vdot(r, r).real
        p = (rs / rs_old) * p + r
    return x


# ------------------------------------------- learned-preconditioner convs
# No reference counterpart (the learned preconditioner is an extension):
# float64 restatement of a 3x3x3 "same" cross-correlation layer and its
# backward pass, pinned by scipy.ndimage.correlate and finite differences.
def conv3d(x, w, b=None, relu=False):
    """x [B, Cin, Z, Y, X], w [Cout, Cin, 3, 3, 3] -> [B, Cout, Z, Y, X]:
    out[o](p) = sum_{c, d} w[o, c, d + 1] x[c](p + d), zero outside."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    B, C, Z, Y, X = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1), (1, 1)))
    out = np.zeros((B, w.shape[0], Z, Y, X))
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                out += np.einsum("oc,bczyx->bozyx", w[:, :, dz, dy, dx],
                                 xp[:, :, dz:dz + Z, dy:dy + Y, dx:dx + X])
    if b is not None:
        out += np.asarray(b, dtype=np.float64)[None, :, None, None, None]
    return np.maximum(out, 0.0) if relu else out


def conv3d_backward(x, w, g):
    """Gradients of sum(g * conv3d(x, w, b)) (no activation):
    dx[c](p) = sum_{o, d} w[o, c, d + 1] g[o](p - d),
    dw[o, c, d + 1] = sum_p g[o](p) x[c](p + d), db[o] = sum_p g[o](p)."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    B, C, Z, Y, X = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1), (1, 1)))
    gp = np.pad(g, ((0, 0), (0, 0), (1, 1), (1, 1), (1, 1)))
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for dz in range(3):
        for dy in range(3):
            for ddx in range(3):
                # g[o](p - d) with d = (dz - 1, ...) sits at gp index p - d + 1
                gs = gp[:, :, 2 - dz:2 - dz + Z, 2 - dy:2 - dy + Y, 2 - ddx:2 - ddx + X]
                dx += np.einsum("oc,bozyx->bczyx", w[:, :, dz, dy, ddx], gs)
                dw[:, :, dz, dy, ddx] = np.einsum(
                    "bozyx,bczyx->oc", g, xp[:, :, dz:dz + Z, dy:dy + Y, ddx:ddx + X])
    return dx, dw, g.sum(axis=(0, 2, 3, 4))
