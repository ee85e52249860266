"""Synthetic inputs of the benchmark workloads (numpy, seeded).

Trajectories are in centred grid units ``[-N/2, N/2)`` (the NUFFT
convention) unless stated otherwise.
"""

from __future__ import annotations

import numpy as np

__all__ = ["radial_3d", "spiral_2d", "vd_spiral_2d", "ellipse_phantom",
           "ellipsoid_phantom", "smooth_field", "gaussian_coil_maps", "uneven_2d_problems"]


def radial_3d(nspokes: int, nread: int, n: int, rng: np.random.Generator,
              centred: bool = True) -> np.ndarray:
    """3-D radial: directions uniform on the sphere, radius -n/2 + j n/nread.
    ``centred=False`` shifts by n/2 and wraps to [0, n) (grid units)."""
    g = rng.standard_normal((nspokes, 3))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    r = -n / 2 + np.arange(nread) * (n / nread)
    t = r[None, :, None] * u[:, None, :]
    if not centred:
        t = np.mod(t + n / 2, n)
    return t.astype(np.float32)


def spiral_2d(ninter: int, nsamp: int, n: int) -> np.ndarray:
    """Archimedean spiral, ninter interleaves x nsamp samples, radius
    growing linearly to n/2 with n / (2 ninter) turns per interleave."""
    s = np.arange(nsamp) / nsamp
    turns = n / (2.0 * ninter)
    r = (n / 2.0) * s
    out = np.empty((ninter, nsamp, 2))
    for l in range(ninter):
        th = 2 * np.pi * (turns * s + l / ninter)
        out[l, :, 0] = r * np.cos(th)
        out[l, :, 1] = r * np.sin(th)
    return out.astype(np.float32)


def vd_spiral_2d(npts: int, n: int, rng: np.random.Generator, power: float = 1.5) -> np.ndarray:
    """Variable-density spiral: random interleave count, radius
    (n/2) s^power (denser centre for power > 1), npts samples."""
    ninter = int(rng.integers(4, 33))
    per = max(1, -(-npts // ninter))
    s = (np.arange(per) + 0.5) / per
    r = (n / 2.0) * s**power
    turns = max(1.0, n / (4.0 * ninter))
    phase0 = rng.uniform(0, 2 * np.pi)
    out = []
    for l in range(ninter):
        th = 2 * np.pi * (turns * s + l / ninter) + phase0
        out.append(np.stack([r * np.cos(th), r * np.sin(th)], -1))
    k = np.concatenate(out, 0)[:npts]
    return np.clip(k, -n / 2, n / 2 - 1e-3).astype(np.float32)


def ellipse_phantom(n: int, rng: np.random.Generator, nell: int = 10) -> np.ndarray:
    yy, xx = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n), indexing="ij")
    img = np.zeros((n, n), dtype=np.complex128)
    for _ in range(nell):
        cx, cy = rng.uniform(-0.5, 0.5, 2)
        a, b = rng.uniform(0.1, 0.5, 2)
        th = rng.uniform(0, np.pi)
        u = (xx - cx) * np.cos(th) + (yy - cy) * np.sin(th)
        v = -(xx - cx) * np.sin(th) + (yy - cy) * np.cos(th)
        amp = rng.uniform(0.2, 1.0) * np.exp(1j * rng.uniform(0, 2 * np.pi))
        img[(u / a) ** 2 + (v / b) ** 2 <= 1] += amp
    return img.astype(np.complex64)


def ellipsoid_phantom(n: int, rng: np.random.Generator, nell: int = 8) -> np.ndarray:
    """Sum of random axis-aligned ellipsoids with random complex amplitudes."""
    g = np.linspace(-1, 1, n, dtype=np.float32)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    img = np.zeros((n, n, n), dtype=np.complex64)
    for _ in range(nell):
        c = rng.uniform(-0.4, 0.4, 3)
        ax = rng.uniform(0.1, 0.45, 3)
        amp = np.complex64(rng.uniform(0.2, 1.0) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
        m = (((xx - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2
             + ((zz - c[2]) / ax[2]) ** 2) <= 1
        img[m] += amp
    return img


def smooth_field(n: int, rng: np.random.Generator, ndim: int = 3, corr: float = 0.15) -> np.ndarray:
    """1 + 0.5 x (complex Gaussian noise low-passed to a correlation length
    of ``corr`` x field of view, unit RMS): smooth, nonzero everywhere."""
    shape = (n,) * ndim
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    k = np.meshgrid(*([np.fft.fftfreq(n)] * ndim), indexing="ij")
    k2 = sum(kk**2 for kk in k)
    f = np.fft.ifftn(np.fft.fftn(z) * np.exp(-2 * (np.pi * corr * n) ** 2 * k2 / n**2))
    f = f / np.sqrt(np.mean(np.abs(f) ** 2))
    return (1.0 + 0.5 * f).astype(np.complex64)


def gaussian_coil_maps(ncoil: int, n: int, rng: np.random.Generator, ndim: int = 3,
                       width: float = 0.6, normalize: bool = True) -> np.ndarray:
    """Smooth coil sensitivities: Gaussian bumps centred on a sphere (circle)
    around the field of view, with random phases; with ``normalize``,
    scaled so that sum_c |S_c|^2 = 1 everywhere."""
    g = np.linspace(-1, 1, n, dtype=np.float32)
    grids = np.meshgrid(*([g] * ndim), indexing="ij")
    maps = np.empty((ncoil,) + (n,) * ndim, dtype=np.complex64)
    for c in range(ncoil):
        d = rng.standard_normal(ndim)
        d = 1.2 * d / np.linalg.norm(d)
        r2 = sum((gr - dc) ** 2 for gr, dc in zip(grids, d))
        maps[c] = np.exp(-r2 / (2 * width**2)) * np.exp(1j * rng.uniform(0, 2 * np.pi))
    if normalize:
        maps /= np.sqrt(np.sum(np.abs(maps) ** 2, axis=0, keepdims=True))
    return maps


def uneven_2d_problems(nprob: int, rng: np.random.Generator, oversamp: float = 2.0,
                       grids=(128, 256, 384, 512), points=(1e4, 5e5), coils=(1, 32),
                       widths=(4.0, 5.0, 6.0)):
    """Independent 2-D NUFFT problems of uneven size: oversampled grid drawn
    from ``grids``, log-uniform point count, uniform coil count and kernel
    width, variable-density spiral with a random density power in [1.2, 2]."""
    out = []
    for _ in range(nprob):
        ng = int(rng.choice(grids))
        n = int(round(ng / oversamp))
        m = int(np.exp(rng.uniform(np.log(points[0]), np.log(points[1]))))
        c = int(rng.integers(coils[0], coils[1] + 1))
        w = float(rng.choice(widths))
        k = vd_spiral_2d(m, n, rng, power=float(rng.uniform(1.2, 2.0)))
        out.append(dict(grid=ng, im=n, npts=m, coils=c, width=w, locs=k))
    return out
