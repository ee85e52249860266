"""Shared helpers for the tests: seeded inputs and error measures."""

from __future__ import annotations

import numpy as np
import torch


def rel_l2(a, b) -> float:
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def crandn(rng: np.random.Generator, shape, dtype=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def radial_3d(nspokes: int, nread: int, n: int, rng: np.random.Generator) -> np.ndarray:
    """3-D radial trajectory in grid units: spoke directions uniform on the
    sphere, radius -n/2 + j * n / nread, centred at n/2 and wrapped to [0, n)."""
    g = rng.standard_normal((nspokes, 3))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    r = -n / 2 + np.arange(nread) * (n / nread)
    t = r[None, :, None] * u[:, None, :] + n / 2
    return np.mod(t, n).astype(np.float32)


def to_dev(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def spiral_2d(ninter: int, nsamp: int, n: int) -> np.ndarray:
    """Archimedean spiral in centred grid units [-n/2, n/2): ninter
    interleaves of nsamp samples, radius growing linearly to n/2 with
    n / (2 * ninter) turns per interleave (Nyquist spacing between turns)."""
    s = np.arange(nsamp) / nsamp
    turns = n / (2.0 * ninter)
    r = (n / 2.0) * s
    out = np.empty((ninter, nsamp, 2), dtype=np.float64)
    for l in range(ninter):
        th = 2 * np.pi * (turns * s + l / ninter)
        out[l, :, 0] = r * np.cos(th)
        out[l, :, 1] = r * np.sin(th)
    return out.astype(np.float32)


def ellipse_phantom(n: int, rng: np.random.Generator, nell: int = 10) -> np.ndarray:
    """Sum of random ellipses with random complex amplitudes on an n x n grid."""
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


def vd_spiral_2d(npts: int, n: int, rng: np.random.Generator, power: float = 1.5) -> np.ndarray:
    """Variable-density spiral in centred grid units: a random number of
    interleaves, radius (n/2) * s^power (denser centre for power > 1)."""
    ninter = int(rng.integers(4, 33))
    per = max(1, npts // ninter)
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
