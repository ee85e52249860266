"""Shared helpers for the tests: seeded inputs and error measures."""

from __future__ import annotations

import numpy as np
import torch


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
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
