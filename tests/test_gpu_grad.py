"""Trajectory gradient (d/dlocs of interpolate / interpolate_adjoint) on
the GPU against the float64 oracle (oracle.grad_locs, itself pinned by
finite differences in test_oracle.py) and, end to end through the NUFFT,
against float64 central finite differences (h = 1e-4) of the loss
||A(k) x - y||^2 evaluated by the oracle.

Tolerances: the kernel accumulates in float32; against the float64 oracle
its relative L2 error is a few 1e-6 on these inputs, bounded here at 1e-4.
"""

import numpy as np
import pytest
import torch

import torchlinops_amd.functional as F
from oracle import oracle as O
from _util import crandn, ellipse_phantom, rel_l2, spiral_2d, to_dev

pytestmark = pytest.mark.gpu


CASES = [
    (1, "kaiser_bessel", 1, "circular", True, (64,), 4.0),
    (2, "kaiser_bessel", 1, "circular", True, (40, 36), 4.0),
    (2, "kaiser_bessel", 1, "zero", False, (40, 36), 5.0),
    (2, "kaiser_bessel", 2, "circular", True, (32, 32), 4.0),
    (2, "spline", 1, "circular", True, (32, 30), 3.0),
    (3, "kaiser_bessel", 1, "circular", True, (16, 18, 20), 4.0),
    (3, "kaiser_bessel", 1, "zero", True, (16, 16, 16), 3.0),
    (3, "spline", 2, "zero", False, (12, 12, 12), 2.0),
]


def _inputs(rng, ndim, grid, cplx, B=3, M=500):
    x = crandn(rng, (B, *grid)) if cplx else rng.standard_normal((B, *grid)).astype(np.float32)
    h = crandn(rng, (B, M)) if cplx else rng.standard_normal((B, M)).astype(np.float32)
    locs = rng.uniform(-2.0, np.array(grid) + 2.0, (M, ndim)).astype(np.float32)
    return x, h, locs


@pytest.mark.parametrize("ndim,kernel,norm,pad,cplx,grid,width", CASES)
def test_grad_kernel_matches_oracle(dev, ndim, kernel, norm, pad, cplx, grid, width):
    rng = np.random.default_rng(100 + ndim * 7 + norm)
    x, h, locs = _inputs(rng, ndim, grid, cplx)
    kp = {"beta": 6.0}
    got = F.interp_grad_locs(to_dev(x, dev), to_dev(h, dev), to_dev(locs, dev), width,
                             kernel=kernel, norm=norm, pad_mode=pad, kernel_params=kp)
    want = O.grad_locs(x, h, locs.astype(np.float64), width, kernel=kernel, norm=norm,
                       pad_mode=pad, beta=6.0)
    assert got.dtype == torch.float32 and got.shape == locs.shape
    assert rel_l2(got.cpu().numpy(), want) < 1e-4


def test_autograd_interpolate_locs(dev):
    rng = np.random.default_rng(3)
    x, c, locs = _inputs(rng, 2, (48, 40), True, B=2, M=700)
    xt = to_dev(x, dev)
    lt = to_dev(locs, dev).requires_grad_(True)
    y = F.interpolate(xt, lt, 4.0, kernel_params={"beta": 7.0})
    loss = (torch.conj(to_dev(c, dev)) * y).real.sum()
    loss.backward()
    want = O.grad_locs(x, c, locs.astype(np.float64), 4.0, beta=7.0)
    assert rel_l2(lt.grad.cpu().numpy(), want) < 1e-4


def test_autograd_interpolate_adjoint_locs_and_vals(dev):
    rng = np.random.default_rng(4)
    G, v, locs = _inputs(rng, 3, (16, 16, 16), True, B=2, M=400)
    vt = to_dev(v, dev).requires_grad_(True)
    lt = to_dev(locs, dev).requires_grad_(True)
    X = F.interpolate_adjoint(vt, lt, (16, 16, 16), 4.0, kernel_params={"beta": 7.0})
    loss = (torch.conj(to_dev(G, dev)) * X).real.sum()
    loss.backward()
    want = O.grad_locs(G, v, locs.astype(np.float64), 4.0, beta=7.0)
    assert rel_l2(lt.grad.cpu().numpy(), want) < 1e-4
    # the vals gradient still follows the reference (ungrid of the upstream)
    want_v = O.ungrid(G, locs.astype(np.float64), 4.0, beta=7.0, precision="f64")
    assert rel_l2(vt.grad.cpu().numpy(), want_v) < 1e-5


def test_grad_wider_than_capacity_raises(dev):
    x = torch.zeros(1, 64, dtype=torch.complex64, device=dev)
    h = torch.zeros(1, 3, dtype=torch.complex64, device=dev)
    l = torch.zeros(3, 1, device=dev)
    with pytest.raises(ValueError, match="widths below 16"):
        F.interp_grad_locs(x, h, l, 20.0)


def test_trajectory_gradient_spiral_nufft_matches_fp64_fd(dev):
    """256^2 image (ellipses + N(0, 0.01) noise), grid 512^2, 64 spiral
    interleaves x 2048 samples; y from the trajectory perturbed by
    N(0, 0.1^2) grid units; d/dk ||A(k) x - y||^2 from autograd (the
    trajectory-gradient kernel) against float64 central differences of the
    oracle's NUFFT along random directions.  Directions are zeroed on
    coordinates within 3 steps of a tap-set change (where the loss is not
    differentiable)."""
    rng = np.random.default_rng(2026)
    N, os_, width, h = 256, 2.0, 4.0, 1e-4
    k = spiral_2d(64, 2048, N).astype(np.float64)
    x = ellipse_phantom(N, rng)
    x = (x + 0.01 * (rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape))).astype(np.complex64)
    k_true = k + 0.1 * rng.standard_normal(k.shape)
    y = O.nufft(x.astype(np.complex128), k_true, os_, width)

    kt = to_dev(k.astype(np.float32), dev).requires_grad_(True)
    A = F.nufft(to_dev(x, dev), kt, oversamp=os_, width=width)
    loss = (A - to_dev(y.astype(np.complex64), dev)).abs().pow(2).sum()
    loss.backward()
    g = kt.grad.double().cpu().numpy()

    k32 = k.astype(np.float32).astype(np.float64)  # the point the GPU saw
    lp = O.prep_locs(k32, (N, N), (int(N * os_),) * 2)
    keep = (np.mod(lp - width / 2, 1.0) > 3 * h * os_) & (np.mod(lp - width / 2, 1.0) < 1 - 3 * h * os_)

    def loss64(kk):
        r = O.nufft(x.astype(np.complex128), kk, os_, width) - y
        return float(np.sum(np.abs(r) ** 2))

    errs = []
    for _ in range(3):
        d = rng.standard_normal(k.shape) * keep
        fd = (loss64(k32 + h * d) - loss64(k32 - h * d)) / (2 * h)
        errs.append(abs(np.sum(g * d) - fd) / abs(fd))
    print("directional-derivative relative errors vs FP64 FD:", errs)
    assert max(errs) < 1e-3
