"""The package's host (CPU) path against the oracle.

Host tensors take ``functional/_host.py`` (the reference's own dispatch on
``is_cuda`` sends them to its torch fallback).  Checked here, on the CPU:
  * the C1 workload through the public API: 2-D interpolate / adjoint on a
    128x128 complex64 grid at 16 384 uniform points (KB width 4, beta
    pi sqrt(8.2)) within 1e-5 of the reference-order float32 oracle and of
    the float64 oracle; array_to_blocks on 1x128x128 complex64 with 8x8
    blocks, stride 4, and blocks_to_array back, bit-exact;
  * 1-D / 3-D, both kernels, both norms, both pads, integer coordinates and
    coordinates outside the grid; float64 / complex128 against the float64
    oracle at 1e-12 (spline) / 5e-7 (Kaiser-Bessel: the oracle's A&S window
    is good to 2.5e-7);
  * fold order: bit-identical to the oracle's sequential lexicographic sum
    on overlapping blocks (a changed order would change the bits);
  * index / index_adjoint bit-exact (sequential ``index_add_``);
  * the trajectory gradient against the float64 oracle gradient;
  * the NUFFT linop and the differentiable Toeplitz form on host tensors.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests._util import crandn, rel_l2

BETA = float(O.nufft_beta(4.0, 2.0))


def test_c1_interpolate_and_blocks_through_the_api():
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(1)
    x = crandn(rng, (128, 128))
    locs = rng.uniform(0, 128, (16384, 2)).astype(np.float32)
    y = crandn(rng, (16384,))
    xt, lt, yt = (torch.from_numpy(a) for a in (x, locs, y))
    kp = {"beta": BETA}
    u = F.interpolate(xt, lt, 4.0, kernel_params=kp)
    g = F.interpolate_adjoint(yt, lt, (128, 128), 4.0, kernel_params=kp)
    assert u.dtype == torch.complex64 and u.shape == (16384,)
    assert g.shape == (128, 128)
    assert rel_l2(u.numpy(), O.ungrid(x, locs, 4.0, beta=BETA)) < 1e-5
    assert rel_l2(g.numpy(), O.grid(y, locs, (128, 128), 4.0, beta=BETA)) < 1e-5
    assert rel_l2(u.numpy(), O.ungrid(x, locs, 4.0, beta=BETA, precision="f64")) < 1e-5
    assert rel_l2(g.numpy(), O.grid(y, locs, (128, 128), 4.0, beta=BETA,
                                     precision="f64")) < 1e-5
    a = crandn(rng, (1, 128, 128))
    b = F.array_to_blocks(torch.from_numpy(a), (8, 8), (4, 4))
    assert b.shape == (1, 31, 31, 8, 8)
    assert np.array_equal(b.numpy(), O.unfold(a, (8, 8), (4, 4)))
    back = F.blocks_to_array(b, (128, 128), (8, 8), (4, 4))
    assert np.array_equal(back.numpy(), O.fold(b.numpy(), (128, 128), (8, 8), (4, 4)))


CASES = [
    # (grid, width, kernel, norm, pad)
    ((40,), 4.0, "kaiser_bessel", 1, "circular"),
    ((40,), 3.0, "spline", 1, "zero"),
    ((20, 24), 4.0, "kaiser_bessel", 1, "zero"),
    ((20, 24), (3.0, 5.0), "spline", 1, "circular"),
    ((20, 24), 4.0, "kaiser_bessel", 2, "circular"),
    ((10, 12, 14), 4.0, "kaiser_bessel", 1, "circular"),
    ((10, 12, 14), 6.0, "kaiser_bessel", 1, "zero"),
    ((10, 12, 14), 3.0, "spline", 2, "zero"),
]


@pytest.mark.parametrize("grid_size,width,kernel,norm,pad", CASES)
def test_host_interp_matches_oracle(grid_size, width, kernel, norm, pad):
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(len(grid_size) * 7 + norm)
    nd = len(grid_size)
    x = crandn(rng, (3, *grid_size))
    locs = rng.uniform(-3, 3 + max(grid_size), (500, nd)).astype(np.float32)
    locs[:50] = np.round(locs[:50])  # integer coordinates: ntap + 1 taps
    y = crandn(rng, (3, 500))
    beta = BETA if kernel == "kaiser_bessel" else 1.0
    kw = dict(kernel=kernel, norm=norm, pad_mode=pad)
    kp = {"beta": beta}
    u = F.interpolate(torch.from_numpy(x), torch.from_numpy(locs), width,
                      kernel_params=kp, **kw).numpy()
    g = F.interpolate_adjoint(torch.from_numpy(y), torch.from_numpy(locs), grid_size,
                              width, kernel_params=kp, **kw).numpy()
    assert rel_l2(u, O.ungrid(x, locs, width, beta=beta, **kw)) < 1e-5
    assert rel_l2(g, O.grid(y, locs, grid_size, width, beta=beta, **kw)) < 1e-5
    # float64 / complex128 computes in double precision
    u64 = F.interpolate(torch.from_numpy(x.astype(np.complex128)),
                        torch.from_numpy(locs.astype(np.float64)), width,
                        kernel_params=kp, **kw).numpy()
    g64 = F.interpolate_adjoint(torch.from_numpy(y.astype(np.complex128)),
                                torch.from_numpy(locs.astype(np.float64)), grid_size,
                                width, kernel_params=kp, **kw).numpy()
    assert u64.dtype == np.complex128
    # the oracle's Kaiser-Bessel window is the A&S polynomial (relative error
    # below 2.5e-7 of I0); the host path evaluates I0 itself
    tol = 5e-7 if kernel == "kaiser_bessel" else 1e-12
    assert rel_l2(u64, O.ungrid(x, locs, width, beta=beta, precision="f64", **kw)) < tol
    assert rel_l2(g64, O.grid(y, locs, grid_size, width, beta=beta, precision="f64",
                              **kw)) < tol


def test_host_planted_mistake_is_caught():
    """A transposed coordinate tuple must fail the bar the host path meets."""
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(5)
    x = crandn(rng, (20, 24))
    locs = rng.uniform(0, 20, (300, 2)).astype(np.float32)
    u = F.interpolate(torch.from_numpy(x), torch.from_numpy(locs[:, ::-1].copy()), 4.0,
                      kernel_params={"beta": BETA}).numpy()
    assert rel_l2(u, O.ungrid(x, locs, 4.0, beta=BETA)) > 1e-2


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64])
@pytest.mark.parametrize("shape,block,stride", [
    ((2, 30), (7,), (3,)),
    ((2, 17, 19), (5, 4), (2, 3)),
    ((1, 11, 12, 13), (4, 4, 4), (2, 2, 3)),
    ((2, 9, 9), (3, 3), (1, 1)),
])
def test_host_unfold_fold_bit_exact(dtype, shape, block, stride):
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape)
    if np.dtype(dtype).kind == "c":
        x = x + 1j * rng.standard_normal(shape)
    x = x.astype(dtype)
    nd = len(block)
    b = F.array_to_blocks(torch.from_numpy(x), block, stride)
    assert np.array_equal(b.numpy(), O.unfold(x, block, stride))
    im = shape[-nd:]
    f = F.blocks_to_array(b, im, block, stride).numpy()
    ref = O.fold(b.numpy(), im, block, stride)
    assert f.dtype == ref.dtype and np.array_equal(f.view(np.uint8), ref.view(np.uint8))


def test_host_fold_order_is_pinned():
    """Summing the covering blocks in reverse order changes float32 bits, so
    bit-equality above pins the lexicographic order."""
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(3)
    blocks = (rng.standard_normal((1, 14, 14, 4, 4)) * 10 ** rng.uniform(-3, 3, (1, 14, 14, 4, 4))
              ).astype(np.float32)
    f = F.blocks_to_array(torch.from_numpy(blocks), (17, 17), (4, 4), (1, 1)).numpy()
    # the oracle's own order agrees bit for bit
    assert np.array_equal(f, O.fold(blocks, (17, 17), (4, 4), (1, 1)))
    # a reversed walk over the blocks gives different bits
    out = np.zeros((17, 17), np.float32)
    for i in reversed(range(14)):
        for j in reversed(range(14)):
            out[i:i + 4, j:j + 4] += blocks[0, i, j]
    assert not np.array_equal(out, f)


def test_host_index_and_adjoint_bit_exact():
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(4)
    x = crandn(rng, (3, 12, 10))
    idx = (rng.integers(0, 12, 400), rng.integers(0, 10, 400))
    tidx = tuple(torch.from_numpy(i) for i in idx)
    out = F.index(torch.from_numpy(x), tidx).numpy()
    assert np.array_equal(out, O.index(x, idx))
    y = crandn(rng, (3, 400))
    adj = F.index_adjoint(torch.from_numpy(y), tidx, (12, 10)).numpy()
    assert np.array_equal(adj, O.index_adjoint(y, idx, (12, 10)))


def test_host_trajectory_gradient():
    from torchlinops_amd import functional as F

    rng = np.random.default_rng(6)
    x = crandn(rng, (2, 16, 18)).astype(np.complex128)
    locs = rng.uniform(0, 16, (200, 2))
    gy = crandn(rng, (2, 200)).astype(np.complex128)
    got = F.interp_grad_locs(torch.from_numpy(x), torch.from_numpy(gy),
                             torch.from_numpy(locs), 4.0,
                             kernel_params={"beta": BETA}).numpy()
    ref = O.grad_locs(x, gy, locs, 4.0, beta=BETA)
    # the oracle's window and its derivative are the A&S I0 / I1 polynomials
    assert rel_l2(got, ref) < 5e-7
    # through autograd on the public op
    lt = torch.from_numpy(locs).requires_grad_(True)
    u = F.interpolate(torch.from_numpy(x), lt, 4.0, kernel_params={"beta": BETA})
    (u.conj() * torch.from_numpy(gy)).real.sum().backward()
    assert rel_l2(lt.grad.numpy(), ref) < 5e-7


def test_host_nufft_and_toeplitz():
    from torchlinops_amd.linops import NUFFT

    rng = np.random.default_rng(7)
    n = 24
    locs = rng.uniform(-n / 2, n / 2, (600, 2)).astype(np.float32)
    x = crandn(rng, (n, n))
    A = NUFFT(torch.from_numpy(locs), (n, n), output_shape=("K",), oversamp=1.25, width=4.0)
    y = A(torch.from_numpy(x)).numpy()
    assert rel_l2(y, O.nufft(x, locs, 1.25, 4.0)) < 1e-5
    z = crandn(rng, (600,))
    xa = A.H(torch.from_numpy(z)).numpy()
    assert rel_l2(xa, O.nufft_adjoint(z, locs, (n, n), 1.25, 4.0)) < 1e-5
    # Toeplitz normal (2x PSF grid; the approximation is tight at os 2)
    A2 = NUFFT(torch.from_numpy(locs), (n, n), output_shape=("K",), oversamp=2.0)
    T = NUFFT(torch.from_numpy(locs), (n, n), output_shape=("K",), oversamp=2.0,
              toeplitz=True)
    Tc = NUFFT(torch.from_numpy(locs), (n, n), output_shape=("K",), oversamp=2.0,
               toeplitz=True, toeplitz_fused=False)
    xt = torch.from_numpy(x).requires_grad_(True)
    out = T.N(xt)
    explicit = A2.H(A2(torch.from_numpy(x))).numpy()
    assert rel_l2(out.detach().numpy(), explicit) < 2e-3
    # the host form of the fused operator equals the reference-shaped chain
    chain = Tc.N(torch.from_numpy(x)).numpy()
    assert rel_l2(out.detach().numpy(), chain) < 1e-5
    # and is differentiable: d/dx ||T x||^2 = 2 T^H T x
    out.abs().pow(2).sum().backward()
    want = 2 * Tc.N(Tc.N(torch.from_numpy(x))).numpy()
    assert rel_l2(xt.grad.numpy(), want) < 1e-5
