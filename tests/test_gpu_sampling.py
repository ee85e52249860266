"""GPU parity of integer-index sampling (tnl_index / tnl_index_adjoint)
against the oracle: the gather bit-exact for every element size, the
scatter-add bit-exact against the oracle's sequential sum (repeated cells
included), autograd, the Sampling linop contract, the sampling-mode NUFFT,
and a round trip at a 256^3 x 8-coil size."""

from math import prod

import numpy as np
import pytest
import torch

import torchlinops_amd as tl
import torchlinops_amd.functional as F
from oracle import oracle as O
from torchlinops_amd.functional import _index as FI
from torchlinops_amd.testing import BaseNamedLinopTests
from tests._util import rel_l2

pytestmark = pytest.mark.gpu


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("GPU is required but not available")
    return torch.device("cuda:0")


def _rand(shape, dtype, g):
    if dtype.is_complex:
        return torch.randn(shape, dtype=dtype, generator=g)
    if dtype.is_floating_point:
        return torch.randn(shape, generator=g).to(dtype)
    return torch.randint(-1000, 1000, shape, generator=g).to(dtype)


GATHER_CASES = [
    # (batch, grid, idx shape, dtype)
    ((), (37,), (50,), torch.float32),
    ((3,), (16, 12), (7, 9), torch.complex64),
    ((2, 3), (9, 8, 7), (33,), torch.complex128),
    ((5,), (20, 21), (4, 5, 6), torch.float64),
    ((4,), (64, 64), (13, 17), torch.int16),
    ((1,), (10, 11, 12), (1000,), torch.uint8),
    ((6,), (8,), (0,), torch.float32),  # no points
]


@pytest.mark.parametrize("batch,grid,ishape,dtype", GATHER_CASES)
def test_index_matches_oracle(dev, batch, grid, ishape, dtype):
    print(f"index batch={batch} grid={grid} idx={ishape} {dtype}", flush=True)
    g = torch.Generator().manual_seed(sum(grid) + len(batch))
    x = _rand(batch + grid, dtype, g)
    # out-of-range and negative components wrap per axis, as multi_index
    idx = tuple(torch.randint(-n, 2 * n, ishape, generator=g) for n in grid)
    got = F.index(x.to(dev), tuple(i.to(dev) for i in idx)).cpu()
    want = O.index(x.numpy(), tuple(i.numpy() for i in idx))
    assert got.shape == batch + ishape
    assert np.array_equal(got.numpy(), want)


ADJ_CASES = [
    ((), (37,), (50,), torch.float32),
    ((3,), (16, 12), (400,), torch.complex64),  # ~2 points per cell
    ((2,), (5, 4, 3), (7, 300), torch.complex64),  # heavy repeats
    ((5,), (20, 21), (4, 5, 6), torch.float64),
    ((4, 2), (9, 8, 7), (33,), torch.complex128),
]


@pytest.mark.parametrize("batch,grid,ishape,dtype", ADJ_CASES)
def test_index_adjoint_matches_sequential_sum(dev, batch, grid, ishape, dtype):
    print(f"index_adjoint batch={batch} grid={grid} idx={ishape} {dtype}", flush=True)
    g = torch.Generator().manual_seed(7 + sum(grid))
    y = _rand(batch + ishape, dtype, g)
    idx = tuple(torch.randint(0, n, ishape, generator=g) for n in grid)
    counts = np.bincount(O.ravel(tuple(i.numpy() for i in idx), grid).reshape(-1))
    assert counts.max() <= FI.HEAVY_MIN  # every cell on the sequential path
    got = F.index_adjoint(y.to(dev), tuple(i.to(dev) for i in idx), grid).cpu()
    want = O.index_adjoint(y.numpy(), tuple(i.numpy() for i in idx), grid)
    assert got.shape == batch + grid
    assert np.array_equal(got.numpy(), want)


@pytest.mark.parametrize("dtype", [torch.complex64, torch.float64])
def test_index_adjoint_heavy_cells(dev, dtype):
    """Rounded radial-like points: a few cells hold thousands of points
    (summed by a workgroup each, fixed order), the rest few (sequential,
    bit-exact)."""
    print(f"index_adjoint heavy cells {dtype}", flush=True)
    g = torch.Generator().manual_seed(21)
    n, spokes, nread = 24, 2000, 24
    u = torch.randn(spokes, 3, generator=g)
    u = u / u.norm(dim=1, keepdim=True)
    r = torch.arange(nread) - nread / 2
    t = torch.remainder(torch.round(r[None, :, None] * u[:, None, :]) + n // 2, n).long()
    idx = tuple(t[..., d] for d in range(3))
    y = _rand((5, spokes, nread), dtype, g)
    counts = np.bincount(O.ravel(tuple(i.numpy() for i in idx), (n,) * 3).reshape(-1),
                         minlength=n**3)
    assert counts.max() > 1000 and (counts > FI.HEAVY_MIN).sum() >= 1
    imap = FI.index_map(tuple(i.to(dev) for i in idx), (n,) * 3, dev)
    got = FI.index_adjoint_with_map(y.to(dev), imap)
    for _ in range(2):  # deterministic
        assert torch.equal(FI.index_adjoint_with_map(y.to(dev), imap), got)
    got = got.cpu().numpy().reshape(5, -1)
    want32 = O.index_adjoint(y.numpy(), idx, (n,) * 3).reshape(5, -1)
    light = counts <= FI.HEAVY_MIN
    assert np.array_equal(got[:, light], want32[:, light])
    want64 = O.index_adjoint(y.numpy().astype(np.complex128), idx, (n,) * 3).reshape(5, -1)
    heavy = ~light
    err = np.abs(got[:, heavy] - want64[:, heavy]).max() / np.abs(want64[:, heavy]).max()
    assert err < (2e-6 if dtype == torch.complex64 else 1e-14), err


@pytest.mark.parametrize("dtype", [torch.int8, torch.uint8, torch.int16, torch.int32,
                                   torch.int64])
def test_index_adjoint_integer_types(dev, dtype):
    """Integer scatter-add is exact and wraps like the sequential sum of the
    same type, on light and heavy (workgroup-summed) cells alike."""
    print(f"index_adjoint integer {dtype}", flush=True)
    g = torch.Generator().manual_seed(17)
    grid, npts = (6, 5), 6000  # ~200 points per cell: every cell heavy
    info = torch.iinfo(dtype)
    # values near the type's range so that the sums wrap many times over
    lo, hi = max(info.min, -(1 << 62)), min(info.max, 1 << 62)
    y = torch.randint(lo, hi, (3, npts), generator=g, dtype=torch.int64).to(dtype)
    idx = tuple(torch.randint(0, n, (npts,), generator=g) for n in grid)
    light = (torch.randint(0, 40, (3, 50), generator=g, dtype=torch.int64).to(dtype),
             tuple(torch.randint(0, n, (50,), generator=g) for n in (9, 11)))
    for vals, ix, gs in [(y, idx, grid), (light[0], light[1], (9, 11))]:
        got = F.index_adjoint(vals.to(dev), tuple(i.to(dev) for i in ix), gs)
        assert got.dtype == dtype
        want = O.index_adjoint(vals.numpy(), tuple(i.numpy() for i in ix), gs)
        assert np.array_equal(got.cpu().numpy(), want)
    counts = np.bincount(O.ravel(tuple(i.numpy() for i in idx), grid).reshape(-1))
    assert counts.min() > FI.HEAVY_MIN


def test_index_adjoint_is_deterministic(dev):
    g = torch.Generator().manual_seed(3)
    y = torch.randn(4, 20000, dtype=torch.complex64, generator=g).to(dev)
    idx = (torch.randint(0, 8, (20000,), generator=g).to(dev),
           torch.randint(0, 8, (20000,), generator=g).to(dev))
    a = F.index_adjoint(y, idx, (8, 8))
    for _ in range(3):
        assert torch.equal(F.index_adjoint(y, idx, (8, 8)), a)


def test_index_with_slices_matches_torch_indexing(dev):
    g = torch.Generator().manual_seed(4)
    x = torch.randn(3, 10, 9, 8, generator=g)
    t = torch.randint(0, 9, (5,), generator=g)
    u = torch.randint(0, 8, (5,), generator=g)
    for idx in [(slice(None), t, u), (slice(2, 7), t, slice(None, None, 2)),
                (torch.randint(0, 10, (5,), generator=g), slice(1, 4), u)]:
        want = x[(slice(None),) + idx]
        got = F.index(x.to(dev), tuple(i.to(dev) if torch.is_tensor(i) else i for i in idx))
        assert got.shape == want.shape
        assert torch.equal(got.cpu(), want)


def test_index_autograd(dev):
    g = torch.Generator().manual_seed(5)
    x = torch.randn(2, 12, 10, dtype=torch.complex64, generator=g).to(dev).requires_grad_()
    idx = (torch.randint(0, 12, (40,), generator=g).to(dev),
           torch.randint(0, 10, (40,), generator=g).to(dev))
    w = torch.randn(2, 40, dtype=torch.complex64, generator=g).to(dev)
    y = F.index(x, idx)
    (y * w.conj()).real.sum().backward()
    # d/dx Re<w, Sx> = S^H w
    want = F.index_adjoint(w, idx, (12, 10))
    assert torch.allclose(x.grad, want, rtol=1e-6, atol=1e-6)
    v = torch.randn(2, 40, dtype=torch.complex64, generator=g).to(dev).requires_grad_()
    z = F.index_adjoint(v, idx, (12, 10))
    u = torch.randn(2, 12, 10, dtype=torch.complex64, generator=g).to(dev)
    (z * u.conj()).real.sum().backward()
    assert torch.equal(v.grad, F.index(u, idx))


class TestSampling(BaseNamedLinopTests):
    """The reference's TestSampling (tests/test_sampling.py:9-38) on the GPU."""

    equality_check = "approx"
    isclose_kwargs = {"rtol": 1e-4}

    @pytest.fixture(scope="class")
    def linop_input_output(self):
        dev = _gpu()
        g = torch.Generator().manual_seed(0)
        N, R, K, B = 64, 13, 17, 3
        idx = torch.randint(0, N - 1, (R, K, 2), generator=g).to(dev)
        x = torch.randn(B, N, N, generator=g).to(dev)
        y = torch.randn(B, R, K, generator=g).to(dev)
        return tl.Sampling.from_stacked_idx(idx, (N, N), ("R", "K")), x, y

    def test_split(self, linop_input_output):
        A, x, _ = linop_input_output
        As = A.split(A, {"R": slice(2, 5), "K": slice(4, 10)})
        assert (A(x)[:, 2:5, 4:10] == As(x)).all()

    def test_values(self, linop_input_output):
        A, x, y = linop_input_output
        idx = tuple(i.cpu().numpy() for i in A.idx)
        assert np.array_equal(A(x).cpu().numpy(), O.index(x.cpu().numpy(), idx))
        assert np.array_equal(A.H(y).cpu().numpy(),
                              O.index_adjoint(y.cpu().numpy(), idx, (64, 64)))


def test_sampling_from_mask_values(dev):
    mask = torch.zeros(8, 8, dtype=torch.bool)
    mask[2, 3] = mask[5, 6] = True
    A = tl.Sampling.from_mask(mask.to(dev), input_size=(8, 8))
    x = torch.randn(8, 8).to(dev)
    r = A(x)
    assert r.shape == (2,) and r[0] == x[2, 3] and r[1] == x[5, 6]


@pytest.mark.parametrize("grid_size,oversamp", [((16, 16), 1.25), ((12, 10), 1.5), ((10, 12, 9), 2.0)])
def test_nufft_sampling_matches_oracle(dev, grid_size, oversamp):
    print(f"nufft sampling grid={grid_size} os={oversamp}", flush=True)
    rng = np.random.default_rng(11)
    ndim = len(grid_size)
    padded = tuple(int(n * oversamp) for n in grid_size)
    m = np.stack([rng.integers(-(p // 2), p - p // 2, (6, 30)) for p in padded], -1)
    locs = (m * np.array(grid_size) / np.array(padded)).astype(np.float32)
    x = (rng.standard_normal((2,) + grid_size) + 1j * rng.standard_normal((2,) + grid_size))
    y = rng.standard_normal((2, 6, 30)) + 1j * rng.standard_normal((2, 6, 30))
    A = tl.NUFFT(torch.from_numpy(locs).to(dev), grid_size, output_shape=("R", "K"),
                 oversamp=oversamp, mode="sampling")
    lp = A.interp.locs.cpu().numpy()
    assert np.array_equal(lp, O.prep_locs(locs, grid_size, padded, "sampling"))
    got = A(torch.from_numpy(x).to(torch.complex64).to(dev)).cpu().numpy()
    want = O.nufft_sampling(x, locs.astype(np.float64), oversamp)
    assert rel_l2(got, want) < 1e-5
    got_h = A.H(torch.from_numpy(y).to(torch.complex64).to(dev)).cpu().numpy()
    want_h = O.nufft_sampling_adjoint(y, locs.astype(np.float64), grid_size, oversamp)
    assert rel_l2(got_h, want_h) < 1e-5
    # the exact DFT at the lattice frequencies
    dft = O.dft(x, locs.astype(np.float64)) * np.sqrt(prod(grid_size) / prod(padded))
    assert rel_l2(got, dft) < 1e-5


def test_sampling_round_trip_full_size(dev):
    """8 coils on a 256^3 grid, 4M distinct cells: S^H S x keeps exactly
    the sampled cells of x and zeroes the rest; S S^H y = y."""
    print("sampling round trip 8 x 256^3, 4M cells", flush=True)
    n, C, M = 256, 8, 1 << 22
    g = torch.Generator(device=dev).manual_seed(12)
    cells = torch.randperm(n**3, generator=g, device=dev)[:M]
    idx = (cells // (n * n), (cells // n) % n, cells % n)
    A = tl.Sampling(idx, (n, n, n), output_shape=("K",))
    x = torch.randn(C, n, n, n, dtype=torch.complex64, device=dev, generator=g)
    y = A(x)
    back = A.H(y)
    mask = torch.zeros(n**3, dtype=torch.bool, device=dev)
    mask[cells] = True
    mask = mask.view(n, n, n)
    assert torch.equal(back, torch.where(mask, x, torch.zeros((), dtype=x.dtype, device=dev)))
    assert torch.equal(A(back), y)
    # checksum of the gather against a torch advanced-indexing read
    assert torch.equal(y, x.reshape(C, -1)[:, cells])


def test_sampling_edge_cases(dev):
    """Empty batch, empty point set (the scatter-add then writes zeros),
    a single cell, and a batch that is not a multiple of the 4 entries a
    lane handles."""
    g = torch.Generator().manual_seed(30)
    x0 = torch.randn(0, 6, 5, dtype=torch.complex64, generator=g).to(dev)
    idx = (torch.tensor([1, 2], device=dev), torch.tensor([3, 4], device=dev))
    assert F.index(x0, idx).shape == (0, 2)
    assert F.index_adjoint(torch.zeros(0, 2, dtype=torch.complex64, device=dev), idx,
                           (6, 5)).shape == (0, 6, 5)
    empty = (torch.zeros(0, dtype=torch.long, device=dev),) * 2
    z = F.index_adjoint(torch.zeros(3, 0, dtype=torch.complex64, device=dev), empty, (6, 5))
    assert z.shape == (3, 6, 5) and not z.abs().any()
    one = torch.randn(7, 1, 1, generator=g).to(dev)
    ii = (torch.zeros(9, dtype=torch.long, device=dev),) * 2
    got = F.index(one, ii)
    assert torch.equal(got, one.reshape(7, 1).expand(7, 9))
    back = F.index_adjoint(got, ii, (1, 1))
    want = O.index_adjoint(got.cpu().numpy(), tuple(i.cpu().numpy() for i in ii), (1, 1))
    assert np.array_equal(back.cpu().numpy(), want)


def test_wrapped_pad_crop_empty_batch(dev):
    from torchlinops_amd.linops.toeplitz import crop_wrapped, pad_wrapped

    x = torch.zeros(0, 4, 4, dtype=torch.complex64, device=dev)
    assert pad_wrapped(x, (4, 4), (8, 8)).shape == (0, 8, 8)
    y = torch.zeros(0, 8, 8, dtype=torch.complex64, device=dev)
    assert crop_wrapped(y, (4, 4), (8, 8)).shape == (0, 4, 4)


@pytest.mark.parametrize("grid_size", [(16, 16), (10, 12, 9)])
def test_fused_sampling_nufft_matches_chain(dev, grid_size):
    """Sampling-mode NUFFT: the fused path (scaled wrapped pad, uncentred
    FFT, gather at relabelled cells) against the Pad / centred FFT /
    Sampling chain, forward, adjoint and normal."""
    rng = np.random.default_rng(31)
    padded = tuple(int(n * 1.25) for n in grid_size)
    m = np.stack([rng.integers(-(p // 2), p - p // 2, (5, 40)) for p in padded], -1)
    locs = torch.from_numpy((m * np.array(grid_size) / np.array(padded)).astype(np.float32))
    A = tl.NUFFT(locs.to(dev), grid_size, output_shape=("R", "K"), batch_shape=("C",),
                 oversamp=1.25, mode="sampling")
    B = tl.NUFFT(locs.to(dev), grid_size, output_shape=("R", "K"), batch_shape=("C",),
                 oversamp=1.25, mode="sampling", fused=False)
    assert A._fused_on and not B._fused_on
    x = torch.randn(3, *grid_size, dtype=torch.complex64, device=dev)
    y = torch.randn(3, 5, 40, dtype=torch.complex64, device=dev)
    assert rel_l2(A(x).cpu().numpy(), B(x).cpu().numpy()) < 1e-5
    assert rel_l2(A.H(y).cpu().numpy(), B.H(y).cpu().numpy()) < 1e-5
    assert rel_l2(A.N(x).cpu().numpy(), B.N(x).cpu().numpy()) < 1e-5
