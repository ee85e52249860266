"""Host-side (CPU) checks of the named-dimension core and the linops that
are plain torch ops (Pad, FFT, Diagonal, Scalar, Chain algebra, NUFFT
construction and constants).  The interpolation linops themselves run only
on the GPU (tests/test_gpu_linops.py)."""

import numpy as np
import pytest
import torch

import torchlinops_amd as tl
from torchlinops_amd import config
from torchlinops_amd.nameddim import ND, NS, Dim, isequal, resolve_wildcards
from oracle import oracle as O


def test_dim_parsing_examples():
    assert Dim("ABCD") == ("A", "B", "C", "D")
    assert Dim("NxNyNz") == ("Nx", "Ny", "Nz")
    assert Dim("A1B2Kx1Ky2") == ("A1", "B2", "Kx1", "Ky2")
    assert Dim("()A()B") == ("()", "A", "()", "B")
    assert Dim("(1)A(23)") == ("(1)", "A", "(23)")
    assert Dim("") == ()


def test_named_dimension():
    assert repr(ND("A")) == "A" and repr(ND("A", 1)) == "A1"
    assert ND("A") == "A" and ND.infer("B2") == ND("B", 2)
    assert ND("A").next_unused(("A", "A1")) == "A2"
    assert ND("...").next_unused(("...",)) == "..."


@pytest.mark.parametrize("a,b,want", [
    (("A", "B"), ("A", "B"), True), (("A", "C"), ("A",), False), (("A", "C"), (), False),
    (("A", "C"), ("...",), True), (("A", "C", "..."), ("...",), True),
    (("A", "B", "C"), ("A", "...", "C"), True), (("...", "A", "C", "..."), ("...",), True),
    (("...", "A", "C"), ("B", "C"), False), (("A", "B"), ("A", "()"), True),
    (("A",), ("()", "()"), False),
])
def test_isequal_table(a, b, want):
    assert isequal(a, b)[0] is want


def test_shared_dim_rename_and_shapes():
    s = NS(("A", "B"), ("B", "C"))
    s.ishape = ("D", "E")
    assert s.ishape == (ND("D"), ND("E")) and s.oshape == (ND("E"), ND("C"))
    s2 = NS(("...", "Nx"), ("...", "Kx"), extra=("Nx",))
    s2.ishape = ("C", "Nx")
    assert s2.ishape == ("C", "Nx") and s2.oshape == ("...", "Kx")
    assert NS(("A", "B"), ("C",)).N.oshape == ("A1", "B1")
    assert NS(("A", "B"), ("B",)).N.oshape == ("A1", "B")
    assert NS(("A",), ("B",)).H.ishape == ("B",)
    assert resolve_wildcards(("C", "..."), ("C", "Kx", "Ky")) == ("C", "Kx", "Ky")
    assert resolve_wildcards(("X", "..."), ("A", "B")) == ("X", "B")


def test_pad_rules_and_numerics():
    assert tl.linops.pad_to_size((4, 5), (9, 10)) == [3, 2, 2, 3]
    p = tl.Pad((9, 10), (4, 5))
    x = torch.randn(2, 4, 5)
    y = p(x)
    assert y.shape == (2, 9, 10)
    want = O.center_pad(x.numpy(), (4, 5), (9, 10))
    assert np.array_equal(y.numpy(), want)
    assert torch.equal(p.H(y), x)
    assert torch.allclose(p.N(x), x)
    with pytest.raises(ValueError):
        p(torch.randn(3, 3))


def test_fft_unitary_and_centered():
    f = tl.FFT(2, centered=True)
    x = torch.randn(3, 8, 6, dtype=torch.complex64)
    want = O.cfftn(x.numpy(), 2)
    assert np.allclose(f(x).numpy(), want, atol=1e-5)
    assert torch.allclose(f.H(f(x)), x, atol=1e-5)
    assert isinstance(f.N, tl.Identity)


def test_diagonal_scalar_algebra():
    w = torch.randn(4, 5, dtype=torch.complex64)
    d = tl.Diagonal(w, ("Nx", "Ny"))
    x = torch.randn(4, 5, dtype=torch.complex64)
    assert torch.allclose(d.H(x), x * w.conj())
    assert torch.allclose(d.N(x), x * w.abs() ** 2)
    s = 3.0 * d
    assert torch.allclose(s(x), 3 * w * x)
    a = d + d
    assert torch.allclose(a(x), 2 * w * x)
    assert torch.allclose((d - d)(x), torch.zeros_like(x))
    with config.using(cache_adjoint_normal=True):
        assert d.H is d.H
    with pytest.raises(ValueError):
        tl.Diagonal(torch.randn(2, 3, 4), ("A", "B"))


def test_chain_mismatch_raises():
    a = tl.Diagonal(torch.randn(3), ("A",))
    b = tl.Diagonal(torch.randn(3), ("B",))
    with pytest.raises(ValueError, match="Mismatched shape"):
        tl.Chain(a, b)


def test_nufft_structure_and_constants():
    locs = torch.rand(4, 5, 3) * 32 - 16
    A = tl.NUFFT(locs, (32, 32, 32), output_shape=("R", "K"), oversamp=1.25)
    assert [l.name for l in A.linops] == ["Apodize", "Pad", "FFT", "Interpolate", "Scalar"]
    assert [l.repr_name for l in A.N.linops] == [
        "Apodize", "Pad", "FFT", "Interpolate", "Scalar", "Interpolate.H", "FFT.H", "Crop.H", "Apodize"]
    beta = float(tl.NUFFT.beta(4.0, 1.25))
    assert abs(beta - O.nufft_beta(4.0, 1.25)) < 1e-12
    apod = tl.NUFFT.apodize_weights((32, 32, 32), (40, 40, 40), 1.25, 4.0, beta)
    want = O.apodize_weights((32, 32, 32), (40, 40, 40), 4.0, beta)
    assert np.allclose(apod.numpy(), want, rtol=1e-5)
    lp = tl.NUFFT.prep_locs(locs, (32, 32, 32), (40, 40, 40))
    assert np.allclose(lp.numpy(), O.prep_locs(locs.numpy(), (32, 32, 32), (40, 40, 40)), atol=1e-4)
    assert float(lp.min()) >= 0 and float(lp.max()) < 40
    assert A.size("R") == 4 and A.size("K") == 5
    As = tl.NUFFT(locs, (32, 32, 32), output_shape=("R", "K"), mode="sampling")
    assert [l.name for l in As.linops] == ["Pad", "FFT", "Sampling"]
    assert As.interp.locs.dtype == torch.int64 and As.size("K") == 5
    with pytest.raises(ValueError, match="Nan/Inf"):
        tl.NUFFT(locs, (32, 32, 32), output_shape=("R", "K"), oversamp=1.0, width=4.0)


def test_cg_reference_semantics_cpu():
    # A = 2I: one step lands on y/2 with r == 0 exactly.  The loss change of
    # that step is above ltol, so a second step would run and divide 0/0,
    # in the reference's loop (pcg.py:67-87) as here: run one step.
    g = torch.Generator().manual_seed(5)
    y = torch.randn(7, dtype=torch.complex64, generator=g)
    x = tl.conjugate_gradients(lambda v: 2 * v, y, max_num_iters=1)
    assert torch.allclose(x, y / 2, atol=1e-6)
    M = torch.randn(6, 6, dtype=torch.complex128, generator=g)
    A = M.conj().T @ M + 0.5 * torch.eye(6, dtype=torch.complex128)
    b = torch.randn(6, dtype=torch.complex128, generator=g)
    x = tl.conjugate_gradients(lambda v: A @ v, b, max_num_iters=12, disable_tracking=True)
    assert torch.allclose(x, torch.linalg.solve(A, b), atol=1e-8)
    s = tl.DeviceCG(lambda v: A @ v, lam=0.0, check_every=2)
    x2 = s.solve(b, tol=1e-10, max_iters=40)
    assert torch.allclose(x2, torch.linalg.solve(A, b), atol=1e-8)
    assert s.history[-1] <= 1e-10
    # preconditioned with the exact inverse: one step
    Ainv = torch.linalg.inv(A)
    s3 = tl.DeviceCG(lambda v: A @ v, M=lambda r: Ainv @ r, check_every=1)
    s3.solve(b, tol=1e-9, max_iters=5)
    assert s3.iters <= 2


def test_shifted_layout_replaces_centring_shifts():
    """The batched / SENSE normal operators skip fftshift / ifftshift: the
    image written at the wrapped cells of shifted_layout and transformed
    without shifts equals the centred transform relabelled by N // 2 (even
    and odd sizes, 1-D to 3-D), and the adjoint crop reads the same cells."""
    from torchlinops_amd.linops.nufft import cfftn, cifftn, shifted_layout

    g = torch.Generator().manual_seed(3)
    for im, pd in (((5, 6), (10, 12)), ((7,), (15,)), ((4, 5, 3), (8, 9, 7))):
        x = torch.randn(2, *im, dtype=torch.complex64, generator=g)
        nd = len(im)
        dim = tuple(range(-nd, 0))
        ix, shift = shifted_layout(im, pd, "cpu")
        full = torch.zeros(2, *pd, dtype=torch.complex64)
        full[(Ellipsis,) + tuple(tl.linops.crop_slice_from_pad(tl.linops.pad_to_size(im, pd)))] = x
        want = cfftn(full, dim=dim)
        wrapped = torch.zeros(2, *pd, dtype=torch.complex64)
        wrapped[(Ellipsis,) + ix] = x
        got = torch.fft.fftn(wrapped, dim=dim, norm="ortho")
        rolled = torch.roll(got, shifts=[int(s) for s in shift], dims=dim)
        assert torch.allclose(rolled, want, atol=1e-5)
        y = torch.randn(2, *pd, dtype=torch.complex64, generator=g)
        crop = tuple(tl.linops.crop_slice_from_pad(tl.linops.pad_to_size(im, pd)))
        want_adj = cifftn(y, dim=dim)[(Ellipsis,) + crop]
        y_unshifted = torch.roll(y, shifts=[-int(s) for s in shift], dims=dim)
        got_adj = torch.fft.ifftn(y_unshifted, dim=dim, norm="ortho")[(Ellipsis,) + ix]
        assert torch.allclose(got_adj, want_adj, atol=1e-5)


# ------------------------------------------------------------ sampling
def test_sampling_linop_shapes_split_and_size():
    """The reference's TestSampling shape checks (tests/test_sampling.py)."""
    g = torch.Generator().manual_seed(0)
    N, R, K = 64, 13, 17
    idx = torch.randint(0, N - 1, (R, K, 2), generator=g)
    A = tl.Sampling.from_stacked_idx(idx, (N, N), ("R", "K"))
    assert A.size("R") == 13 and A.size("K") == 17 and A.size("N") is None
    As = A.split(A, {"R": slice(2, 5), "K": slice(4, 10)})
    assert torch.equal(As.locs, idx[2:5, 4:10])
    assert A.locs.shape == (R, K, 2)
    A.register_shape("custom", ("A", "B"))
    assert len(A._shape["custom"]) == 2
    # output axes "..." cannot be split: the same linop comes back
    B = tl.Sampling(tuple(idx[..., d] for d in range(2)), (N, N))
    assert B.split(B, {"R": slice(0, 2)}) is B


def test_sampling_from_mask_and_range_error():
    mask = torch.zeros(8, 8, dtype=torch.bool)
    mask[2, 3] = mask[5, 6] = True
    A = tl.Sampling.from_mask(mask, input_size=(8, 8))
    assert torch.equal(A.locs, torch.tensor([[2, 3], [5, 6]]))
    with pytest.raises(ValueError, match="range"):
        tl.Sampling((torch.tensor([0, 8]),), input_size=(8,))
    with pytest.raises(ValueError, match="doesn't match"):
        tl.Sampling((torch.tensor([0, 1]),), input_size=(8, 8))


def test_sampling_host_path():
    """Host tensors take index_select / sequential index_add_, like the
    reference's CPU branch."""
    A = tl.Sampling((torch.tensor([0, 3, 3]),), input_size=(8,))
    x = torch.randn(8)
    assert torch.equal(A(x), x[[0, 3, 3]])
    y = torch.tensor([1.0, 2.0, 4.0])
    want = torch.zeros(8)
    want[0], want[3] = 1.0, 6.0
    assert torch.equal(A.H(y), want)
    assert torch.equal(tl.functional.index(x, (torch.tensor([1]),)), x[[1]])


def test_index_adjoint_bounds_error():
    with pytest.raises(IndexError, match="Out-of-bounds"):
        tl.functional.index_adjoint(torch.randn(2), (torch.tensor([0, 8]),), (8,))


def test_nufft_sampling_chain_and_toeplitz_error():
    from torchlinops_amd.linops.nufft import toeplitz_psf

    locs = torch.randint(0, 19, (10, 2))
    A = tl.NUFFT(locs.clone(), (16, 16), output_shape=Dim("K"), mode="sampling",
                 do_prep_locs=False)
    assert [type(l).__name__ for l in A.linops] == ["Pad", "FFT", "Sampling"]
    assert A.interp.input_size == (20, 20)
    with pytest.raises(NotImplementedError, match="Sampling"):
        toeplitz_psf(A)
    with pytest.raises(ValueError, match="integer-type"):
        tl.NUFFT(locs.float(), (16, 16), output_shape=Dim("K"), mode="sampling",
                 do_prep_locs=False)
    with pytest.raises(ValueError, match="Unrecognized NUFFT mode"):
        tl.NUFFT(locs.float(), (16, 16), output_shape=Dim("K"), mode="gridding")


def test_prep_locs_sampling_matches_oracle():
    torch.manual_seed(0)
    locs = torch.rand(1000, 3) * 64 - 32
    locs = torch.round(locs * 1.25) / 1.25
    got = tl.NUFFT.prep_locs(locs, (64, 64, 64), (80, 80, 80), nufft_mode="sampling")
    want = O.prep_locs(locs.numpy(), (64, 64, 64), (80, 80, 80), "sampling")
    assert got.dtype == torch.int64
    assert np.array_equal(got.numpy(), want)


def test_index_helpers():
    F = tl.functional
    assert tuple(F.slice2range(slice(None, None, -1), 3)) == (2, 1, 0)
    idx = F.ensure_tensor_indexing((torch.tensor([1, 2]),), (4, 5))
    assert idx[0].shape == (4, 1) and torch.equal(idx[1], torch.tensor([1, 2]))
    with pytest.raises(ValueError, match="boolean"):
        F.mask2idx(torch.zeros(3))
    t = torch.tensor([[1, -1], [7, 2]])
    assert torch.equal(F.ravel(t, (4, 3), dim=-1), torch.tensor([1 * 3 + 2, 3 * 3 + 2]))
    assert len(F.canonicalize_idx(torch.zeros(5, 3, dtype=torch.long))) == 3


# ------------------------------------------------ polynomial preconditioner
def test_poly_coefficients_closed_forms():
    from torchlinops_amd.alg import poly as P

    assert np.array_equal(P.ifista_coeffs(0), [1.0])
    assert np.array_equal(P.ifista_coeffs(2), [3.0, -3.0, 1.0])
    # degree 0: c0 = int x / int x^2 (l_2) and 2 / (a + b) (l_inf)
    a, b = 0.2, 1.7
    c2, _ = P.l_2_opt(0, a, b)
    assert abs(c2[0] - 1.5 * (b**2 - a**2) / (b**3 - a**3)) < 1e-6
    ci, _ = P.l_inf_opt(0, a, b)
    assert abs(ci[0] - 2 / (a + b)) < 1e-6


@pytest.mark.parametrize("degree", [1, 2, 4])
def test_poly_l2_is_the_least_squares_fit(degree):
    """l_2 coefficients equal a Gauss-Legendre least-squares fit of
    x p(x) ~ 1 (exact quadrature for these degrees)."""
    from torchlinops_amd.alg import poly as P

    a, b = 0.0, 1.0
    c, _ = P.l_2_opt(degree, a, b)
    xg, wg = np.polynomial.legendre.leggauss(degree + 4)
    x = (b - a) / 2 * xg + (a + b) / 2
    w = np.sqrt(wg * (b - a) / 2)
    V = x[:, None] ** (np.arange(degree + 1)[None, :] + 1)
    ref = np.linalg.lstsq(V * w[:, None], w, rcond=None)[0]
    assert np.allclose(c, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("degree", [1, 3, 5])
def test_poly_linf_equioscillates(degree):
    """|1 - x p(x)| <= 1/T_{d+1}((b+a)/(b-a)) on [a, b], with the bound
    reached at d + 2 points (Chebyshev alternation)."""
    from torchlinops_amd.alg import poly as P

    a, b = 0.05, 1.0
    _, p = P.l_inf_opt(degree, a, b)
    x = np.linspace(a, b, 200001)
    r = 1 - x * p(x)
    bound = 1 / np.cosh((degree + 1) * np.arccosh((b + a) / (b - a)))
    assert np.abs(r).max() <= bound * (1 + 1e-6)
    peaks = np.sum(np.abs(np.abs(r) - bound) < 1e-6 * bound + 1e-12)
    assert peaks >= degree + 2


def test_polynomial_preconditioner_linop_on_a_diagonal():
    """P(T) x on a diagonal T equals p(lambda) x, for the linop and for
    Horner's rule on tensors."""
    from torchlinops_amd.alg import poly as P

    lam = torch.linspace(0.01, 1.0, 50, dtype=torch.float64)
    T = tl.Diagonal(lam, ioshape=("N",))
    x = torch.randn(50, dtype=torch.float64)
    # (l_inf on [0, 1] at even d + 1 gives p(1) = 0, which the injectivity
    # check rejects as the reference's does)
    for norm, lo in (("l_2", 0.0), ("l_inf", 0.05), ("ifista", 0.0)):
        c = P._coeffs(3, norm, lo, 1.0)
        want = np.polynomial.polynomial.polyval(lam.numpy(), np.asarray(c, float)) * x.numpy()
        got = P.polynomial_preconditioner(T, 3, norm, lower_eig=lo)(x)
        assert np.allclose(got.numpy(), want, rtol=1e-6)
        got2 = P.apply_polynomial(T, c, x)
        assert np.allclose(got2.numpy(), want, rtol=1e-6)
    assert isinstance(P.polynomial_preconditioner(T, -1), tl.Identity)
    with pytest.raises(AssertionError, match="injective"):
        P.l_inf_opt(3, 0.0, 1.0)
    with pytest.raises(ValueError, match="Unknown"):
        P.polynomial_preconditioner(T, 2, "l_1")


@pytest.mark.parametrize("inner_type", ["simple_inner", "dense_inner", None])
def test_toeplitz_normal_with_inner(inner_type):
    """The reference's Toeplitz cases (tests/test_toeplitz.py:62-131) on host
    tensors: a Diagonal inner broadcast over the batch and a Dense inner
    that couples the two batch entries (A -> A1).  The Toeplitz normal must
    hold no NUFFT, equal A^H inner A up to the Toeplitz approximation, equal
    the reference-shaped chain (toeplitz_fused=False) and pass the adjoint
    test."""
    from tests._util import rel_l2

    torch.manual_seed(11)
    n, R, K = (16, 16), 4, 100
    locs = (torch.rand(R, K, 2) - 0.5) * 16
    kw = dict(output_shape=("R", "K"), batch_shape=("A",), oversamp=2.0)
    A = tl.NUFFT(locs.clone(), n, **kw)
    T = tl.NUFFT(locs.clone(), n, toeplitz=True, **kw)
    Tc = tl.NUFFT(locs.clone(), n, toeplitz=True, toeplitz_fused=False, **kw)
    inner = None
    if inner_type == "simple_inner":
        inner = tl.Diagonal(torch.randn(R, K), ioshape=("A", "R", "K"), broadcast_dims=("A",))
    elif inner_type == "dense_inner":
        inner = tl.Dense(torch.randn(2, 2, dtype=torch.complex64), weightshape=("A", "A1"),
                         ishape=("A", "R", "K"), oshape=("A1", "R", "K"))
    N, Nc = T.normal(inner), Tc.normal(inner)
    assert not any(isinstance(m, tl.NUFFT) for m in N.modules())
    if inner_type == "dense_inner":
        assert tuple(map(str, N.oshape))[0] == "A1"
    x = torch.randn(2, *n, dtype=torch.complex64)
    want = A.H(A(x) if inner is None else inner(A(x)))
    got = N(x)
    assert rel_l2(got.numpy(), want.numpy()) < 2e-3
    assert rel_l2(got.numpy(), Nc(x).numpy()) < 1e-5
    y = torch.randn(2, *n, dtype=torch.complex64)
    lhs = torch.vdot(y.flatten(), got.flatten())
    rhs = torch.vdot(N.H(y).flatten(), x.flatten())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_toeplitz_psf_shapes_and_errors():
    """psf_sizing / all_indices (nufft.py:501-541): a coupling kernel is a
    Dense over (A, A1, k-space); an inner whose in and out shapes differ in
    length, or whose batch dims are both unspecified, is rejected."""
    from torchlinops_amd.linops.nufft import all_indices, toeplitz_psf

    assert list(all_indices((2, 3))) == [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2)]
    locs = (torch.rand(3, 20, 2) - 0.5) * 8
    A = tl.NUFFT(locs, (8, 8), output_shape=("R", "K"), batch_shape=("A",), oversamp=2.0)
    dense = tl.Dense(torch.randn(2, 2, dtype=torch.complex64), weightshape=("A", "A1"),
                     ishape=("A", "R", "K"), oshape=("A1", "R", "K"))
    k = toeplitz_psf(A, dense)
    assert isinstance(k, tl.Dense) and tuple(k.weight.shape) == (2, 2, 16, 16)
    assert tuple(map(str, k.weightshape))[:2] == ("A", "A1")
    bad = tl.Dense(torch.randn(2, dtype=torch.complex64), weightshape=("A",),
                   ishape=("A", "R", "K"), oshape=("R", "K"))
    with pytest.raises(ValueError, match="identical input and output shape lengths"):
        toeplitz_psf(A, bad)
