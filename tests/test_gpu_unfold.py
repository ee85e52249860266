"""GPU parity of array_to_blocks / blocks_to_array: bit-exact against the
oracle (unfold is a copy; fold sums covering blocks in the reference's
lexicographic order)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests._util import crandn, to_dev

pytestmark = pytest.mark.gpu

SHAPES = [
    ((3, 40), (7,), (3,)),
    ((2, 17, 23), (5, 4), (2, 3)),
    ((1, 128, 128), (8, 8), (4, 4)),
    ((2, 9, 10, 8), (3, 4, 2), (2, 3, 1)),
    ((2, 33, 43), (7, 7), (2, 2)),
    ((1, 20, 20, 20), (8, 8, 8), (4, 4, 4)),
    ((2, 11, 12, 10), (11, 12, 10), (1, 2, 1)),
]


def _arr(rng, shape, dtype):
    if dtype == "c64":
        return crandn(rng, shape)
    return rng.standard_normal(shape).astype(np.float32 if dtype == "f32" else np.float64)


@pytest.mark.parametrize("shape,block,stride", SHAPES)
@pytest.mark.parametrize("dtype", ["f32", "f64", "c64"])
def test_unfold_fold_bit_exact(dev, shape, block, stride, dtype):
    import torchlinops_amd.functional as F

    rng = np.random.default_rng(21)
    x = _arr(rng, shape, dtype)
    got = F.array_to_blocks(to_dev(x, dev), block, stride).cpu().numpy()
    want = O.unfold(x, block, stride)
    assert got.shape == want.shape and np.array_equal(got, want)
    xb = _arr(rng, want.shape, dtype)
    gf = F.blocks_to_array(to_dev(xb, dev), shape[-len(block):], block, stride).cpu().numpy()
    wf = O.fold(xb, shape[-len(block):], block, stride)
    assert gf.shape == wf.shape and np.array_equal(gf, wf)


def test_c2_shape_unfold_fold_bit_exact(dev):
    """8 x 128^3 complex64, blocks 8^3, stride 4 (31^3 blocks)."""
    import torchlinops_amd.functional as F

    rng = np.random.default_rng(3)
    x = crandn(rng, (8, 128, 128, 128))
    y = F.array_to_blocks(to_dev(x, dev), (8, 8, 8), (4, 4, 4))
    assert tuple(y.shape) == (8, 31, 31, 31, 8, 8, 8)
    want = O.unfold(x, (8, 8, 8), (4, 4, 4))
    assert np.array_equal(y.cpu().numpy(), want)
    del want
    back = F.blocks_to_array(y, (128, 128, 128), (8, 8, 8), (4, 4, 4)).cpu().numpy()
    wf = O.fold(y.cpu().numpy(), (128, 128, 128), (8, 8, 8), (4, 4, 4))
    assert np.array_equal(back, wf)
    # norm weights: fold(unfold(x)) / fold(unfold(1)) == x where covered
    wts = F.get_norm_weights((128, 128, 128), (8, 8, 8), (4, 4, 4), device=dev).cpu().numpy()
    cov = wts > 0
    assert np.allclose((back / np.where(cov, wts, 1))[:, cov], x[:, cov], rtol=1e-6, atol=1e-6)


def test_mask_and_output_buffers(dev):
    import torchlinops_amd.functional as F

    x = torch.randn(2, 2, 33, 43, device=dev)
    mask = torch.zeros(7, 7, dtype=torch.bool, device=dev)
    mask[3, :] = True
    mask[:, 3] = True
    full = F.array_to_blocks(x, (7, 7), (2, 2))
    sub = F.array_to_blocks(x, (7, 7), (2, 2), mask)
    assert torch.equal(sub, full[..., mask])
    back = F.blocks_to_array(sub, (33, 43), (7, 7), (2, 2), mask)
    ref = full.clone()
    ref[..., ~mask] = 0
    assert torch.equal(back, F.blocks_to_array(ref, (33, 43), (7, 7), (2, 2)))
    # output buffers keep their storage
    y = F.array_to_blocks(x, (7, 7), (2, 2))
    y2 = F.array_to_blocks(torch.randn_like(x), (7, 7), (2, 2), out=y)
    assert y2.data_ptr() == y.data_ptr()
    from torchlinops_amd.functional import fold, unfold

    f = fold(y, (33, 43), (7, 7), (2, 2))
    f2 = fold(torch.randn_like(y), (33, 43), (7, 7), (2, 2), output=f)
    assert f2.data_ptr() == f.data_ptr()
    u = unfold(x, (7, 7), (2, 2))
    u2 = unfold(x, (7, 7), (2, 2), output=u)
    assert u2.data_ptr() == u.data_ptr() and torch.equal(u2, full)


def test_autograd_and_adjoint(dev):
    import torchlinops_amd.functional as F

    x = torch.randn(1, 15, 15, 15, dtype=torch.complex64, device=dev, requires_grad=True)
    y = F.array_to_blocks(x, (4, 4, 4), (2, 2, 2))
    y.real.sum().backward()
    want = F.blocks_to_array(torch.ones_like(y), (15, 15, 15), (4, 4, 4), (2, 2, 2))
    assert torch.equal(x.grad, want)
    z = torch.randn_like(y)
    lhs = torch.vdot(z.flatten(), y.detach().flatten())
    rhs = torch.vdot(F.blocks_to_array(z, (15, 15, 15), (4, 4, 4), (2, 2, 2)).flatten(),
                     x.detach().flatten())
    assert torch.allclose(lhs, rhs, rtol=1e-4)


def test_determinism(dev):
    import torchlinops_amd.functional as F

    xb = torch.randn(3, 19, 19, 8, 8, device=dev) * 1e3
    a = F.blocks_to_array(xb, (83, 83), (8, 8), (4, 4))
    b = F.blocks_to_array(xb, (83, 83), (8, 8), (4, 4))
    assert torch.equal(a, b)
