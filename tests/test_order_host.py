"""The plan's order test (functional._plan.order_coherence), on the CPU:
spokes in acquisition order keep neighbouring samples in one interpolation
bin, the same samples shuffled do not, so only the shuffled list counts as
scattered (planned interpolation then writes through the point-major
scratch)."""

import numpy as np
import torch

from tests._util import radial_3d
from torchlinops_amd.functional._plan import SCATTERED_BELOW, order_coherence


def test_spokes_in_order_are_coherent_and_shuffled_are_scattered():
    rng = np.random.default_rng(3)
    locs = torch.from_numpy(radial_3d(200, 128, 64, rng).reshape(-1, 3).astype(np.float32))
    ordered = order_coherence(locs, (8, 8, 8))
    shuffled = order_coherence(locs[torch.randperm(locs.shape[0])], (8, 8, 8))
    print(f"coherence ordered {ordered:.3f} shuffled {shuffled:.3f}")
    assert ordered > 0.8 > SCATTERED_BELOW > 0.05 > shuffled
    # long lists are sampled in runs: the estimate stays on the same side
    sampled = order_coherence(locs, (8, 8, 8), runs=8, run_len=256)
    sampled_shuffled = order_coherence(locs[torch.randperm(locs.shape[0])], (8, 8, 8),
                                       runs=8, run_len=256)
    assert abs(sampled - ordered) < 0.1 and sampled_shuffled < 0.05


def test_spiral_2d_and_tiny_inputs():
    t = torch.linspace(0, 1, 4096)
    r, a = 60 * t, 2 * np.pi * 16 * t
    locs = torch.stack([64 + r * torch.cos(a), 64 + r * torch.sin(a)], dim=1)
    assert order_coherence(locs, (32, 32)) > 0.9
    assert order_coherence(torch.zeros(1, 2), (16, 16)) == 1.0
    assert order_coherence(torch.zeros(0, 3), (8, 8, 8)) == 1.0
