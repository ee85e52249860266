from . import learned
from .batched import BatchedNUFFTNormal, SegmentedCG
from .cg import CGRun, DeviceCG, conjugate_gradients, inner

__all__ = ["conjugate_gradients", "CGRun", "DeviceCG", "inner", "BatchedNUFFTNormal",
           "SegmentedCG", "learned"]
