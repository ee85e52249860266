from . import learned, poly
from .batched import BatchedNUFFTNormal, SegmentedCG
from .cg import CGRun, DeviceCG, conjugate_gradients, inner
from .poly import apply_polynomial, polynomial_preconditioner

__all__ = ["conjugate_gradients", "CGRun", "DeviceCG", "inner", "BatchedNUFFTNormal",
           "SegmentedCG", "learned", "poly", "polynomial_preconditioner", "apply_polynomial"]
