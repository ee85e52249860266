"""torchlinops_amd -- MI355X-native gridding path of torch-named-linops.

Drop-in replacements for the reference's interpolation / block operators
and the NUFFT built on them, with the arithmetic in hand-written HIP
kernels (libtnl.so, C ABI in include/tnl.h).
"""

from . import alg, config, functional, linops, nameddim
from ._lib import library_path
from .alg import DeviceCG, conjugate_gradients
from .linops import (
    FFT,
    NUFFT,
    Add,
    ArrayToBlocks,
    BlocksToArray,
    Chain,
    Dense,
    Diagonal,
    Identity,
    Interpolate,
    NamedLinop,
    Pad,
    Sampling,
    Scalar,
    toeplitz_psf,
)
from .nameddim import ANY, ELLIPSES, Dim, NamedDimension, NamedShape

__version__ = "0.1.0"

__all__ = [
    "config", "functional", "linops", "alg", "nameddim", "library_path",
    "NamedLinop", "Chain", "Add", "Identity", "Diagonal", "Scalar", "Pad", "FFT", "Dense",
    "Interpolate", "ArrayToBlocks", "BlocksToArray", "NUFFT", "Sampling", "toeplitz_psf",
    "conjugate_gradients", "DeviceCG", "Dim", "NamedDimension", "NamedShape",
    "ELLIPSES", "ANY",
]
