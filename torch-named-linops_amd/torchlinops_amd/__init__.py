"""torchlinops_amd -- MI355X-native gridding path of torch-named-linops.

Drop-in replacements for the reference's interpolation / block operators
and the NUFFT built on them, with the arithmetic in hand-written HIP
kernels (libtnl.so, C ABI in include/tnl.h).
"""

from . import config, functional
from ._lib import library_path

__version__ = "0.1.0"

__all__ = ["config", "functional", "library_path"]
