from .basic import FFT, Dense, Diagonal, Identity, Pad, Scalar, crop_slice_from_pad, pad_to_size
from .chain import Add, Chain
from .interp import ArrayToBlocks, BlocksToArray, Interpolate
from .namedlinop import NamedLinop
from .nufft import NUFFT, cfftn, cifftn, toeplitz_psf
from .sampling import Sampling
from .toeplitz import ToeplitzNormal

__all__ = [
    "NamedLinop", "Chain", "Add", "Identity", "Diagonal", "Scalar", "Pad", "FFT", "Dense",
    "Interpolate", "ArrayToBlocks", "BlocksToArray", "NUFFT", "Sampling", "ToeplitzNormal", "toeplitz_psf",
    "cfftn", "cifftn", "pad_to_size", "crop_slice_from_pad",
]
