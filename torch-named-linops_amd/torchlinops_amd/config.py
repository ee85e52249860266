"""Global configuration flags.

The first four names and their defaults follow the reference's module-level
flags (src/torchlinops/config.py:32-42) and its `using` context manager
(config.py:51-120).  `interp_plan` and `interp_point_major` are specific
to this library.

reduce_identity_in_normal : bool
    Drop an inner Identity inside ``normal(inner)``.  Default True.
cache_adjoint_normal : bool
    Cache ``.H`` and ``.N``.  Default True.
log_device_transfers : bool
    Kept for API compatibility; this library performs no device transfers.
shape_inference : bool
    Resolve wildcard dims in Chain.  Default False.
interp_plan : str
    ``"auto"`` (default) builds and caches a point-binning plan for 2-D/3-D
    interpolation when the problem is large enough to profit; ``"always"``
    / ``"never"`` force the binned or the direct kernels.
interp_point_major : str
    How planned interpolation writes its output.  ``"auto"`` (default)
    writes each point's channels side by side into a scratch and transposes
    them into the output when the plan found the trajectory listed in
    scattered order and there are at least 8 floats per point (4 complex
    channels); ``"always"`` / ``"never"`` force that form or the direct
    channel-major stores.  Results are the same bits either way.
"""

from __future__ import annotations

import sys

reduce_identity_in_normal: bool = True
cache_adjoint_normal: bool = True
log_device_transfers: bool = True
shape_inference: bool = False
interp_plan: str = "auto"
interp_point_major: str = "auto"


def inner_not_relevant(inner) -> bool:
    from .linops.basic import Identity

    return inner is None or (isinstance(inner, Identity) and reduce_identity_in_normal)


class ConfigContext:
    """Temporarily set config values inside a ``with`` block."""

    VALID_KEYS = {
        "reduce_identity_in_normal",
        "cache_adjoint_normal",
        "log_device_transfers",
        "shape_inference",
        "interp_plan",
        "interp_point_major",
    }

    def __init__(self, **kwargs) -> None:
        for key in kwargs:
            if key not in self.VALID_KEYS:
                raise ValueError(
                    f"Invalid config key: {key!r}. Valid keys: {self.VALID_KEYS}"
                )
        self._changes = kwargs

    def __enter__(self) -> "ConfigContext":
        mod = sys.modules[__name__]
        self._saved = {k: getattr(mod, k) for k in self._changes}
        for k, v in self._changes.items():
            setattr(mod, k, v)
        return self

    def __exit__(self, *args) -> None:
        mod = sys.modules[__name__]
        for k, v in self._saved.items():
            setattr(mod, k, v)


def using(**kwargs) -> ConfigContext:
    return ConfigContext(**kwargs)
