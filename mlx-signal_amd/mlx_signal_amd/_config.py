"""Global configuration: dispatch policy, dtype policy, fallback warnings.

Mirrors the reference's configuration surface (src/mlx_signal_processing/_config.py:42-121).

Dispatch modes
--------------
- ``"auto"`` (default): run the HIP path when the amount of work is at least
  ``gpu_min_size`` elements and a GPU is present, otherwise scipy.signal
  (launch overhead makes the GPU a net loss on tiny inputs).
- ``"gpu"``: always run the HIP path.  Calls it cannot handle raise
  ``NotImplementedError`` instead of falling back.  ``"mlx"`` is accepted as
  an alias so code written for the reference keeps working.
- ``"scipy"``: always use scipy.signal (results still come back as tensors).

Precision (no counterpart in the reference)
-------------------------------------------
- ``"fp32"`` (default): the FP32 kernels.
- ``"fp16_split"`` / ``"bf16_split"``: welch means with
  ``nperseg == nfft == 256`` (real input, constant or no detrend) run their
  two 16-point DFT stages on the matrix cores with hi + lo split FP16 / BF16
  operands and FP32 accumulation.  Every other call keeps the FP32 kernels.
"""

from __future__ import annotations

import contextlib
import dataclasses
import warnings

from ._device import gpu_available

__all__ = [
    "Config",
    "DowncastWarning",
    "FallbackWarning",
    "config_context",
    "get_config",
    "set_config",
]

_DISPATCH_MODES = ("auto", "gpu", "mlx", "scipy")
_FLOAT64_MODES = ("downcast", "strict")
_PRECISIONS = ("fp32", "fp16_split", "bf16_split")


class FallbackWarning(UserWarning):
    """A call was routed to scipy.signal because the GPU path does not cover it."""


class DowncastWarning(UserWarning):
    """float64/complex128 input was downcast to float32/complex64."""


@dataclasses.dataclass
class Config:
    dispatch: str = "auto"
    #: elements of work below which "auto" dispatch uses scipy
    gpu_min_size: int = 1 << 15
    float64: str = "downcast"  # or "strict" (raise on float64 input)
    warn_on_downcast: bool = True
    warn_on_fallback: bool = True
    precision: str = "fp32"  # "fp16_split" / "bf16_split": matrix-core 256-point welch


_config = Config()


def get_config() -> Config:
    """Return a copy of the current global configuration."""
    return dataclasses.replace(_config)


def precision() -> str:
    """The configured arithmetic form of the fused means (no copy: hot path)."""
    return _config.precision


def set_config(**kwargs) -> None:
    """Update global configuration options atomically (all validated first)."""
    for key in kwargs:
        if not hasattr(_config, key):
            raise TypeError(f"unknown config option {key!r}")
    if kwargs.get("dispatch", _config.dispatch) not in _DISPATCH_MODES:
        raise ValueError(f"dispatch must be one of {_DISPATCH_MODES}")
    if kwargs.get("float64", _config.float64) not in _FLOAT64_MODES:
        raise ValueError(f"float64 must be one of {_FLOAT64_MODES}")
    if kwargs.get("precision", _config.precision) not in _PRECISIONS:
        raise ValueError(f"precision must be one of {_PRECISIONS}")
    size = kwargs.get("gpu_min_size", _config.gpu_min_size)
    if not isinstance(size, int) or isinstance(size, bool) or size < 0:
        raise ValueError("gpu_min_size must be a non-negative integer")
    for key, value in kwargs.items():
        setattr(_config, key, value)


@contextlib.contextmanager
def config_context(**kwargs):
    """Temporarily override configuration options within a ``with`` block."""
    old = dataclasses.asdict(_config)
    try:
        set_config(**kwargs)
        yield
    finally:
        _config.__dict__.update(old)


def gpu_pinned() -> bool:
    return _config.dispatch in ("gpu", "mlx")


def use_gpu(work_size: int) -> bool:
    """Decide whether the HIP path should run for ``work_size`` elements
    (reference: _config.use_mlx, _config.py:94-100)."""
    if gpu_pinned():
        return True
    if _config.dispatch == "scipy":
        return False
    return work_size >= _config.gpu_min_size and gpu_available()


def capability_fallback(func_name: str, reason: str) -> None:
    """Record a scipy fallback taken because the GPU path can't handle the call
    (reference: _config.py:103-121).  Raises under dispatch="gpu"."""
    if gpu_pinned():
        raise NotImplementedError(
            f"mlx_signal_amd.{func_name}: {reason} has no GPU path; "
            "use dispatch='auto' to allow the scipy fallback"
        )
    if _config.warn_on_fallback:
        warnings.warn(
            f"mlx_signal_amd.{func_name}: {reason}; falling back to scipy.signal",
            FallbackWarning,
            stacklevel=3,
        )
