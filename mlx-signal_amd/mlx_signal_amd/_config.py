"""Process-wide settings: where calls run, what happens to 64-bit input, and
which warnings are raised.

The option names, defaults, exception types and messages are those of the
reference's configuration surface (src/mlx_signal_processing/_config.py), so
code that configures the reference keeps working.  ``dispatch`` accepts
``"gpu"`` (the HIP path) and keeps ``"mlx"`` as its alias; ``precision`` has
no counterpart there.

dispatch
    ``"auto"``: the HIP path for calls with at least ``gpu_min_size`` elements
    when a GPU is visible, scipy.signal otherwise.  ``"gpu"`` / ``"mlx"``: the
    HIP path always; a call it cannot serve raises NotImplementedError.
    ``"scipy"``: scipy.signal always (results are still tensors).
float64
    ``"downcast"``: 64-bit input is computed in 32 bits (DowncastWarning
    unless ``warn_on_downcast`` is off).  ``"strict"``: 64-bit input raises.
precision
    ``"fp32"``: the FP32 kernels.  ``"fp16_split"`` / ``"bf16_split"``: welch
    means with nperseg == nfft == 256 run their two 16-point DFT stages on
    the matrix cores with hi + lo split operands; other calls keep FP32.
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


class FallbackWarning(UserWarning):
    """A call went to scipy.signal because the GPU path does not cover it."""


class DowncastWarning(UserWarning):
    """A 64-bit input was computed in 32 bits."""


def _one_of(choices):
    def check(name, value):
        if value not in choices:
            return ValueError, f"{name} must be one of {choices}"
        return None

    return check


def _non_negative_int(name, value):
    if isinstance(value, bool) or not isinstance(value, int) or value < 0:
        return ValueError, f"{name} must be a non-negative integer"
    return None


def _anything(name, value):
    return None


# name -> (default, validator); the order is the order of Config's fields
_OPTIONS = {
    "dispatch": ("auto", _one_of(("auto", "gpu", "mlx", "scipy"))),
    "gpu_min_size": (1 << 15, _non_negative_int),
    "float64": ("downcast", _one_of(("downcast", "strict"))),
    "warn_on_downcast": (True, _anything),
    "warn_on_fallback": (True, _anything),
    "precision": ("fp32", _one_of(("fp32", "fp16_split", "bf16_split"))),
}

# A snapshot type with one field per option, for get_config().
Config = dataclasses.make_dataclass(
    "Config",
    [(name, type(default), dataclasses.field(default=default))
     for name, (default, _) in _OPTIONS.items()],
    namespace={"__doc__": "A snapshot of the settings (see the module docstring)."},
)
Config.__module__ = __name__


class _Settings:
    """The live values, read attribute-wise by the hot path (no copies)."""

    def __init__(self):
        self.__dict__.update({name: default for name, (default, _) in _OPTIONS.items()})

    def values(self) -> dict:
        return dict(self.__dict__)


_config = _Settings()


def get_config() -> Config:
    """A copy of the current settings."""
    return Config(**_config.values())


def precision() -> str:
    """The arithmetic form of the fused means."""
    return _config.precision


def set_config(**changes) -> None:
    """Change settings.  Every change is checked before any is applied, so a
    rejected call leaves the settings as they were."""
    merged = _config.values()
    for name in changes:
        if name not in _OPTIONS:
            raise TypeError(f"unknown config option {name!r}")
    merged.update(changes)
    for name, (_, validate) in _OPTIONS.items():
        problem = validate(name, merged[name])
        if problem is not None:
            exc, message = problem
            raise exc(message)
    _config.__dict__.update(changes)


@contextlib.contextmanager
def config_context(**changes):
    """Apply ``changes`` for the duration of a ``with`` block."""
    saved = _config.values()
    try:
        set_config(**changes)
        yield
    finally:
        _config.__dict__.clear()
        _config.__dict__.update(saved)


def gpu_pinned() -> bool:
    """True when dispatch names the GPU path (no scipy fallback allowed)."""
    return _config.dispatch in ("gpu", "mlx")


def use_gpu(work_size: int) -> bool:
    """Whether a call with ``work_size`` elements of work runs on the HIP path
    (counterpart of the reference's use_mlx)."""
    mode = _config.dispatch
    if mode == "scipy":
        return False
    if mode == "auto":
        return work_size >= _config.gpu_min_size and gpu_available()
    return True


def capability_fallback(func_name: str, reason: str) -> None:
    """Handle a call the HIP path cannot serve: raise under a pinned GPU
    dispatch, else warn (FallbackWarning) that scipy.signal takes it."""
    if gpu_pinned():
        raise NotImplementedError(
            f"mlx_signal_amd.{func_name}: {reason} has no GPU path; "
            "use dispatch='auto' to allow the scipy fallback"
        )
    if not _config.warn_on_fallback:
        return
    warnings.warn(
        f"mlx_signal_amd.{func_name}: {reason}; falling back to scipy.signal",
        FallbackWarning,
        stacklevel=3,
    )
