"""mlx_signal_amd: the fused STFT/Welch path of mlx-signal, MI355X-native.

Drop-in for the spectral part of ``mlx_signal_processing``
(``import mlx_signal_amd as msig`` in place of
``import mlx_signal_processing as msig``): same scipy.signal-style functions,
arguments, defaults, warnings and errors; results are float32/complex64 torch
tensors on the ROCm device instead of MLX arrays.  The work runs in
hand-written gfx950 HIP kernels reached through a ctypes C ABI (libmsig.so).
"""

from ._config import (
    Config,
    DowncastWarning,
    FallbackWarning,
    config_context,
    get_config,
    set_config,
)
from .spectral import coherence, csd, istft, periodogram, spectrogram, stft, welch
from .windows import get_window

__all__ = [
    "Config",
    "DowncastWarning",
    "FallbackWarning",
    "coherence",
    "config_context",
    "csd",
    "get_config",
    "get_window",
    "istft",
    "periodogram",
    "set_config",
    "spectrogram",
    "stft",
    "welch",
]

__version__ = "0.1.0"
