"""Window functions: a cached wrapper over scipy.signal.get_window.

As in the reference (src/mlx_signal_processing/windows.py:19-35) the window
values are computed on the host in float64 (so sum(w) and sum(w**2) keep full
precision) and handed to the device as float32.
"""

from __future__ import annotations

import functools

import numpy as np
import torch
from scipy.signal import get_window as _sp_get_window

__all__ = ["get_window"]


@functools.lru_cache(maxsize=256)
def _window_np(window, Nx: int, fftbins: bool = True) -> np.ndarray:
    """Cached float64 window values (hashable window spec only)."""
    w = _sp_get_window(window, Nx, fftbins=fftbins)
    w = np.ascontiguousarray(w, dtype=np.float64)
    w.setflags(write=False)
    return w


def get_window(window, Nx: int, fftbins: bool = True, device=None) -> torch.Tensor:
    """Return a window of a given length and type as a float32 tensor.

    Accepts everything ``scipy.signal.get_window`` accepts: a name
    (``"hann"``), a ``(name, param)`` tuple (``("kaiser", 8.6)``) or a float
    (kaiser beta).
    """
    from ._device import result_device

    key = tuple(window) if isinstance(window, (list, tuple)) else window
    w = _window_np(key, int(Nx), bool(fftbins))
    return torch.from_numpy(w.astype(np.float32)).to(device or result_device())
