"""Windows: scipy.signal.get_window's values, kept in float64 on the host.

The spectral scale factors (sum w, sum w^2) are computed from the float64
values; the kernels read a float32 copy on the device.  Named windows are
computed once per (spec, length, periodic) and shared read-only.
"""

from __future__ import annotations

import collections

import numpy as np
import torch
from scipy.signal import get_window as _scipy_window

__all__ = ["get_window"]

_MAX_CACHED = 256
_cache: "collections.OrderedDict" = collections.OrderedDict()


def window_values(spec, length: int, periodic: bool = True) -> np.ndarray:
    """float64 values of a named window spec (a name, a (name, params...) tuple
    or a Kaiser beta), read-only and shared between callers."""
    key = (spec, int(length), bool(periodic))
    values = _cache.get(key)
    if values is None:
        values = np.array(_scipy_window(spec, int(length), fftbins=bool(periodic)),
                          dtype=np.float64, copy=True)
        values.setflags(write=False)
        _cache[key] = values
        if len(_cache) > _MAX_CACHED:
            _cache.popitem(last=False)
    else:
        _cache.move_to_end(key)
    return values


def get_window(window, Nx: int, fftbins: bool = True, device=None) -> torch.Tensor:
    """The window ``window`` of length ``Nx`` as a float32 tensor.

    ``window`` is anything scipy.signal.get_window takes: a name ("hann"), a
    tuple of a name and its parameters (("kaiser", 8.6)) or a float (a Kaiser
    beta).  ``fftbins`` selects the periodic form (True) or the symmetric
    one.  The tensor is on ``device`` (default: the result device)."""
    from ._device import result_device

    spec = tuple(window) if isinstance(window, list) else window
    values = window_values(spec, Nx, fftbins)
    return torch.tensor(values, dtype=torch.float32, device=device or result_device())
