"""Array conversion and dtype policy: float32/complex64 in, device tensors out.

Counterpart of the reference's MLX boundary (src/mlx_signal_processing/_array.py:32-130):
NumPy arrays, lists, scalars and torch tensors become float32 or complex64
tensors on the compute device; float64/complex128 inputs are downcast with a
DowncastWarning, or rejected under float64='strict'.
"""

from __future__ import annotations

import warnings

import numpy as np
import torch

from . import _ingest
from ._config import DowncastWarning, _config


def _downcast_notice(src: str) -> None:
    if _config.float64 == "strict":
        raise TypeError(
            f"{src} input is not supported with float64='strict' (the GPU path "
            "computes in float32). Cast to float32/complex64, or set "
            "mlx_signal_amd.set_config(float64='downcast')."
        )
    if _config.warn_on_downcast:
        warnings.warn(
            f"{src} input downcast to 32-bit for the GPU path; results are float32-accurate",
            DowncastWarning,
            stacklevel=4,
        )


def to_device(x, device: torch.device) -> torch.Tensor:
    """Convert input to a float32 or complex64 tensor on ``device``.

    Lists/scalars convert silently; explicit float64/complex128 arrays and
    tensors are downcast with a DowncastWarning (or raise under "strict")."""
    if isinstance(x, torch.Tensor):
        d = x.dtype
        if d in (torch.float32, torch.complex64):
            return x.to(device)
        if d == torch.float64:
            _downcast_notice("float64")
            return x.to(device=device, dtype=torch.float32)
        if d == torch.complex128:
            _downcast_notice("complex128")
            return x.to(device=device, dtype=torch.complex64)
        if d.is_complex:
            return x.to(device=device, dtype=torch.complex64)
        return x.to(device=device, dtype=torch.float32)
    if not isinstance(x, np.ndarray):
        a = np.asarray(x)
        if a.dtype.kind == "c":
            a = a.astype(np.complex64)
        elif a.dtype.kind in "fiub":
            a = a.astype(np.float32)
        return to_device(a, device)
    if x.dtype.kind in "fc":
        cplx = x.dtype.kind == "c"
        if (cplx and x.dtype != np.complex64) or (not cplx and x.dtype.itemsize > 4):
            _downcast_notice(str(x.dtype))
        out_dtype = torch.complex64 if cplx else torch.float32
        if device.type == "cuda" and x.size * (8 if cplx else 4) >= _ingest.MIN_BYTES:
            # pipelined pinned upload; the downcast happens inside its copy
            return _ingest.upload(x, out_dtype, device)
        np_dtype = np.complex64 if cplx else np.float32
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np_dtype)).to(device)
    if x.dtype.kind in "iub":
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    raise TypeError(f"unsupported input dtype: {x.dtype}")


def check_strict(x) -> None:
    """Raise under float64='strict' for 64-bit signal inputs on ANY dispatch path."""
    if _config.float64 != "strict":
        return
    bad = False
    if isinstance(x, torch.Tensor):
        bad = x.dtype in (torch.float64, torch.complex128)
    elif isinstance(x, np.ndarray):
        bad = (x.dtype.kind == "f" and x.dtype.itemsize > 4) or (
            x.dtype.kind == "c" and x.dtype.itemsize > 8
        )
    if bad:
        raise TypeError(
            "float64/complex128 input is not supported with float64='strict' "
            "(results are float32 on every dispatch path). Cast to "
            "float32/complex64, or set mlx_signal_amd.set_config(float64='downcast')."
        )


def to_numpy(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def signal_np(x) -> np.ndarray:
    """to_numpy for signal arguments of scipy fallbacks: enforces the strict policy."""
    check_strict(x)
    return to_numpy(x)


def result_to_device(a, device: torch.device) -> torch.Tensor:
    """A scipy-route result as a float32/complex64 tensor on ``device``
    (reference: _array.py:108-123)."""
    a = np.asarray(a)
    if a.dtype.kind == "f" and a.dtype.itemsize > 4:
        a = a.astype(np.float32)
    elif a.dtype.kind == "c" and a.dtype.itemsize > 8:
        a = a.astype(np.complex64)
    elif a.dtype.kind in "iub":
        a = a.astype(np.float32)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def input_size(x) -> int:
    """Number of elements in an array-like, without forcing a conversion."""
    if isinstance(x, torch.Tensor):
        return x.numel()
    return int(np.size(x))


def input_shape(x) -> tuple:
    if isinstance(x, torch.Tensor):
        return tuple(x.shape)
    return tuple(np.shape(x))
