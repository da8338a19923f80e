"""Inputs and results at the API boundary: float32 / complex64 in, tensors out.

NumPy arrays, lists, scalars and torch tensors become float32 or complex64
tensors on the compute device.  64-bit floating input is computed in 32
bits with a DowncastWarning, or rejected under float64='strict' (the
reference's dtype policy, src/mlx_signal_processing/_array.py).  scipy-route
results come back as tensors of the same 32-bit dtypes.
"""

from __future__ import annotations

import warnings

import numpy as np
import torch

from . import _ingest
from ._config import DowncastWarning, _config

_STRICT_HINT = ("Cast to float32/complex64, or set "
                "mlx_signal_amd.set_config(float64='downcast').")


def _notice_downcast(what: str) -> None:
    """Apply the 64-bit policy to an input about to be computed in 32 bits."""
    if _config.float64 == "strict":
        raise TypeError(
            f"{what} input is not supported with float64='strict' (the GPU path "
            f"computes in float32). {_STRICT_HINT}"
        )
    if _config.warn_on_downcast:
        warnings.warn(
            f"{what} input downcast to 32-bit for the GPU path; results are float32-accurate",
            DowncastWarning,
            stacklevel=4,
        )


_TORCH_TARGET = {torch.float64: torch.float32, torch.complex128: torch.complex64}


def to_device(x, device: torch.device) -> torch.Tensor:
    """``x`` as a float32 or complex64 tensor on ``device``.

    Lists and scalars convert silently; explicit float64/complex128 arrays and
    tensors go through the 64-bit policy; integer and boolean data become
    float32."""
    if isinstance(x, torch.Tensor):
        if x.dtype in (torch.float32, torch.complex64):
            return x.to(device)
        wide = _TORCH_TARGET.get(x.dtype)
        if wide is not None:
            _notice_downcast(str(x.dtype).replace("torch.", ""))
            return x.to(device=device, dtype=wide)
        return x.to(device=device, dtype=torch.complex64 if x.dtype.is_complex else torch.float32)
    if not isinstance(x, np.ndarray):
        a = np.asarray(x)
        if a.dtype.kind not in "fciub":
            raise TypeError(f"unsupported input dtype: {a.dtype}")
        return to_device(a.astype(np.complex64 if a.dtype.kind == "c" else np.float32), device)
    kind = x.dtype.kind
    if kind in "iub":
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    if kind not in "fc":
        raise TypeError(f"unsupported input dtype: {x.dtype}")
    cplx = kind == "c"
    if x.dtype.itemsize > (8 if cplx else 4):
        _notice_downcast(str(x.dtype))
    target = torch.complex64 if cplx else torch.float32
    if device.type == "cuda" and x.size * (8 if cplx else 4) >= _ingest.MIN_BYTES:
        # pipelined pinned upload; the downcast happens inside its copy
        return _ingest.upload(x, target, device)
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex64 if cplx else np.float32)).to(device)


def _is_64bit(x) -> bool:
    if isinstance(x, torch.Tensor):
        return x.dtype in _TORCH_TARGET
    if isinstance(x, np.ndarray):
        return x.dtype.kind in "fc" and x.dtype.itemsize == (16 if x.dtype.kind == "c" else 8)
    return False


def check_strict(x) -> None:
    """Raise under float64='strict' for a 64-bit array or tensor, whichever
    path (GPU or scipy) the call takes: results are 32-bit on both."""
    if _config.float64 == "strict" and _is_64bit(x):
        raise TypeError(
            "float64/complex128 input is not supported with float64='strict' "
            f"(results are float32 on every dispatch path). {_STRICT_HINT}"
        )


def to_numpy(x) -> np.ndarray:
    """A host NumPy view or copy of an array-like."""
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def signal_np(x) -> np.ndarray:
    """A signal argument for a scipy.signal call, after the 64-bit policy."""
    check_strict(x)
    return to_numpy(x)


_NARROW = {"f": np.float32, "c": np.complex64, "i": np.float32, "u": np.float32, "b": np.float32}


def result_to_device(a, device: torch.device) -> torch.Tensor:
    """A scipy-route result as a float32 / complex64 tensor on ``device``."""
    a = np.asarray(a)
    target = _NARROW.get(a.dtype.kind)
    if target is not None and a.dtype != target:
        a = a.astype(target)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def input_size(x) -> int:
    """Element count of an array-like without converting it."""
    return x.numel() if isinstance(x, torch.Tensor) else int(np.size(x))


def input_shape(x) -> tuple:
    """Shape of an array-like without converting it."""
    return tuple(x.shape) if isinstance(x, torch.Tensor) else tuple(np.shape(x))
