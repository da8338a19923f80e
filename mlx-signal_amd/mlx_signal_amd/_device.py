"""Which device results live on.

``torch.cuda`` is ROCm/HIP on this build of PyTorch; the HIP path computes on
the current device and returns tensors there.  Without a GPU the scipy route
returns CPU tensors.
"""

from __future__ import annotations

import functools

import torch


@functools.lru_cache(maxsize=1)
def gpu_available() -> bool:
    return bool(torch.cuda.is_available()) and torch.cuda.device_count() > 0


def compute_device() -> torch.device:
    """The device the HIP path runs on; raises when there is no GPU."""
    if not gpu_available():
        raise RuntimeError(
            "mlx_signal_amd: the HIP path needs a ROCm GPU and none is visible "
            "(use dispatch='auto' or 'scipy' on machines without one)"
        )
    return torch.device("cuda", torch.cuda.current_device())


def result_device() -> torch.device:
    return compute_device() if gpu_available() else torch.device("cpu")
