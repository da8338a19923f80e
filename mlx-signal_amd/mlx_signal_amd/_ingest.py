"""Pipelined host-to-device upload of large NumPy inputs.

The reference's NumPy boundary is one host-to-unified-memory copy
(src/mlx_signal_processing/_array.py:59-60), 54% of its end-to-end welch
time (README.md:60).  Here a large array is uploaded in chunks through two
pinned staging buffers:

* host threads copy chunk k into staging buffer k % 2 (numpy's copy releases
  the GIL, so the memcpy runs on several cores; a float64 -> float32 or
  complex128 -> complex64 downcast happens inside the same copy);
* the DMA of chunk k runs on a dedicated copy stream while chunk k + 1 is
  being staged, and a buffer is refilled only after its previous DMA has
  completed (a recorded event);
* the caller's stream waits for the copy stream before any kernel reads the
  tensor.

The staging buffers are allocated once per device with the allocator's
pinned-memory path and reused.
"""

from __future__ import annotations

import concurrent.futures
import os
import threading

import numpy as np
import torch

__all__ = ["upload", "MIN_BYTES"]

MIN_BYTES = 16 << 20          # smaller arrays take one plain copy
CHUNK_BYTES = 32 << 20
_THREADS = max(1, min(8, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                      else (os.cpu_count() or 1)))

_lock = threading.Lock()
_pool: concurrent.futures.ThreadPoolExecutor | None = None
_state: dict = {}  # device -> (staging buffers, events, copy stream)


def _executor() -> concurrent.futures.ThreadPoolExecutor:
    global _pool
    if _pool is None:
        _pool = concurrent.futures.ThreadPoolExecutor(_THREADS, thread_name_prefix="msig-ingest")
    return _pool


def _device_state(device: torch.device):
    key = str(device)
    st = _state.get(key)
    if st is None:
        bufs = [torch.empty(CHUNK_BYTES, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        events = [None, None]
        stream = torch.cuda.Stream(device=device)  # default priority
        st = (bufs, events, stream)
        _state[key] = st
    return st


def _parallel_copy(dst: np.ndarray, src: np.ndarray) -> None:
    n = src.shape[0]
    parts = min(_THREADS, max(1, n // (1 << 18)))
    if parts <= 1:
        np.copyto(dst, src, casting="same_kind")
        return
    bounds = np.linspace(0, n, parts + 1).astype(np.int64)
    futs = [_executor().submit(np.copyto, dst[a:b], src[a:b], casting="same_kind")
            for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
    for f in futs:
        f.result()


def upload(x: np.ndarray, dtype: torch.dtype, device: torch.device) -> torch.Tensor:
    """x (any real/complex NumPy array) as a contiguous ``dtype`` tensor on
    ``device`` (float32 or complex64), uploaded in pinned chunks."""
    np_dtype = np.float32 if dtype == torch.float32 else np.complex64
    out = torch.empty(x.shape, dtype=dtype, device=device)
    src = np.ascontiguousarray(x).reshape(-1)
    if src.size == 0:
        return out
    esz = np.dtype(np_dtype).itemsize
    per = CHUNK_BYTES // esz
    flat = out.view(-1)
    with _lock:
        bufs, events, stream = _device_state(device)
        stream.wait_stream(torch.cuda.current_stream(device))  # out allocated on the caller's stream
        for k, start in enumerate(range(0, src.size, per)):
            stop = min(src.size, start + per)
            i = k % 2
            if events[i] is not None:
                events[i].synchronize()  # the previous DMA out of this buffer is done
            stage = bufs[i][: (stop - start) * esz].view(dtype)
            _parallel_copy(stage.numpy(), src[start:stop])
            with torch.cuda.stream(stream):
                flat[start:stop].copy_(stage, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
            events[i] = ev
        torch.cuda.current_stream(device).wait_stream(stream)
        out.record_stream(stream)
    return out
