"""ctypes binding of libmsig.so, the gfx950 kernels' C ABI (include/msig.h).

The library is built in-tree (``python mlx-signal_amd/build.py``) and loaded
from this directory; there is no other implementation to fall back to, so a
missing or stale library raises.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch

import os

# MSIG_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = Path(os.environ.get("MSIG_LIBRARY") or Path(__file__).with_name("libmsig.so"))
ABI_VERSION = 8

MSIG_EINVAL = -22
MSIG_ENOTSUP = -95
MSIG_EIO = -5

MIN_NFFT = 16
MAX_NFFT = 4096
NOALLOC_SCRATCH_BYTES = 16 << 20  # MSIG_NOALLOC_SCRATCH_BYTES

# kinds
PSD_MEAN, CSD_MEAN, COH_TRIO_MEAN, COHERENCE = 0, 1, 2, 3
FRAMES_COMPLEX, FRAMES_POWER, CSD_FRAMES, COH_FRAMES = 4, 5, 6, 7

# flags
F_COMPLEX_INPUT = 0x1
F_ONESIDED = 0x2
F_APPLY_FACTOR = 0x4
F_DETREND_CONST = 0x8
F_MFMA_F16 = 0x10
F_MFMA_BF16 = 0x20
MFMA_NFFT = 256


class MsigJob(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("rows", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("ld_x", ctypes.c_int64),
        ("ld_y", ctypes.c_int64),
        ("nperseg", ctypes.c_int32),
        ("nstep", ctypes.c_int32),
        ("nfft", ctypes.c_int32),
        ("nseg", ctypes.c_int32),
        ("scale", ctypes.c_double),
        ("lead", ctypes.c_int64),
        ("nvalid", ctypes.c_int64),
        ("side_src", ctypes.c_void_p),
        ("side_dst", ctypes.c_void_p),
        ("side_len", ctypes.c_int64),
    ]


class MsigBatchJob(ctypes.Structure):
    _fields_ = [
        ("job", MsigJob),
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("win", ctypes.c_void_p),
        ("out0", ctypes.c_void_p),
        ("out1", ctypes.c_void_p),
        ("out2", ctypes.c_void_p),
    ]


class MsigTuning(ctypes.Structure):
    """Per FFT size (index log2(nfft)); 0 selects the library default."""

    _fields_ = [
        ("max_run", ctypes.c_int32 * 13),
        ("target_blocks", ctypes.c_int32 * 13),
        ("blocks_per_cu", ctypes.c_int32 * 13),
        ("radix_schedule", ctypes.c_int32 * 13),
        ("per_class_launches", ctypes.c_int32),
    ]


class MsigTile(ctypes.Structure):
    """One candidate tile shape for one FFT size (msig_tuning's fields at it)."""

    _fields_ = [
        ("max_run", ctypes.c_int32),
        ("target_blocks", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("radix_schedule", ctypes.c_int32),
    ]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t
_JOBP = ctypes.POINTER(MsigJob)

# name -> (restype, argtypes), exactly the declarations of include/msig.h
SIGNATURES = {
    "msig_abi_version": (ctypes.c_int, []),
    "msig_last_error": (ctypes.c_char_p, []),
    "msig_device_count": (ctypes.c_int, []),
    "msig_init": (ctypes.c_int, [ctypes.c_int]),
    "msig_eligible": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32]),
    "msig_workspace_bytes": (ctypes.c_int, [_JOBP, ctypes.POINTER(_sz)]),
    "msig_mean_plan": (ctypes.c_int, [_JOBP, ctypes.POINTER(_i32)]),
    "msig_spectral_mean": (ctypes.c_int, [_JOBP, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "msig_spectral_frames": (ctypes.c_int, [_JOBP, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "msig_spectral_mean_noalloc": (ctypes.c_int, [_JOBP, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "msig_rfft_frames": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i32, _i32, _i32, _i32, _vp, _vp]),
    "msig_rfft_power_mean": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i32, _i32, _i32, _vp, _vp, _sz, _vp]),
    "msig_csd_frames": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _i32, _i32, _i32, _vp, _vp]),
    "msig_coherence_frames": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "msig_csd_mean": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _i32, _i32, _i32, _vp, _vp, _sz, _vp]),
    "msig_coherence_mean": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _sz, _vp]),
    "msig_mean_workspace_bytes": (ctypes.c_int, [_i32, _i64, _i64, _i32, _i32, ctypes.POINTER(_sz)]),
    "msig_batch_create": (ctypes.c_int, [ctypes.POINTER(MsigBatchJob), _i32,
                                         ctypes.POINTER(MsigTuning), ctypes.POINTER(_vp)]),
    "msig_batch_descriptor_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "msig_batch_pack": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "msig_batch_workspace_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "msig_batch_run": (ctypes.c_int, [_vp, _vp, _vp, _sz, _vp]),
    "msig_batch_launches": (ctypes.c_int, [_vp]),
    "msig_batch_destroy": (ctypes.c_int, [_vp]),
    "msig_batch_autotune": (ctypes.c_int, [ctypes.POINTER(MsigBatchJob), _i32,
                                           ctypes.POINTER(MsigTuning), ctypes.POINTER(MsigTile),
                                           _i32, _i32, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(_i32), _vp]),
    "msig_istft_frames": (ctypes.c_int, [_vp, _i64, _i32, _i32, _vp, _vp, _vp]),
    "msig_overlap_add": (ctypes.c_int, [_vp, _i64, _i32, _i32, _i32, _i32, _vp, _i64, _i64, _vp, _vp]),
    "msig_istft_ola": (ctypes.c_int, [_vp, _i64, _i32, _i32, _vp, _vp, _i64, _i64, _vp, _vp, _vp,
                                      _i64, _vp]),
    "msig_median_frames": (ctypes.c_int, [_vp, _i64, _i32, _i32, _i32, ctypes.c_double, _i32, _i32, _vp, _vp]),
    "msig_kernel_timing_arm": (ctypes.c_int, [_i32]),
    "msig_kernel_timing_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i32)]),
}

_lib = None
_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Load libmsig.so once and bind every entry point's prototype."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python mlx-signal_amd/build.py` "
                "(the fused spectral path has no non-HIP implementation)"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        v = lib.msig_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libmsig.so ABI {v} != expected {ABI_VERSION}: rebuild it")
        _lib = lib
    return _lib


def check(rc: int, what: str) -> None:
    """Turn a status code into the exception the reference would raise."""
    if rc == 0:
        return
    msg = load().msig_last_error().decode(errors="replace")
    text = f"{what}: {msg}"
    if rc == MSIG_EINVAL:
        raise ValueError(text)
    if rc == MSIG_ENOTSUP:
        raise NotImplementedError(text)
    raise RuntimeError(text)


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_of_device(dev: torch.device) -> int:
    """The current HIP stream handle of `dev` as an integer."""
    if _raw_stream is not None:
        return _raw_stream(torch.cuda.current_device() if dev.index is None else dev.index)
    return torch.cuda.current_stream(dev).cuda_stream


def stream_of(t: torch.Tensor):
    """The current HIP stream of t's device, as the raw handle the ABI takes
    (torch's raw accessor when present: the Stream object costs a few us)."""
    if _raw_stream is not None:
        idx = t.device.index
        return ctypes.c_void_p(_raw_stream(torch.cuda.current_device() if idx is None else idx))
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())
