"""The C ABI: libmsig.so loads on a machine without a GPU and exports every
entry point include/msig.h declares (no compute calls here)."""

import ctypes
import re
from pathlib import Path

import pytest

from mlx_signal_amd import _hip

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "msig.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(msig_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_hip.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = ctypes.CDLL(str(_hip.LIB_PATH))
    for name in _declared():
        assert hasattr(lib, name), name
    b = _hip.load()
    assert b.msig_abi_version() == _hip.ABI_VERSION
    assert b.msig_device_count() >= 0


def test_eligibility_predicate():
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()
    for n in (16, 64, 256, 1024, 4096):
        assert lib.msig_eligible(n, n, 1, 1, 0) == 1
        assert lib.msig_eligible(n // 2 + 1, n, 1, 0, 0) == 1  # zero-padded segments
    assert lib.msig_eligible(100, 100, 1, 1, 0) == 0   # non power of two
    assert lib.msig_eligible(8192, 8192, 1, 1, 0) == 0  # beyond the fused range
    assert lib.msig_eligible(256, 128, 1, 1, 0) == 0   # nfft < nperseg
    assert lib.msig_eligible(8, 8, 1, 1, 0) == 0


def test_bad_descriptor_rejected_without_gpu():
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()
    job = _hip.MsigJob(_hip.PSD_MEAN, _hip.F_ONESIDED, 4, 1000, 1000, 1000, 256, 128, 256, 99, 1.0)
    sz = ctypes.c_size_t(0)
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_EINVAL
    assert b"nseg" in lib.msig_last_error()
    job.nseg = (1000 - 256) // 128 + 1
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == 0
    assert sz.value > 0
    job.nfft = 100
    job.nperseg = 100
    job.nseg = (1000 - 100) // 128 + 1
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_ENOTSUP


def test_kernel_timing_contract_without_gpu():
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()
    assert lib.msig_kernel_timing_arm(-1) == _hip.MSIG_EINVAL
    tot = ctypes.c_double(-1.0)
    n = ctypes.c_int32(-1)
    # nothing armed: reads zero launches without touching a device
    assert lib.msig_kernel_timing_read(ctypes.byref(tot), ctypes.byref(n)) == 0
    assert (tot.value, n.value) == (0.0, 0)
    assert lib.msig_kernel_timing_read(None, ctypes.byref(n)) == _hip.MSIG_EINVAL


def test_tuning_struct_matches_header():
    """msig_tuning's fields, in order, are the ctypes structure's."""
    import re

    text = (ROOT / "include" / "msig.h").read_text()
    body = re.search(r"typedef struct msig_tuning \{(.*?)\} msig_tuning;", text, re.S).group(1)
    fields = re.findall(r"int32_t (\w+)(\[13\])?;", body)
    assert [f for f, _ in fields] == [f for f, _ in _hip.MsigTuning._fields_]
    assert ctypes.sizeof(_hip.MsigTuning) == sum(4 * 13 if arr else 4 for _, arr in fields)
    abi = int(re.search(r"#define MSIG_ABI_VERSION (\d+)", text).group(1))
    assert abi == _hip.ABI_VERSION


def test_tile_struct_matches_header():
    """msig_tile's fields, in order, are the ctypes structure's."""
    text = (ROOT / "include" / "msig.h").read_text()
    body = re.search(r"typedef struct msig_tile \{(.*?)\} msig_tile;", text, re.S).group(1)
    fields = re.findall(r"int32_t (\w+);", body)
    assert fields == [f for f, _ in _hip.MsigTile._fields_]
    assert ctypes.sizeof(_hip.MsigTile) == 4 * len(fields)


def test_autotune_rejects_bad_arguments_without_gpu():
    """msig_batch_autotune validates its arguments before any device call."""
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()

    def job(nfft):
        j = _hip.MsigJob(_hip.PSD_MEAN, _hip.F_ONESIDED, 1, 4096, 4096, 4096, nfft, nfft // 2,
                         nfft, (4096 - nfft) // (nfft // 2) + 1, 1.0)
        return _hip.MsigBatchJob(j, 1, None, 1, 1, None, None)  # never dereferenced here

    jobs = (_hip.MsigBatchJob * 2)(job(256), job(1024))
    tiles = (_hip.MsigTile * 1)(_hip.MsigTile(8, 1024, 4, 0))
    ms = (ctypes.c_float * 1)()
    best = ctypes.c_int32(-1)
    rc = lib.msig_batch_autotune(jobs, 2, None, tiles, 1, 3, ms, ctypes.byref(best), None)
    assert rc == _hip.MSIG_EINVAL and b"share nfft" in lib.msig_last_error()
    assert lib.msig_batch_autotune(jobs, 1, None, tiles, 0, 3, ms, ctypes.byref(best), None) == _hip.MSIG_EINVAL
    assert lib.msig_batch_autotune(jobs, 1, None, tiles, 1, 0, ms, ctypes.byref(best), None) == _hip.MSIG_EINVAL
    bad = (_hip.MsigTile * 1)(_hip.MsigTile(8, 1024, 4, 2))
    assert lib.msig_batch_autotune(jobs, 1, None, bad, 1, 3, ms, ctypes.byref(best), None) == _hip.MSIG_EINVAL
    assert best.value == -1


def test_job_struct_matches_header():
    """msig_job's fields, in order, are the ctypes structure's."""
    text = (ROOT / "include" / "msig.h").read_text()
    body = re.search(r"typedef struct msig_job \{(.*?)\} msig_job;", text, re.S).group(1)
    fields = re.findall(r"^\s*(?:int32_t|int64_t|double|float\*|const float\*)\s+(\w+);",
                        body, re.M)
    assert fields == [f for f, _ in _hip.MsigJob._fields_]


def test_zero_extension_is_for_frame_kinds_only():
    """A job with a virtual lead is refused for mean kinds and checked for
    range before any device call."""
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()
    sz = ctypes.c_size_t(0)
    # stft-like: 1000 stored samples, 128 zeros in front, extended length 1000 + 256 + 72
    n_ext = 1000 + 256 + 72
    nseg = (n_ext - 256) // 128 + 1
    job = _hip.MsigJob(_hip.PSD_MEAN, _hip.F_ONESIDED, 2, n_ext, 1000, 1000, 256, 128, 256, nseg,
                       1.0, 128, 1000)
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_ENOTSUP
    assert b"per-frame" in lib.msig_last_error()
    job.kind = _hip.FRAMES_COMPLEX
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == 0
    job.nvalid = n_ext  # lead + stored past the extended row
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_EINVAL
    job.nvalid, job.lead = 1000, -1
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_EINVAL


def test_fused_istft_rejects_bad_arguments_without_gpu():
    """msig_istft_ola validates before any device call (fake pointers are
    never dereferenced on these paths)."""
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()
    p = ctypes.c_void_p(16)
    def call(z, rows, nperseg, first, count):
        return lib.msig_istft_ola(z, rows, 4, nperseg, p, p, first, count, p, None, None, 0, None)

    assert call(None, 1, 256, 0, 10) == _hip.MSIG_EINVAL
    assert call(p, 0, 256, 0, 10) == _hip.MSIG_EINVAL
    assert call(p, 1, 100, 0, 10) == _hip.MSIG_ENOTSUP
    assert call(p, 1, 8192, 0, 10) == _hip.MSIG_ENOTSUP
    # the output spans (nseg + 1) * nperseg / 2 = 640 samples
    assert call(p, 1, 256, 1, 640) == _hip.MSIG_EINVAL
    assert b"outside" in lib.msig_last_error()
    assert call(p, 1, 256, -1, 10) == _hip.MSIG_EINVAL
    assert call(p, 1, 256, 640, 0) == 0  # nothing to do


def test_side_copy_arguments_checked_without_gpu():
    """A side copy needs both buffers and a non-negative length; batch plans
    take none (checked before any device call)."""
    if not _hip.LIB_PATH.exists():
        pytest.skip("libmsig.so not built")
    lib = _hip.load()
    job = _hip.MsigJob(_hip.PSD_MEAN, _hip.F_ONESIDED, 1, 1024, 1024, 1024, 256, 128, 256,
                       7, 1.0, 0, 0, 16, None, 8)
    sz = ctypes.c_size_t(0)
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_EINVAL
    assert b"side copy" in lib.msig_last_error()
    job.side_dst = 32
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == 0
    job.side_len = -1
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_EINVAL
    p = ctypes.c_void_p(16)
    assert lib.msig_istft_ola(p, 1, 4, 256, p, p, 0, 10, p, p, None, 4, None) == _hip.MSIG_EINVAL


def test_noalloc_scratch_limit_matches_header():
    text = (Path(__file__).resolve().parents[1] / "include" / "msig.h").read_text()
    m = re.search(r"#define MSIG_NOALLOC_SCRATCH_BYTES \(\(size_t\)(\d+) << (\d+)\)", text)
    assert m, "MSIG_NOALLOC_SCRATCH_BYTES not found"
    assert int(m.group(1)) << int(m.group(2)) == _hip.NOALLOC_SCRATCH_BYTES
