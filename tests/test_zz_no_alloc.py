"""msig_spectral_mean_noalloc (block partials in the library's static device
scratch, no allocation) against msig_spectral_mean (caller's workspace): the
same kernels and plan, so the outputs must be bit-identical.  Each kind runs
two calls in a row, the second on a smaller input, so a stale partial left
in the shared scratch by the first call would show.  Jobs whose partials
exceed the scratch are rejected before anything is launched."""

import ctypes

import numpy as np
import pytest
import torch

from mlx_signal_amd import _hip, _kernels

pytestmark = pytest.mark.gpu

KINDS = {"psd": _hip.PSD_MEAN, "csd": _hip.CSD_MEAN, "trio": _hip.COH_TRIO_MEAN,
         "coherence": _hip.COHERENCE}


def _outputs(code, rows, bins, dev):
    f32 = dict(dtype=torch.float32, device=dev)
    c64 = dict(dtype=torch.complex64, device=dev)
    if code in (_hip.PSD_MEAN, _hip.COHERENCE):
        return [torch.full((rows, bins), float("nan"), **f32)]
    if code == _hip.CSD_MEAN:
        return [torch.full((rows, bins), float("nan"), **c64)]
    return [torch.full((rows, bins), float("nan"), **f32),
            torch.full((rows, bins), float("nan"), **f32),
            torch.full((rows, bins), float("nan"), **c64)]


def _both(kind, rows, n, N, detrend, seed):
    """(workspace outputs, no-allocation outputs) of one job."""
    lib = _hip.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    _hip.check(lib.msig_init(dev.index), "msig_init")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    x = torch.randn(rows, n, device=dev, generator=g)
    y = 0.5 * x + torch.randn(rows, n, device=dev, generator=g)
    win = torch.hann_window(N, periodic=True, dtype=torch.float32, device=dev)
    code = KINDS[kind]
    yy = None if code == _hip.PSD_MEAN else y
    job = _kernels._job(code, x, yy, N, N // 2, N, detrend, True, code != _hip.COHERENCE, 0.25)
    bins = N // 2 + 1
    sz = ctypes.c_size_t(0)
    _hip.check(lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)), "ws")
    assert sz.value <= _hip.NOALLOC_SCRATCH_BYTES
    ws = torch.empty(sz.value, dtype=torch.uint8, device=dev)
    a = _outputs(code, rows, bins, dev)
    b = _outputs(code, rows, bins, dev)
    pa = [_hip.ptr(t) for t in a] + [_hip.ptr(None)] * (3 - len(a))
    pb = [_hip.ptr(t) for t in b] + [_hip.ptr(None)] * (3 - len(b))
    s = _hip.stream_of(x)
    _hip.check(lib.msig_spectral_mean(ctypes.byref(job), _hip.ptr(x), _hip.ptr(yy), _hip.ptr(win),
                                      *pa, _hip.ptr(ws), ws.numel(), s), "mean")
    _hip.check(lib.msig_spectral_mean_noalloc(ctypes.byref(job), _hip.ptr(x), _hip.ptr(yy),
                                              _hip.ptr(win), *pb, s), "mean_noalloc")
    torch.cuda.synchronize()
    return a, b


@pytest.mark.parametrize("kind", ["psd", "csd", "trio", "coherence"])
def test_noalloc_equals_workspace_entry_two_calls(kind):
    calls = [(16, 1 << 20, 1024, "constant"), (3, 40_000, 256, False)]  # second is smaller
    for i, (rows, n, N, detrend) in enumerate(calls):
        a, b = _both(kind, rows, n, N, detrend, seed=10 + i)
        for ta, tb in zip(a, b):
            assert torch.isfinite(ta).all()
            assert torch.equal(ta, tb), (kind, i)


def test_noalloc_rejects_partials_beyond_scratch():
    lib = _hip.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    _hip.check(lib.msig_init(dev.index), "msig_init")
    rows, N = 4096, 4096
    n = N + (N // 2) * 4000
    job = _hip.MsigJob(_hip.COH_TRIO_MEAN, _hip.F_ONESIDED, rows, n, n, n, N, N // 2, N,
                       (n - N) // (N // 2) + 1, 1.0, 0, 0)
    sz = ctypes.c_size_t(0)
    _hip.check(lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)), "ws")
    assert sz.value > _hip.NOALLOC_SCRATCH_BYTES
    one = torch.empty(1, device=dev)
    p = _hip.ptr(one)  # never dereferenced: the size check fails first
    rc = lib.msig_spectral_mean_noalloc(ctypes.byref(job), p, p, p, p, p, p,
                                        _hip.stream_of(one))
    assert rc == _hip.MSIG_EINVAL
    assert b"exceed" in lib.msig_last_error()
