"""Two checks the reference suite makes on its GPU path, restated here:
peak device memory of the fused means stays far below the per-frame spectra
they avoid (reference: tests/test_spectral.py:233-263), and the dispatch
modes agree (reference: tests/test_dispatch_dtypes.py:12-22)."""

import numpy as np
import pytest
import scipy.signal as sps
import torch

import mlx_signal_amd as msig

pytestmark = pytest.mark.gpu

MIB = 1024 * 1024


def _peak_delta(fn):
    """Peak extra device memory (MiB) of one call, after a warm-up call."""
    fn()
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    out = fn()
    torch.cuda.synchronize()
    peak = (torch.cuda.max_memory_allocated() - base) / MIB
    del out
    return peak


def test_fused_means_peak_memory(rng):
    # 64 ch x 2^19, nperseg 1024: 1023 segments; per-frame coherence spectra
    # would take about 512 MiB
    x = torch.from_numpy(rng.standard_normal((64, 1 << 19)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal((64, 1 << 19)).astype(np.float32)).cuda()
    welch_mib = _peak_delta(lambda: msig.welch(x, nperseg=1024))
    coh_mib = _peak_delta(lambda: msig.coherence(x, y, nperseg=1024))
    csd_mib = _peak_delta(lambda: msig.csd(x, y, nperseg=1024))
    assert welch_mib < 16.0, f"welch peak memory {welch_mib:.1f} MiB"
    assert coh_mib < 64.0, f"coherence peak memory {coh_mib:.1f} MiB"
    assert csd_mib < 64.0, f"csd peak memory {csd_mib:.1f} MiB"


def test_dispatch_modes_agree(rng):
    x = rng.standard_normal(4096).astype(np.float32)
    res = {}
    for mode, extra in (("auto", {"gpu_min_size": 0}), ("gpu", {}), ("mlx", {}),
                        ("scipy", {}), ("auto-default", {})):
        cfg = dict(dispatch="auto" if mode == "auto-default" else mode, **extra)
        with msig.config_context(**cfg):
            f, p = msig.welch(x, nperseg=256)
        assert isinstance(p, torch.Tensor) and p.dtype == torch.float32, mode
        res[mode] = p.cpu()
    # auto above the size threshold runs the HIP path: the same bits
    assert torch.equal(res["auto"], res["gpu"]) and torch.equal(res["mlx"], res["gpu"])
    # auto below it (4096 < the default 32768) is the scipy route
    assert torch.equal(res["auto-default"], res["scipy"])
    s = res["scipy"].numpy()
    np.testing.assert_allclose(res["gpu"].numpy(), s, rtol=1e-4, atol=1e-5 * s.max())
    np.testing.assert_allclose(s, sps.welch(x, nperseg=256)[1], rtol=1e-6)
