"""Continuous-batching server: requests arriving while others run are batched
and overlapped on several streams, and every result is bitwise identical to
the same request computed alone (same tuning), and within 1e-5 x peak of the
FP64 oracle."""

import numpy as np
import pytest
import torch

import oracle
from _utils import assert_peak_close
from mlx_signal_amd.batch import SpectralBatch, SpectralRequest
from mlx_signal_amd.serving import SpectralServer, latency_stats

pytestmark = pytest.mark.gpu


def _schedule(rng, dev, n=40, rate=4000.0):
    t, sched = 0.0, []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        ch = int(rng.integers(1, 9))
        length = int(2 ** rng.uniform(12, 16))
        nps = int(rng.choice([256, 512, 1024]))
        kind = ["welch", "csd", "coherence"][i % 3]
        x = torch.randn(ch, length, device=dev)
        y = torch.randn(ch, length, device=dev) + 0.3 * x if kind != "welch" else None
        sched.append((t, SpectralRequest(kind, x, y, fs=100.0, nperseg=nps)))
    return sched


@pytest.mark.parametrize("tuning", [None, {256: {"target_blocks": 64, "max_run": 4},
                                          1024: {"blocks_per_cu": 2}}])
def test_served_results_equal_alone(rng, tuning):
    dev = torch.device("cuda")
    sched = _schedule(rng, dev)
    srv = SpectralServer(hbm_budget=1 << 30, streams=3, max_inflight=2, max_batch=8,
                         tuning=tuning)
    recs = srv.serve(sched)
    torch.cuda.synchronize()
    assert len(recs) == len(sched)
    assert srv.batches > 1  # requests were batched in several launches
    assert srv.bytes_in_flight == 0
    for (t, r), rec in zip(sched, recs):
        assert rec.done >= rec.admitted >= rec.arrival - 1e-9
        alone = SpectralBatch([r], tuning=tuning).run()[0]
        torch.cuda.synchronize()
        assert torch.equal(alone, rec.result)
    st = latency_stats(recs)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_served_results_match_oracle(rng):
    dev = torch.device("cuda")
    sched = _schedule(rng, dev, n=12)
    recs = SpectralServer(streams=2).serve(sched)
    torch.cuda.synchronize()
    for (_, r), rec in zip(sched, recs):
        x = r.x.cpu().numpy()
        kw = dict(fs=r.fs, nperseg=r.nperseg)
        if r.kind == "welch":
            ref = oracle.welch(x, **kw)[1]
        elif r.kind == "csd":
            ref = oracle.csd(x, r.y.cpu().numpy(), **kw)[1]
        else:
            ref = oracle.coherence(x, r.y.cpu().numpy(), **kw)[1]
        assert_peak_close(rec.result, ref, 1e-5)


def test_admission_respects_budget(rng):
    dev = torch.device("cuda")
    sched = _schedule(rng, dev, n=10)
    from mlx_signal_amd.serving import request_bytes

    biggest = max(request_bytes(r) for _, r in sched)
    srv = SpectralServer(hbm_budget=biggest, streams=2, max_inflight=2)
    peak = 0
    for t, r in sched:
        srv.submit(r, arrival=t)
    while srv.queue or srv.inflight:
        srv.step()
        peak = max(peak, srv.bytes_in_flight)
    assert peak <= biggest
    with pytest.raises(ValueError, match="budget"):
        SpectralServer(hbm_budget=16).submit(sched[0][1])
