"""Ragged batches (msig_batch_*): parity with the FP64 oracle, bitwise equality
of each request alone and in the mix, and HIP-graph replays equal to eager
runs."""

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
import oracle
from _utils import assert_peak_close, make_sig
from mlx_signal_amd import _kernels
from mlx_signal_amd.batch import SpectralBatch, SpectralRequest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _requests(rng, dev):
    specs = [
        ("welch", 3, 5000, 256, None, "hann", "constant"),
        ("welch", 1, 1 << 16, 1024, None, "hann", "constant"),
        ("csd", 2, 9000, 256, 100, "blackman", False),
        ("coherence", 4, 20000, 512, None, "hann", "constant"),
        ("welch", 2, 40000, 4096, None, "blackman", False),
        ("csd", 1, 3333, 64, None, "hann", "constant"),
        ("welch", 5, 7777, 128, 32, "hamming", "constant"),
        ("coherence", 2, 12000, 1024, 700, "hann", False),
        ("csd", 3, 50000, 4096, None, "hann", "constant"),
    ]
    reqs, host = [], []
    for kind, ch, n, nps, nov, win, det in specs:
        x = make_sig(rng, (ch, n))
        y = (0.5 * x + make_sig(rng, (ch, n))).astype(np.float32) if kind != "welch" else None
        host.append((kind, x, y, dict(fs=1000.0, window=win, nperseg=nps, noverlap=nov, detrend=det)))
        reqs.append(SpectralRequest(kind, torch.from_numpy(x).to(dev),
                                    None if y is None else torch.from_numpy(y).to(dev),
                                    fs=1000.0, window=win, nperseg=nps, noverlap=nov, detrend=det))
    return reqs, host


def _oracle(kind, x, y, kw):
    if kind == "welch":
        return oracle.welch(x, **kw)[1]
    if kind == "csd":
        return oracle.csd(x, y, **kw)[1]
    kw = dict(kw)
    return oracle.coherence(x, y, **kw)[1]


def test_batch_matches_oracle_and_standalone(rng):
    dev = torch.device("cuda")
    reqs, host = _requests(rng, dev)
    b = SpectralBatch(reqs)
    assert b.launches >= 2
    outs = [o.clone() for o in b.run()]
    torch.cuda.synchronize()
    for (kind, x, y, kw), out in zip(host, outs):
        assert_peak_close(out, _oracle(kind, x, y, kw), TOL)
        # the single-call API runs the same block geometry: same bits
        fn = {"welch": lambda: msig.welch(x, **kw)[1], "csd": lambda: msig.csd(x, y, **kw)[1],
              "coherence": lambda: msig.coherence(x, y, **kw)[1]}[kind]
        assert torch.equal(fn(), out), kind


@pytest.mark.parametrize("tuning", [None, {256: {"target_blocks": 64, "max_run": 4, "blocks_per_cu": 2},
                                          1024: {"target_blocks": 128, "max_run": 8},
                                          4096: {"blocks_per_cu": 1}},
                                    {64: {"radix_schedule": 1}, 128: {"radix_schedule": 1},
                                     512: {"radix_schedule": 1},
                                     1024: {"radix_schedule": 1, "max_run": 8}}])
def test_batch_alone_equals_in_mix(rng, tuning):
    dev = torch.device("cuda")
    reqs, _ = _requests(rng, dev)
    mix = SpectralBatch(reqs, tuning=tuning)
    mixed = [o.clone() for o in mix.run()]
    for i, r in enumerate(reqs):
        alone = SpectralBatch([r], tuning=tuning).run()[0]
        assert torch.equal(alone, mixed[i]), i


def test_batch_graph_replays_match_eager(rng):
    dev = torch.device("cuda")
    reqs, _ = _requests(rng, dev)
    b = SpectralBatch(reqs)
    eager = [o.clone() for o in b.run()]
    g = b.graph()
    other = torch.randn(1 << 20, device=dev)
    for rep in range(3):
        for o in b.outputs:
            o.zero_()
        g.replay()
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(b.outputs, eager)):
            assert torch.equal(o, e), (rep, i)
        # eager work between replays
        other = torch.fft.rfft(other).abs()[: 1 << 19].repeat(2)
        msig.welch(torch.randn(4, 1 << 14, device=dev), nperseg=256)
        torch.cuda.synchronize()


def test_batch_rejects_ineligible(rng):
    dev = torch.device("cuda")
    x = torch.randn(2, 4000, device=dev)
    with pytest.raises(ValueError):
        SpectralBatch([SpectralRequest("welch", x, nperseg=100)])  # non power-of-two nfft
    with pytest.raises(ValueError):
        SpectralBatch([SpectralRequest("welch", x, nperseg=256, detrend="linear")])
    with pytest.raises(ValueError):
        SpectralBatch([SpectralRequest("csd", x, nperseg=256)])  # y missing


def test_batch_radix_schedule_matches_oracle(rng):
    """Remainder-radix-first kernels (nfft 64, 128, 512, 1024) against the
    oracle; alone-equals-in-mix under such a tuning is in the test above."""
    dev = torch.device("cuda")
    reqs, host = _requests(rng, dev)
    tuning = {n: {"radix_schedule": 1} for n in (64, 128, 512, 1024)}
    outs = [o.clone() for o in SpectralBatch(reqs, tuning=tuning).run()]
    for (kind, x, y, kw), out in zip(host, outs):
        assert_peak_close(out, _oracle(kind, x, y, kw), TOL)


def test_native_autotune_times_every_candidate(rng):
    """msig_batch_autotune (through tuning.time_candidates) returns a positive
    median per candidate and the index of the smallest; the winner's plan runs
    and matches the oracle."""
    from mlx_signal_amd.tuning import candidates_for, time_candidates

    dev = torch.device("cuda")
    reqs, host = [], []
    for ch, n, kind in ((2, 20000, "welch"), (3, 9000, "csd"), (1, 70000, "welch")):
        x = make_sig(rng, (ch, n))
        y = (0.5 * x + make_sig(rng, (ch, n))).astype(np.float32) if kind == "csd" else None
        host.append((kind, x, y, dict(fs=1.0, window="hann", nperseg=1024)))
        reqs.append(SpectralRequest(kind, torch.from_numpy(x).to(dev),
                                    None if y is None else torch.from_numpy(y).to(dev), nperseg=1024))
    cands = candidates_for(1024, [dict(target_blocks=256, max_run=8, blocks_per_cu=2),
                                  dict(target_blocks=1024, max_run=32, blocks_per_cu=4),
                                  dict(target_blocks=4096, max_run=8, blocks_per_cu=8)])
    assert len(cands) == 6  # 1024 = 16 * 16 * 4: two radix schedules
    ms, best = time_candidates(reqs, 1024, cands, reps=3)
    assert len(ms) == len(cands) and all(t > 0 for t in ms)
    assert best == min(range(len(ms)), key=lambda i: ms[i])
    b = SpectralBatch(reqs, tuning={1024: cands[best]})
    b.run()
    for out, (kind, x, y, kw) in zip(b.outputs, host):
        assert_peak_close(out, _oracle(kind, x, y, kw), TOL)
    b.close()
