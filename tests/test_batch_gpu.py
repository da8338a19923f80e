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


def _half_requests(rng, dev, sizes=(16, 32, 64, 128, 256, 512, 1024, 2048, 4096)):
    """welch and csd at hop = nperseg/2 over every fused size: the jobs of the
    single persistent launch (odd and even segment counts, 1-3 rows)."""
    reqs, host = [], []
    for i, nps in enumerate(sizes):
        for kind in ("welch", "csd"):
            ch = 1 + (i + len(kind)) % 3
            n = nps * int(rng.integers(1, 12)) + int(rng.integers(0, nps))
            win = "hann" if (i % 2) else "blackman"
            x = make_sig(rng, (ch, n))
            y = (0.5 * x + make_sig(rng, (ch, n))).astype(np.float32) if kind == "csd" else None
            kw = dict(fs=1000.0, window=win, nperseg=nps, detrend="constant" if i % 3 else False)
            host.append((kind, x, y, kw))
            reqs.append(SpectralRequest(kind, torch.from_numpy(x).to(dev),
                                        None if y is None else torch.from_numpy(y).to(dev), **kw))
    return reqs, host


def test_single_launch_matches_per_class_and_oracle(rng):
    """All welch / csd jobs at hop = nfft/2 run in one persistent launch (plus
    the finalize): bitwise the per-class launches' results, and the oracle's
    to 1e-5 of the peak, over all nine fused sizes."""
    dev = torch.device("cuda")
    reqs, host = _half_requests(rng, dev)
    one = SpectralBatch(reqs)
    assert one.launches == 2
    per = SpectralBatch(reqs, per_class_launches=True)
    assert per.launches == 2 * 9 + 1
    a = [o.clone() for o in one.run()]
    b = [o.clone() for o in per.run()]
    torch.cuda.synchronize()
    for i, ((kind, x, y, kw), oa, ob) in enumerate(zip(host, a, b)):
        assert torch.equal(oa, ob), (i, kind, kw["nperseg"])
        assert_peak_close(oa, _oracle(kind, x, y, kw), TOL)


def test_single_launch_mixed_with_other_classes(rng):
    """A batch with coherence, other hops and a radix-1 size keeps those in
    per-class launches beside the single launch; every result is the oracle's
    and bitwise the per-class plan's."""
    dev = torch.device("cuda")
    reqs, host = _requests(rng, dev)
    tuning = {512: {"radix_schedule": 1}}
    one = SpectralBatch(reqs, tuning=tuning)
    per = SpectralBatch(reqs, tuning=tuning, per_class_launches=True)
    # single: welch/csd 256, 1024, 4096, 64 at hop = nfft/2 share one launch
    assert one.launches < per.launches
    a = [o.clone() for o in one.run()]
    b = [o.clone() for o in per.run()]
    for i, ((kind, x, y, kw), oa, ob) in enumerate(zip(host, a, b)):
        assert torch.equal(oa, ob), i
        assert_peak_close(oa, _oracle(kind, x, y, kw), TOL)


def test_single_launch_repeated_runs_and_two_plans(rng):
    """The work counter advances across runs without a reset: five runs of one
    plan, interleaved with runs of a second plan, give the same bits."""
    dev = torch.device("cuda")
    reqs, _ = _half_requests(rng, dev, sizes=(256, 1024, 4096))
    p1 = SpectralBatch(reqs)
    p2 = SpectralBatch(reqs[::-1])
    first = [o.clone() for o in p1.run()]
    second = [o.clone() for o in p2.run()]
    for rep in range(5):
        for o in p1.outputs + p2.outputs:
            o.zero_()
        p1.run()
        p2.run()
        torch.cuda.synchronize()
        for i, o in enumerate(p1.outputs):
            assert torch.equal(o, first[i]), (rep, i)
        for i, o in enumerate(p2.outputs):
            assert torch.equal(o, second[i]), (rep, i)
    for i in range(len(reqs)):
        assert torch.equal(first[i], second[len(reqs) - 1 - i]), i


def test_single_launch_autotune(rng):
    """Tuning for the single launch times tile shapes at radix schedule 0 only,
    and the winner keeps every job in the one launch (2 launches a run)."""
    from mlx_signal_amd.tuning import SINGLE_CANDIDATES, candidates_for, time_candidates

    dev = torch.device("cuda")
    reqs, host = _half_requests(rng, dev, sizes=(256,))
    cands = candidates_for(256, per_class_launches=False)
    assert len(cands) == len(SINGLE_CANDIDATES)
    assert all(c["radix_schedule"] == 0 for c in cands)
    ms, best = time_candidates(reqs, 256, cands, reps=2, per_class_launches=False)
    assert len(ms) == len(cands) and all(t > 0 for t in ms)
    b = SpectralBatch(reqs, tuning={256: cands[best]})
    assert b.launches == 2
    b.run()
    for out, (kind, x, y, kw) in zip(b.outputs, host):
        assert_peak_close(out, _oracle(kind, x, y, kw), TOL)
    b.close()
