"""The headline workload itself on the HIP path, against the FP64 oracle.

welch, csd and coherence over 64 channels x 2^20 float32 (x = 3 unit
sinusoids + N(0,1), y = a causal 32-tap FIR of x + 0.1 N(0,1), generated on
the device from a seed, exactly as the benchmark does), fs = 48000,
nperseg = 1024, noverlap = 512, hann, constant detrend, density, mean.  At
this shape every FFT group walks many segment pairs in the hop = N/2 loop
that carries a shared half-block and its detrend sum from one pair to the
next, so these tests pin that loop at the benchmark's own geometry.

Also: shapes whose launch plan gives several pairs per group (run >= 2) with
an odd segment count and an uneven last block, for all three kinds at
nperseg 256 and 1024, with and without detrend; and the committed golden
fixtures (scipy.signal in FP64 on float32 inputs) fed through the GPU path.
"""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
import oracle
from _utils import as_np, assert_peak_close
from mlx_signal_amd import _kernels

pytestmark = pytest.mark.gpu

TOL = 1e-5
GOLDEN = Path(__file__).resolve().parent / "golden"
FS = 48000.0
KW = dict(fs=FS, window="hann", nperseg=1024, noverlap=512)


def _rows_close(got, x, y, kind, rows):
    """Rows `rows` of a device result against the oracle on the same rows."""
    xs = x[rows].cpu().numpy()
    ys = None if y is None else y[rows].cpu().numpy()
    if kind == "welch":
        ref = oracle.welch(xs, **KW)[1]
    elif kind == "csd":
        ref = oracle.csd(xs, ys, **KW)[1]
    else:
        ref = oracle.coherence(xs, ys, **KW)[1]
    g = as_np(got[rows])
    for i in range(len(rows)):  # each channel against its own peak
        assert_peak_close(g[i], ref[i], TOL)


@pytest.fixture(scope="module")
def headline_xy():
    import bench

    dev = torch.device("cuda")
    x, y = bench.make_inputs(torch, bench.C, bench.NS, 11, dev)
    torch.cuda.synchronize()
    yield x, y
    del x, y
    torch.cuda.empty_cache()


ROWS = [0, 31, 63]


def test_headline_welch_matches_oracle(headline_xy):
    x, _ = headline_xy
    before = _kernels.launches["mean_psd"]
    f, p = msig.welch(x, **KW)
    assert _kernels.launches["mean_psd"] == before + 1
    assert tuple(p.shape) == (64, 513)
    np.testing.assert_array_equal(as_np(f), np.fft.rfftfreq(1024, 1 / FS).astype(np.float32))
    _rows_close(p, x, None, "welch", ROWS)


def test_headline_csd_matches_oracle(headline_xy):
    x, y = headline_xy
    before = _kernels.launches["mean_csd"]
    _, p = msig.csd(x, y, **KW)
    assert _kernels.launches["mean_csd"] == before + 1
    assert p.dtype == torch.complex64 and tuple(p.shape) == (64, 513)
    _rows_close(p, x, y, "csd", ROWS)


def test_headline_coherence_matches_oracle(headline_xy):
    x, y = headline_xy
    before = _kernels.launches["mean_coherence"]
    _, c = msig.coherence(x, y, **KW)
    assert _kernels.launches["mean_coherence"] == before + 1
    _rows_close(c, x, y, "coherence", ROWS)


def test_headline_plan_walks_several_pairs_per_group():
    """The headline plan really runs the carried-half loop over many pairs."""
    g = _kernels.mean_geometry("psd", 64, 1 << 20, 1024, 512, 1024)
    assert g["run"] >= 8, g
    assert g["units"] == 1024, g


def _uneven_case(kind, N, detrend, min_run=2):
    """A shape whose plan gives several pairs per group, an odd segment
    count (a partner-less last unit) and a last block that its groups do not
    fill evenly; found by asking the library for its plan."""
    hop = N // 2
    for rows in (1, 2, 3, 5):
        for nseg in range(2 * 64 * 32 + 1, 2 * 64 * 400, 2 * 37 + 2):
            n = N + hop * (nseg - 1) + 13
            g = _kernels.mean_geometry(kind, rows, n, N, hop, N)
            gpb, run, bpr, units = g["gpb"], g["run"], g["bpr"], g["units"]
            last = units - (bpr - 1) * gpb * run
            if run >= min_run and nseg % 2 == 1 and 0 < last < gpb * run and last % run:
                return rows, n, g
    raise AssertionError(f"no uneven shape found for {kind} N={N}")


@pytest.mark.parametrize("N", [256, 1024])
@pytest.mark.parametrize("detrend", [False, "constant"])
@pytest.mark.parametrize("kind", ["psd", "csd", "coherence"])
def test_multi_pair_runs_uneven_last_block(rng, N, detrend, kind):
    rows, n, g = _uneven_case(kind, N, detrend)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    t = np.arange(n, dtype=np.float32)
    x += np.sin(2 * np.pi * 0.1234 * t).astype(np.float32) + 0.25
    y = (0.5 * x + rng.standard_normal((rows, n))).astype(np.float32)
    kw = dict(nperseg=N, noverlap=N // 2, detrend=detrend)
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y).cuda()
    if kind == "psd":
        got, ref = msig.welch(xd, **kw)[1], oracle.welch(x, **kw)[1]
    elif kind == "csd":
        got, ref = msig.csd(xd, yd, **kw)[1], oracle.csd(x, y, **kw)[1]
    else:
        got, ref = msig.coherence(xd, yd, **kw)[1], oracle.coherence(x, y, **kw)[1]
    g_np = as_np(got)
    for r in range(rows):
        assert_peak_close(g_np[r], ref[r], TOL)


def _golden():
    cases = json.loads((GOLDEN / "spectral_golden.json").read_text())
    data = np.load(GOLDEN / "spectral_golden.npz")
    return cases, data


@pytest.mark.parametrize("name", [c["name"] for c in _golden()[0]])
def test_golden_fixture_through_gpu(name):
    """Committed scipy.signal FP64 outputs (tests/golden) against the GPU path
    on the same float32 inputs: grids exact to float32, values within
    1e-5 x peak."""
    cases, data = _golden()
    c = next(c for c in cases if c["name"] == name)
    kw = dict(c["kwargs"])
    if isinstance(kw.get("window"), list):
        kw["window"] = tuple(kw["window"])
    ins = [torch.from_numpy(data[f"{name}/in/{k}"]).cuda() for k in c["inputs"]]
    out = getattr(msig, c["func"])(*ins, **kw)
    assert len(out) == c["noutputs"]
    for i, got in enumerate(out):
        ref = data[f"{name}/out/{i}"]
        assert tuple(got.shape) == ref.shape, (name, i)
        assert got.device.type == "cuda"
        if i < c["noutputs"] - 1:  # frequency / time grids
            np.testing.assert_allclose(as_np(got), ref.astype(np.float32), rtol=0, atol=0)
        else:
            assert_peak_close(got, ref, TOL)
