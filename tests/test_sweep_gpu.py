"""Seeded random sweep over the planner's options on the HIP path, each case
against the FP64 oracle (itself pinned to scipy.signal): nperseg (powers of
two across the fused range and a few others that take the composed path),
noverlap, nfft padding, detrend, scaling, window, sides, average, axis,
complex input and DC-offset signals; and istft over sizes, windows, hops and
boundary trimming.  Frequency and time grids are exact.

Bound: 1e-5 x the peak |ref| bin, or, where a case is ill-conditioned in
FP32 (a Kaiser window's tiny edge weights, a coherence ratio over few
segments, a large DC offset under detrending), 8 x the error of scipy.signal
run on the same float32 data (the reference's own CPU route computes in
float32).  Some of these cases take scipy's float32 route above 1e-5 itself:
1.1e-5 to 1.3e-3 x peak in this sweep."""

import warnings

import numpy as np
import pytest
import scipy.signal as sps

import mlx_signal_amd as msig
import oracle
from _utils import as_np, peak_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
REL_TO_F32 = 8.0


def check(got, ref64, f32_ref):
    """got vs the FP64 reference, bounded as the module docstring says."""
    assert tuple(got.shape) == np.shape(ref64)
    e = peak_rel_err(got, ref64)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        e32 = peak_rel_err(f32_ref(), ref64)
    assert e <= max(TOL, REL_TO_F32 * e32), f"error {e:.3e} x peak (scipy float32: {e32:.3e})"


NPS = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 48, 100, 300]
WINDOWS = ["hann", "hamming", "blackman", ("tukey", 0.25), "boxcar", ("kaiser", 8.0)]


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    nps = int(rng.choice(NPS))
    kw = dict(
        nperseg=nps,
        noverlap=int(rng.integers(0, nps)),
        window=WINDOWS[int(rng.integers(len(WINDOWS)))],
        detrend=[False, "constant", "linear"][int(rng.integers(3))],
        scaling=["density", "spectrum"][int(rng.integers(2))],
        fs=float(rng.choice([1.0, 250.0, 48000.0])),
    )
    if rng.random() < 0.3:
        kw["nfft"] = nps * int(rng.choice([2, 4]))
    rows = int(rng.integers(1, 4))
    n = nps * int(rng.integers(1, 6)) + int(rng.integers(0, nps))
    cplx = rng.random() < 0.15
    x = rng.standard_normal((rows, n)).astype(np.float32)
    y = (0.4 * x + rng.standard_normal((rows, n))).astype(np.float32)
    if rng.random() < 0.2:
        x = x + np.float32(rng.uniform(-50, 50))  # a DC offset the detrend must remove
    if cplx:
        x = (x + 1j * rng.standard_normal((rows, n))).astype(np.complex64)
    onesided = bool(rng.random() < 0.8)
    axis = -1
    if rng.random() < 0.25:
        x, y = x.T.copy(), y.T.copy()
        axis = 0
    return rng, kw, x, y, onesided, axis


@pytest.mark.parametrize("seed", range(48))
def test_random_welch_csd(seed):
    rng, kw, x, y, onesided, axis = _case(seed)
    average = "median" if rng.random() < 0.25 else "mean"
    opts = dict(return_onesided=onesided, axis=axis, average=average, **kw)
    f, p = msig.welch(x, **opts)
    f_ref, p_ref = oracle.welch(x, **opts)
    np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
    check(p, p_ref, lambda: sps.welch(x, **opts)[1])
    if not np.iscomplexobj(x):
        f, c = msig.csd(x, y, **opts)
        f_ref, c_ref = oracle.csd(x, y, **opts)
        np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
        check(c, c_ref, lambda: sps.csd(x, y, **opts)[1])


@pytest.mark.parametrize("seed", range(24))
def test_random_coherence_spectrogram_stft(seed):
    rng, kw, x, y, onesided, axis = _case(100 + seed)
    kw.pop("scaling")
    if not np.iscomplexobj(x):
        f, cxy = msig.coherence(x, y, axis=axis, **kw)
        f_ref, c_ref = oracle.coherence(x, y, axis=axis, **kw)
        np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
        check(cxy, c_ref, lambda: sps.coherence(x, y, axis=axis, **kw)[1])
    mode = ["psd", "complex", "magnitude"][int(rng.integers(3))]
    opts = dict(axis=axis, mode=mode, return_onesided=onesided, **kw)
    f, t, s = msig.spectrogram(x, **opts)
    f_ref, t_ref, s_ref = oracle.spectrogram(x, **opts)
    np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
    np.testing.assert_array_equal(as_np(t), t_ref.astype(np.float32))
    check(s, s_ref, lambda: sps.spectrogram(x, **opts)[2])
    kw.pop("detrend")
    opts = dict(axis=axis, return_onesided=onesided, **kw)
    f, t, z = msig.stft(x, **opts)
    f_ref, t_ref, z_ref = oracle.stft(x, **opts)
    np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
    np.testing.assert_array_equal(as_np(t), t_ref.astype(np.float32))
    check(z, z_ref, lambda: sps.stft(x, **opts)[2])


@pytest.mark.parametrize("seed", range(16))
def test_random_istft(seed):
    rng = np.random.default_rng(5000 + seed)
    nps = int(rng.choice([16, 64, 256, 1024, 4096, 100, 48]))
    window = ["hann", "hamming", "blackman", ("tukey", 0.5)][int(rng.integers(4))]
    noverlap = [nps // 2, (3 * nps) // 4, nps // 2 + 1][int(rng.integers(3))]
    boundary = bool(rng.random() < 0.7)
    kw = dict(fs=float(rng.choice([1.0, 48000.0])), window=window, nperseg=nps, noverlap=noverlap)
    x = rng.standard_normal((int(rng.integers(1, 4)), nps * int(rng.integers(2, 7)) + 7))
    _, _, z = sps.stft(x, boundary="zeros" if boundary else None, padded=boundary, **kw)
    z = z.astype(np.complex64)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # windows without boundary padding can fail NOLA
        t, xr = msig.istft(z, boundary=boundary, **kw)
        t_ref, x_ref = oracle.istft(z, boundary=boundary, **kw)
    np.testing.assert_array_equal(as_np(t), t_ref.astype(np.float32))
    check(xr, x_ref, lambda: sps.istft(z, boundary=boundary, **kw)[1])
