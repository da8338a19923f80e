"""Pinning the FP64 oracle (CPU only).

Each oracle function is checked against something it does not compute from
its own formulas: the committed golden fixtures (scipy.signal FP64), scipy
on the reference suite's case lists, closed forms, an O(N^2) DFT, and
invariants.  The planted-mistake tests show that plausible slips (a dropped
onesided factor, a wrong detrend, an off-by-one segment count) fail here."""

import json
import warnings
from pathlib import Path

import numpy as np
import pytest
import scipy.signal as sps

import oracle
import oracle.spectral_fp64 as osp
from _utils import make_sig, peak_rel_err

GOLDEN = Path(__file__).resolve().parent / "golden"


def _golden_cases():
    cases = json.loads((GOLDEN / "spectral_golden.json").read_text())
    data = np.load(GOLDEN / "spectral_golden.npz")
    return cases, data


def _kwargs(c):
    kw = dict(c["kwargs"])
    if isinstance(kw.get("window"), list):
        kw["window"] = tuple(kw["window"])
    return kw


def _run(module, c, data):
    name = c["name"]
    ins = [data[f"{name}/in/{k}"] for k in ("x", "y") if f"{name}/in/{k}" in data]
    return getattr(module, c["func"])(*ins, **_kwargs(c))


def test_oracle_matches_golden_fixtures():
    cases, data = _golden_cases()
    assert len(cases) >= 20
    for c in cases:
        out = _run(oracle, c, data)
        for i in range(c["noutputs"]):
            ref = data[f"{c['name']}/out/{i}"]
            got = np.asarray(out[i])
            assert got.shape == ref.shape, (c["name"], i)
            assert peak_rel_err(got, ref) < 1e-12, (c["name"], i)


WELCH_CASES = [
    {}, {"nperseg": 128}, {"nperseg": 128, "noverlap": 96}, {"nperseg": 128, "window": "hamming"},
    {"nperseg": 100, "window": ("tukey", 0.25)}, {"nperseg": 128, "detrend": "linear"},
    {"nperseg": 128, "detrend": False}, {"nperseg": 128, "scaling": "spectrum"},
    {"nperseg": 128, "return_onesided": False}, {"nperseg": 128, "average": "median"},
    {"nperseg": 128, "nfft": 256}, {"nperseg": 129, "noverlap": 37, "fs": 250.0},
    {"nperseg": 64, "window": "blackman", "noverlap": 0},
]


@pytest.mark.parametrize("n", [1024, 999, 4093])
@pytest.mark.parametrize("case", WELCH_CASES)
def test_oracle_welch_vs_scipy_fp64(rng, n, case):
    x = make_sig(rng, (n,)).astype(np.float64)
    f_ref, p_ref = sps.welch(x, **case)
    f, p = oracle.welch(x, **case)
    np.testing.assert_array_equal(f, f_ref)
    assert peak_rel_err(p, p_ref) < 1e-12


@pytest.mark.parametrize("shape,axis", [((3, 2000), -1), ((2000, 3), 0), ((2, 1500, 3), 1)])
def test_oracle_multidim_vs_scipy(rng, shape, axis):
    x = make_sig(rng, shape).astype(np.float64)
    y = make_sig(rng, shape).astype(np.float64)
    assert peak_rel_err(oracle.welch(x, nperseg=128, axis=axis)[1],
                        sps.welch(x, nperseg=128, axis=axis)[1]) < 1e-12
    assert peak_rel_err(oracle.csd(x, y, nperseg=128, axis=axis)[1],
                        sps.csd(x, y, nperseg=128, axis=axis)[1]) < 1e-12
    assert peak_rel_err(oracle.coherence(x, y, nperseg=128, axis=axis)[1],
                        sps.coherence(x, y, nperseg=128, axis=axis)[1]) < 1e-12


def test_oracle_csd_broadcast_unequal_complex(rng):
    x = make_sig(rng, (4096,)).astype(np.float64)
    y = make_sig(rng, (3, 4096)).astype(np.float64)
    assert peak_rel_err(oracle.csd(x, y, nperseg=256)[1], sps.csd(x, y, nperseg=256)[1]) < 1e-12
    x2 = make_sig(rng, (3000,)).astype(np.float64)
    assert peak_rel_err(oracle.csd(x2, y[0], nperseg=256)[1],
                        sps.csd(x2, y[0], nperseg=256)[1]) < 1e-12
    xc = make_sig(rng, (2048,), complex_=True).astype(np.complex128)
    with pytest.warns(UserWarning):
        ref = sps.welch(xc, nperseg=256)[1]
    assert peak_rel_err(oracle.welch(xc, nperseg=256)[1], ref) < 1e-12


@pytest.mark.parametrize("case", [{}, {"nperseg": 128, "noverlap": 100}, {"boundary": "even"},
                                  {"boundary": "odd"}, {"boundary": "constant"},
                                  {"boundary": None}, {"padded": False}, {"scaling": "psd"},
                                  {"detrend": "constant"}, {"return_onesided": False},
                                  {"nperseg": 128, "nfft": 256}])
def test_oracle_stft_vs_scipy(rng, case):
    x = make_sig(rng, (4000,)).astype(np.float64)
    f_ref, t_ref, z_ref = sps.stft(x, fs=8000.0, **case)
    f, t, z = oracle.stft(x, fs=8000.0, **case)
    assert z.shape == z_ref.shape
    np.testing.assert_allclose(t, t_ref, rtol=0, atol=1e-15)
    assert peak_rel_err(z, z_ref) < 1e-12


@pytest.mark.parametrize("mode", ["psd", "complex", "magnitude", "angle", "phase"])
def test_oracle_spectrogram_vs_scipy(rng, mode):
    x = make_sig(rng, (4096,)).astype(np.float64)
    _, _, s_ref = sps.spectrogram(x, fs=100.0, mode=mode)
    _, _, s = oracle.spectrogram(x, fs=100.0, mode=mode)
    assert peak_rel_err(s, s_ref) < 1e-12


@pytest.mark.parametrize("case", [{}, {"nfft": 4096}, {"nfft": 1000}, {"scaling": "spectrum"},
                                  {"detrend": False}, {"window": "hann"},
                                  {"return_onesided": False}])
def test_oracle_periodogram_vs_scipy(rng, case):
    x = make_sig(rng, (2048,)).astype(np.float64)
    assert peak_rel_err(oracle.periodogram(x, **case)[1], sps.periodogram(x, **case)[1]) < 1e-12


@pytest.mark.parametrize("name", ["hann", "hamming", "blackman", "boxcar", ("tukey", 0.25),
                                  ("tukey", 0.5), ("tukey", 1.0)])
@pytest.mark.parametrize("n", [1, 2, 7, 64, 100, 256, 1024, 4096])
def test_oracle_windows_vs_scipy(name, n):
    np.testing.assert_allclose(oracle.get_window(name, n), sps.get_window(name, n), rtol=0,
                               atol=1e-15)


@pytest.mark.parametrize("n", [16, 256, 1024, 4096])
def test_hann_sums_closed_form(n):
    w = oracle.get_window("hann", n)
    assert abs(w.sum() - n / 2) < 1e-9
    assert abs((w * w).sum() - 3 * n / 8) < 1e-9


@pytest.mark.parametrize("n,nperseg,noverlap,nfft", [(1 << 20, 1024, 512, None), (1 << 16, 256, 128, None),
                                                     (4093, 129, 37, None), (4000, 128, 100, 256),
                                                     (64, 1, 0, None)])
def test_plan_integers_closed_form(n, nperseg, noverlap, nfft):
    p = oracle.plan(n, "hann", nperseg, noverlap, nfft, fs=48000.0)
    nstep = nperseg - noverlap
    assert p["nstep"] == nstep
    assert p["nseg"] == (n - nperseg) // nstep + 1
    nf = nfft or nperseg
    assert p["bins"] == nf // 2 + 1
    np.testing.assert_allclose(p["freqs"], np.arange(nf // 2 + 1) * 48000.0 / nf, rtol=1e-15)
    assert len(p["t"]) == p["nseg"]
    assert p["t"][0] == (nperseg / 2) / 48000.0


def test_plan_headline_values():
    p = oracle.plan(1 << 20, "hann", 1024, 512, None, fs=48000.0)
    assert p["nseg"] == 2047 and p["bins"] == 513
    assert abs(p["scale"] - 1.0 / (48000.0 * 384.0)) < 1e-20
    q = oracle.plan(1 << 16, "hann", 256, 128, None, fs=48000.0)
    assert q["nseg"] == 511 and q["bins"] == 129
    s = oracle.plan(4000 + 256, "hann", 256, None, None, padded=True)
    assert (s["n"] - 256) % 128 == 0


def test_dft_brute_force_vs_fft(rng):
    for n in (1, 2, 3, 8, 17, 64):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        np.testing.assert_allclose(oracle.dft(x), np.fft.fft(x), atol=1e-12)


def test_welch_tiny_by_brute_force(rng):
    """welch on a tiny input, rebuilt from the O(N^2) DFT and the textbook
    definitions (mean-detrend, periodic hann, density scale, onesided x2)."""
    x = rng.standard_normal(40)
    N, step, fs = 16, 8, 10.0
    w = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(N) / N)
    segs = [x[s : s + N] - x[s : s + N].mean() for s in range(0, len(x) - N + 1, step)]
    P = np.mean([np.abs(oracle.dft(s * w)) ** 2 for s in segs], axis=0)[: N // 2 + 1]
    P = P / (fs * (w * w).sum())
    P[1:-1] *= 2
    np.testing.assert_allclose(oracle.welch(x, fs=fs, nperseg=N, noverlap=N - step)[1], P,
                               rtol=1e-12)


def test_bin_centred_tone_peak_closed_form():
    """A unit-amplitude tone on bin k0 through a hann window, spectrum scaling:
    peak = A^2/4 * 2 (onesided) = 0.5 exactly; neighbours get (1/2)^2 of it."""
    N, k0 = 256, 19
    n = np.arange(N * 8)
    x = np.cos(2 * np.pi * k0 * n / N)
    _, p = oracle.welch(x, nperseg=N, scaling="spectrum", detrend=False)
    assert abs(p[k0] - 0.5) < 1e-12
    assert abs(p[k0 + 1] - 0.125) < 1e-12 and abs(p[k0 - 1] - 0.125) < 1e-12
    assert np.max(np.delete(p, [k0 - 1, k0, k0 + 1])) < 1e-20


def test_parseval_and_identities(rng):
    x = make_sig(rng, (4096,)).astype(np.float64)
    fs = 10.0
    _, p = oracle.periodogram(x, fs=fs, detrend=False)
    np.testing.assert_allclose(p.sum() * fs / len(x), np.mean(x ** 2), rtol=1e-12)
    _, c = oracle.coherence(x, x, nperseg=256)
    np.testing.assert_allclose(c, 1.0, rtol=1e-12)
    np.testing.assert_allclose(oracle.csd(x, x, nperseg=256)[1].real,
                               oracle.welch(x, nperseg=256)[1], rtol=1e-12)
    y = 0.3 * x + make_sig(rng, (4096,))
    _, c = oracle.coherence(x, y, nperseg=256)
    assert np.all(c >= 0) and np.all(c <= 1 + 1e-12)


def test_median_bias_closed_form():
    assert oracle.median_bias(1) == 1.0
    assert oracle.median_bias(2) == 1.0
    assert abs(oracle.median_bias(3) - 5 / 6) < 1e-15
    assert abs(oracle.median_bias(5) - (1 + (1 / 3 - 1 / 2) + (1 / 5 - 1 / 4))) < 1e-15
    np.testing.assert_array_equal(oracle.onesided_factor(8), [1, 2, 2, 2, 1])
    np.testing.assert_array_equal(oracle.onesided_factor(7), [1, 2, 2, 2])


# ---- planted mistakes: each must be caught by the checks above ----


def _golden_welch_err():
    cases, data = _golden_cases()
    c = next(c for c in cases if c["name"] == "welch_1x2p16")
    out = _run(oracle, c, data)
    return peak_rel_err(out[1], data["welch_1x2p16/out/1"])


def test_planted_dropped_onesided_factor_is_caught(monkeypatch):
    monkeypatch.setattr(osp, "onesided_factor", lambda nfft: np.ones(nfft // 2 + 1))
    assert _golden_welch_err() > 0.1


def test_planted_missing_detrend_is_caught(monkeypatch):
    monkeypatch.setattr(osp, "_detrend", lambda d, kind: d)
    assert _golden_welch_err() > 1e-6


def test_planted_segment_off_by_one_is_caught(monkeypatch):
    orig = osp._frames

    def short(x, nperseg, nstep):
        return orig(x, nperseg, nstep)[..., :-1, :]

    monkeypatch.setattr(osp, "_frames", short)
    assert _golden_welch_err() > 1e-6


def test_planted_wrong_window_is_caught(monkeypatch):
    monkeypatch.setitem(osp._COSINE_SUMS, "hann", (0.5, 0.5, 0.0001))
    assert _golden_welch_err() > 1e-6


# ---------------------------------------------------------------------------
# istft (reference: spectral.py:824-992), pinned to scipy.signal.istft in FP64
# ---------------------------------------------------------------------------

ISTFT_CASES = [
    ({}, {}),
    ({"nperseg": 128}, {"nperseg": 128}),
    ({"nperseg": 128, "noverlap": 96, "window": "blackman"},
     {"nperseg": 128, "noverlap": 96, "window": "blackman"}),
    ({"nperseg": 64, "noverlap": 16}, {"nperseg": 64, "noverlap": 16}),
    ({"nperseg": 63}, {"nperseg": 63}),                                  # odd nperseg
    ({"nperseg": 64, "nfft": 128}, {"nperseg": 64, "nfft": 128}),        # padded frames
    ({"nperseg": 64, "window": ("tukey", 0.25)}, {"nperseg": 64, "window": ("tukey", 0.25)}),
    ({"nperseg": 64, "boundary": None, "padded": False}, {"nperseg": 64, "boundary": False}),
    ({"nperseg": 64, "scaling": "psd", "fs": 50.0}, {"nperseg": 64, "scaling": "psd", "fs": 50.0}),
]


@pytest.mark.parametrize("stft_kw,istft_kw", ISTFT_CASES)
def test_istft_matches_scipy_fp64(stft_kw, istft_kw):
    import scipy.signal as sps

    rng = np.random.default_rng(21)
    x = rng.standard_normal(1500)
    _, _, z = sps.stft(x, **stft_kw)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        t_ref, x_ref = sps.istft(z, **istft_kw)
        t, xo = oracle.istft(z, **istft_kw)
    np.testing.assert_allclose(xo, x_ref, rtol=0, atol=1e-12 * max(1.0, np.abs(x_ref).max()))
    np.testing.assert_allclose(t, t_ref, rtol=1e-15)


def test_istft_two_sided_complex_matches_scipy():
    import scipy.signal as sps

    rng = np.random.default_rng(22)
    x = rng.standard_normal(900) + 1j * rng.standard_normal(900)
    _, _, z = sps.stft(x, nperseg=64, return_onesided=False)
    _, x_ref = sps.istft(z, nperseg=64, input_onesided=False)
    _, xo = oracle.istft(z, nperseg=64, input_onesided=False)
    assert np.iscomplexobj(xo)
    np.testing.assert_allclose(xo, x_ref, rtol=0, atol=1e-12)


def test_istft_axes_match_scipy():
    import scipy.signal as sps

    rng = np.random.default_rng(23)
    x = rng.standard_normal((3, 2, 700))
    _, _, z = sps.stft(x, nperseg=64)                 # (3, 2, freq, time)
    zt = np.moveaxis(z, -2, 0)                        # (freq, 3, 2, time)
    _, x_ref = sps.istft(zt, nperseg=64, freq_axis=0, time_axis=-1)
    t, xo = oracle.istft(zt, nperseg=64, freq_axis=0, time_axis=-1)
    np.testing.assert_allclose(xo, x_ref, rtol=0, atol=1e-12)
    assert len(t) == xo.shape[-1]                     # time vector = output time axis


def test_istft_round_trip_recovers_signal():
    import scipy.signal as sps

    rng = np.random.default_rng(24)
    x = rng.standard_normal((2, 4096))
    for nps, nov in ((256, 128), (256, 192), (1024, 512)):
        _, _, z = sps.stft(x, nperseg=nps, noverlap=nov)
        _, xr = oracle.istft(z, nperseg=nps, noverlap=nov)
        np.testing.assert_allclose(xr[..., :4096], x, rtol=0, atol=1e-12)


def test_istft_nola_warning_and_errors():
    z = np.zeros((33, 10), np.complex128)
    with pytest.warns(UserWarning, match="NOLA"):
        oracle.istft(z, window=np.zeros(64), nperseg=64)
    with pytest.raises(ValueError):
        oracle.istft(np.zeros(10))
    with pytest.raises(ValueError):
        oracle.istft(z, freq_axis=0, time_axis=0)
    with pytest.raises(ValueError):
        oracle.istft(z, nperseg=64, noverlap=64)
    with pytest.raises(ValueError):
        oracle.istft(z, nperseg=64, nfft=32)
    with pytest.raises(ValueError):
        oracle.istft(z, scaling="bogus")


def test_istft_planted_missing_norm_divide_fails():
    """Dropping the window-power divide breaks the round trip."""
    import scipy.signal as sps

    rng = np.random.default_rng(25)
    x = rng.standard_normal(2048)
    _, _, z = sps.stft(x, nperseg=128)
    win = oracle.get_window("hann", 128)
    xs = np.fft.irfft(z.T, n=128) * win.sum() * win
    out = np.zeros(128 + 64 * (xs.shape[0] - 1))
    for s in range(xs.shape[0]):
        out[s * 64:s * 64 + 128] += xs[s]
    out = out[64:-64]
    assert np.abs(out[:2048] - x).max() > 1e-2
