"""istft on the HIP path against the FP64 oracle (itself pinned to
scipy.signal.istft): the fused inverse kernel for one-sided power-of-two
frames, the torch.fft inverse for other shapes, the shared gather
overlap-add, axes, round trips and the NOLA warning.  Error bound: every
sample within 1e-5 x the peak |x| of the FP64 result."""

import warnings

import numpy as np
import pytest
import scipy.signal as sps
import torch

import mlx_signal_amd as msig
import oracle
from _utils import assert_peak_close, peak_rel_err
from mlx_signal_amd import _kernels

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _norm(window, nperseg, nstep, nseg, boundary):
    """FP64 window-power sum of the (trimmed) output."""
    w = oracle.get_window(window, nperseg) if isinstance(window, (str, tuple)) else window
    norm = np.zeros(nperseg + (nseg - 1) * nstep)
    for s in range(nseg):
        norm[s * nstep:s * nstep + nperseg] += w ** 2
    h = nperseg // 2
    return norm[h:-h] if boundary else norm


def assert_istft_close(x, x_ref, norm):
    """Per-sample bound TOL x peak x max(1, sqrt(mean(norm) / norm_i)): the
    divide by the window-power sum amplifies FP32 rounding by 1/w where the
    window is small (frame edges without boundary padding); samples with
    norm <= 1e-10 are not divided and must match exactly."""
    xa = x.cpu().numpy().astype(np.complex128)
    d = np.abs(xa - x_ref)
    peak = np.abs(x_ref).max()
    live = norm > 1e-10
    amp = np.where(live, np.sqrt(norm.mean() / np.where(live, norm, 1.0)), 1.0)
    bound = TOL * peak * np.maximum(1.0, amp)
    assert np.all(d <= bound), f"worst {np.max(d / bound):.2f} x bound"


def _spectra(rng, n, shape=(), **kw):
    x = rng.standard_normal(shape + (n,))
    _, _, z = sps.stft(x, **kw)
    return x, z.astype(np.complex64)


@pytest.mark.parametrize("nperseg", [16, 64, 256, 1024, 4096])
@pytest.mark.parametrize("noverlap_frac", [0.5, 0.75, 0.25])
@pytest.mark.parametrize("boundary", [True, False])
def test_istft_fused_matches_oracle(rng, nperseg, noverlap_frac, boundary):
    noverlap = int(nperseg * noverlap_frac)
    n = 6 * nperseg + 37
    stft_kw = dict(nperseg=nperseg, noverlap=noverlap,
                   boundary="zeros" if boundary else None, padded=boundary)
    _, z = _spectra(rng, n, (2,), **stft_kw)
    kw = dict(nperseg=nperseg, noverlap=noverlap, boundary=boundary)
    fused = "istft_ola" if 2 * noverlap == nperseg else "istft_frames"  # hop = N/2: OLA fused
    before = _kernels.launches[fused]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # no boundary: hann's zero first sample fails NOLA
        t, x = msig.istft(z, **kw)
        t_ref, x_ref = oracle.istft(z, **kw)
    assert _kernels.launches[fused] == before + 1
    assert x.dtype == torch.float32 and x.device.type == "cuda"
    assert tuple(x.shape) == x_ref.shape
    assert_istft_close(x, x_ref, _norm("hann", nperseg, nperseg - noverlap, z.shape[-1], boundary))
    np.testing.assert_allclose(t.cpu().numpy(), t_ref, rtol=1e-6)


def _ola_cases():
    # (nperseg, rows, nseg): one and two frames, odd and even counts, and
    # rows long enough for several units per group and several blocks per
    # row (the overlap carry inside a group and the lead-in across groups)
    return [(16, 2, 1), (16, 1, 2), (16, 3, 1001), (64, 1, 3), (64, 2, 40001),
            (256, 3, 301), (256, 1, 4000), (1024, 1, 5001), (1024, 2, 64),
            (2048, 2, 2049), (4096, 1, 301), (4096, 3, 2)]


@pytest.mark.parametrize("nperseg,rows,nseg", _ola_cases())
def test_fused_istft_equals_frames_then_overlap_add(rng, nperseg, rows, nseg):
    """The fused hop = N/2 kernel gives the two-kernel result, for the whole
    output, the boundary-trimmed span and a random span.  It folds the
    window-power divide into per-sample weights (weff / norm rounded once)
    instead of dividing the sum, so the two differ by a few ulps: bound
    1e-6 x the peak |x|."""
    dev = torch.device("cuda")
    m = nperseg // 2
    z = torch.from_numpy((rng.standard_normal((rows, nseg, m + 1))
                          + 1j * rng.standard_normal((rows, nseg, m + 1))).astype(np.complex64)).to(dev)
    win_np = sps.get_window("hann", nperseg).astype(np.float32)
    win = torch.from_numpy(win_np).to(dev)
    weff = torch.from_numpy((win_np * (win_np.sum() / nperseg)).astype(np.float32)).to(dev)
    frames = _kernels.istft_frames(z, weff, nperseg)
    outlen = (nseg + 1) * m
    a = int(rng.integers(0, outlen))
    spans = [(0, outlen), (m, max(0, outlen - 2 * m)), (a, int(rng.integers(0, outlen - a + 1)))]
    for first, count in spans:
        want = _kernels.overlap_add(frames, win, nperseg, m, first, count)
        got = _kernels.istft_ola(z, weff, win, nperseg, first, count)
        assert got.shape == want.shape
        if count:
            d = float((got - want).abs().max() / want.abs().max().clamp_min(1e-30))
            assert d <= 1e-6, (first, count, d)


def test_istft_routes_half_overlap_to_fused_kernel(rng):
    _, z = _spectra(rng, 1 << 14, (2,), nperseg=256)
    before = dict(_kernels.launches)
    _, x = msig.istft(z, nperseg=256)
    assert _kernels.launches["istft_ola"] == before.get("istft_ola", 0) + 1
    assert _kernels.launches["overlap_add"] == before.get("overlap_add", 0)
    _, x_ref = oracle.istft(z, nperseg=256)
    assert_istft_close(x, x_ref, _norm("hann", 256, 128, z.shape[-1], True))
    msig.istft(z, nperseg=256, noverlap=192)  # other hops keep the two kernels
    assert _kernels.launches["overlap_add"] == before.get("overlap_add", 0) + 1


def test_istft_single_frame_with_boundary_is_empty(rng):
    """One frame with boundary trimming leaves nperseg - 2 (nperseg // 2) = 0
    samples (scipy returns an empty signal)."""
    z = np.ones((129, 1), np.complex64)
    for kw in (dict(nperseg=256), dict(nperseg=256, noverlap=192)):
        _, x = msig.istft(z, **kw)
        _, x_ref = oracle.istft(z, **kw)
        assert tuple(x.shape) == x_ref.shape == (0,)


def test_istft_no_overlap_boxcar(rng):
    """hop == nperseg is invertible with a rectangular window (NOLA holds;
    with a periodic hann the window-power sum is zero at every frame start)."""
    _, z = _spectra(rng, 4096, (2,), nperseg=256, noverlap=0, window="boxcar")
    _, x = msig.istft(z, nperseg=256, noverlap=0, window="boxcar")
    _, x_ref = oracle.istft(z, nperseg=256, noverlap=0, window="boxcar")
    assert peak_rel_err(x, x_ref) <= TOL


@pytest.mark.parametrize("window", ["blackman", ("tukey", 0.25), "hamming"])
@pytest.mark.parametrize("scaling", ["spectrum", "psd"])
def test_istft_windows_and_scaling(rng, window, scaling):
    _, z = _spectra(rng, 3000, (), nperseg=256, window=window, scaling=scaling)
    kw = dict(nperseg=256, window=window, scaling=scaling, fs=8000.0)
    _, x = msig.istft(z, **kw)
    _, x_ref = oracle.istft(z, **kw)
    assert peak_rel_err(x, x_ref) <= TOL


def test_istft_segment_count_edges(rng):
    for nseg_n in (256, 256 + 128, 256 + 256):  # one, two, three frames without boundary
        _, z = _spectra(rng, nseg_n, (), nperseg=256, boundary=None, padded=False)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            _, x = msig.istft(z, nperseg=256, boundary=False)
            _, x_ref = oracle.istft(z, nperseg=256, boundary=False)
        assert_istft_close(x, x_ref, _norm("hann", 256, 128, z.shape[-1], False))


@pytest.mark.parametrize("case", [
    dict(nperseg=63),                      # odd: no fused kernel
    dict(nperseg=64, nfft=128),            # padded frames
    dict(nperseg=100, noverlap=60),        # non power of two
])
def test_istft_composed_shapes(rng, case):
    stft_kw = {k: v for k, v in case.items()}
    _, z = _spectra(rng, 2000, (3,), **stft_kw)
    before = _kernels.launches["istft_frames"]
    _, x = msig.istft(z, **case)
    _, x_ref = oracle.istft(z, **case)
    assert _kernels.launches["istft_frames"] == before
    assert peak_rel_err(x, x_ref) <= TOL


def test_istft_two_sided_complex(rng):
    x = (rng.standard_normal(1500) + 1j * rng.standard_normal(1500)).astype(np.complex64)
    _, _, z = sps.stft(x, nperseg=128, return_onesided=False)
    _, xr = msig.istft(z.astype(np.complex64), nperseg=128, input_onesided=False)
    _, x_ref = oracle.istft(z, nperseg=128, input_onesided=False)
    assert xr.dtype == torch.complex64
    assert peak_rel_err(xr, x_ref) <= TOL


def test_istft_spectrum_width_mismatch(rng):
    """nperseg wider / narrower than the spectra imply: bins padded / cut."""
    _, z = _spectra(rng, 3000, (), nperseg=128)       # 65 bins
    for kw in (dict(nperseg=256, nfft=256), dict(nperseg=64)):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            _, x = msig.istft(z, **kw)
            _, x_ref = oracle.istft(z, **kw)
        assert peak_rel_err(x, x_ref) <= TOL, kw


def test_istft_axes(rng):
    x = rng.standard_normal((3, 2, 900))
    _, _, z = sps.stft(x, nperseg=64)
    zt = np.moveaxis(z, -2, 0).astype(np.complex64)   # (freq, 3, 2, time)
    t, xr = msig.istft(zt, nperseg=64, freq_axis=0, time_axis=-1)
    t_ref, x_ref = oracle.istft(zt, nperseg=64, freq_axis=0, time_axis=-1)
    assert tuple(xr.shape) == x_ref.shape
    assert len(t) == xr.shape[-1]
    assert peak_rel_err(xr, x_ref) <= TOL
    zt2 = np.moveaxis(z, -1, 0).astype(np.complex64)  # (time, 3, 2, freq)
    _, xr2 = msig.istft(zt2, nperseg=64, time_axis=0, freq_axis=-1)
    _, x_ref2 = oracle.istft(zt2, nperseg=64, time_axis=0, freq_axis=-1)
    assert tuple(xr2.shape) == x_ref2.shape
    assert peak_rel_err(xr2, x_ref2) <= TOL


def test_stft_istft_round_trip_on_device(rng):
    x = torch.from_numpy(rng.standard_normal((4, 1 << 15)).astype(np.float32)).cuda()
    for nps, nov in ((256, 128), (1024, 768), (4096, 2048)):
        _, _, z = msig.stft(x, nperseg=nps, noverlap=nov)
        _, xr = msig.istft(z, nperseg=nps, noverlap=nov)
        err = float((xr[:, : x.shape[-1]] - x).abs().max() / x.abs().max())
        assert err <= TOL, (nps, nov, err)


def test_istft_nola_warning(rng):
    z = np.zeros((33, 10), np.complex64)
    with pytest.warns(UserWarning, match="NOLA"):
        msig.istft(z, window=np.zeros(64), nperseg=64)
    with pytest.warns(UserWarning, match="missing boundary"):
        msig.istft(z, window=np.r_[np.zeros(8), np.ones(56)], nperseg=64, boundary=False)


def test_istft_errors(rng):
    z = np.zeros((33, 10), np.complex64)
    with pytest.raises(ValueError):
        msig.istft(np.zeros(10, np.complex64))
    with pytest.raises(ValueError):
        msig.istft(z, freq_axis=0, time_axis=0)
    with pytest.raises(ValueError):
        msig.istft(z, nperseg=64, noverlap=64)
    with pytest.raises(ValueError):
        msig.istft(z, nperseg=64, nfft=32)
    with pytest.raises(ValueError):
        msig.istft(z, scaling="bogus")
    with pytest.raises(ValueError):
        msig.istft(z, window=np.ones(10), nperseg=64)


def test_overlap_add_more_than_65535_rows(rng):
    """The overlap-add kernel strides over all rows (no 65535 grid limit):
    istft with a hop other than nperseg/2 on 70000 short rows."""
    rows, nps, hop, nseg = 70_000, 16, 4, 3
    z = (rng.standard_normal((rows, nps // 2 + 1, nseg))
         + 1j * rng.standard_normal((rows, nps // 2 + 1, nseg))).astype(np.complex64)
    t, x = msig.istft(torch.from_numpy(z).cuda(), nperseg=nps, noverlap=nps - hop)
    for r in (0, 65_535, 69_999):
        _, ref = oracle.istft(z[r], nperseg=nps, noverlap=nps - hop)
        assert_peak_close(x[r], ref, 1e-5)
