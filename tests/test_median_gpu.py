"""The per-bin median kernel (msig_median_frames): exact against numpy.median
on adversarial columns, and welch / csd with average='median' against the
FP64 oracle (every bin within 1e-5 x the peak bin)."""

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
import oracle
from _utils import assert_peak_close, make_sig
from mlx_signal_amd import _kernels

pytestmark = pytest.mark.gpu


def _np_median(a, scale=1.0):
    return (np.median(a.astype(np.float64), axis=1) * scale).astype(np.float32)


@pytest.mark.parametrize("nseg", [1, 2, 3, 4, 7, 64, 255, 256, 512, 513, 1023, 1024, 1025, 2047, 2048,
                                  2049, 4097, 8192, 8193, 16000])
def test_median_kernel_exact_random(rng, nseg):
    rows, bins = 3, 37
    a = rng.standard_normal((rows, nseg, bins)).astype(np.float32)
    a[:, :, 5] = np.round(a[:, :, 5])            # many ties
    a[:, :, 6] = 0.0                              # all equal
    a[:, ::2, 7] = -0.0                           # signed zeros
    a[:, :, 8] = np.abs(a[:, :, 8]) * 1e30        # large exponents
    a[:, :, 9] = np.abs(a[:, :, 9]) * 1e-30       # small exponents
    got = _kernels.median_frames(torch.from_numpy(a).cuda(), 1.0, False, 0).cpu().numpy()
    want = np.median(a, axis=1)                   # numpy in float32 (mean of two middles)
    np.testing.assert_array_equal(got, want.astype(np.float32))


def test_median_kernel_nan_and_factor(rng):
    a = rng.exponential(size=(2, 101, 9)).astype(np.float32)
    a[1, 50, 4] = np.nan
    got = _kernels.median_frames(torch.from_numpy(a).cuda(), 0.5, True, 16).cpu().numpy()
    want = np.median(a, axis=1) * 0.5
    f = np.full(9, 2.0, np.float32)
    f[0] = f[8] = 1.0                              # nfft 16: DC and Nyquist keep factor 1
    want = (want * f).astype(np.float32)
    assert np.isnan(got[1, 4])
    mask = ~np.isnan(want)
    np.testing.assert_allclose(got[mask], want[mask], rtol=1e-6)
    # odd nfft: no Nyquist bin, every bin but DC doubled
    got_odd = _kernels.median_frames(torch.from_numpy(a[:1]).cuda(), 1.0, True, 17).cpu().numpy()
    f_odd = np.full(9, 2.0, np.float32)
    f_odd[0] = 1.0
    np.testing.assert_allclose(got_odd, np.median(a[:1], axis=1) * f_odd, rtol=1e-6)


def test_median_kernel_complex_columns(rng):
    a = (rng.standard_normal((2, 300, 20)) + 1j * rng.standard_normal((2, 300, 20))).astype(np.complex64)
    got = _kernels.median_frames(torch.from_numpy(a).cuda(), 1.0, False, 0).cpu().numpy()
    want = np.median(a.real, axis=1) + 1j * np.median(a.imag, axis=1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)


@pytest.mark.parametrize("n", [1024, 4093, 20000])
@pytest.mark.parametrize("nperseg", [64, 256])
def test_welch_csd_median_match_oracle(rng, n, nperseg):
    x = make_sig(rng, (2, n))
    y = (0.5 * x + make_sig(rng, (2, n))).astype(np.float32)
    before = _kernels.launches["median_frames"]
    _, p = msig.welch(x, nperseg=nperseg, average="median")
    _, c = msig.csd(x, y, nperseg=nperseg, average="median")
    assert _kernels.launches["median_frames"] == before + 2
    assert_peak_close(p, oracle.welch(x, nperseg=nperseg, average="median")[1], 1e-5)
    assert_peak_close(c, oracle.csd(x, y, nperseg=nperseg, average="median")[1], 1e-5)


def test_median_axis_and_two_sided(rng):
    x = make_sig(rng, (3000, 2)).astype(np.float32)
    _, p = msig.welch(x, nperseg=128, average="median", axis=0)
    assert_peak_close(p, oracle.welch(x, nperseg=128, average="median", axis=0)[1], 1e-5)
    _, p2 = msig.welch(x[:, 0], nperseg=128, average="median", return_onesided=False)
    assert_peak_close(p2, oracle.welch(x[:, 0], nperseg=128, average="median",
                                       return_onesided=False)[1], 1e-5)


@pytest.mark.parametrize("nseg", [2047, 2048, 4095, 4096])
def test_median_kernel_adversarial_columns(rng, nseg):
    """Columns that stress the radix select: exponential power values, one
    huge outlier over the rest (the range-relative digits then squeeze the
    other keys into one bucket), an outlier over a narrow cluster, a column
    split into two far groups (for even counts the two middle values lie in
    different buckets), ties and a constant column.  All must equal
    numpy.median exactly."""
    rows, bins = 2, 12
    a = rng.exponential(size=(rows, nseg, bins)).astype(np.float32)
    a[:, 0, 1] = 1e30                                      # outlier: fallback
    a[:, :, 2] = 1.0 + rng.random((rows, nseg)).astype(np.float32) * 1e-3
    a[:, 0, 2] = 1e30                                      # outlier over a narrow cluster
    half = nseg // 2
    lo = 10.0 ** rng.uniform(-10, 0, (rows, half))
    hi = 10.0 ** rng.uniform(10, 20, (rows, nseg - half))
    a[:, :, 3] = np.concatenate([lo, hi], axis=1).astype(np.float32)  # split column
    a[:, :, 4] = np.round(rng.exponential(size=(rows, nseg)) * 4)    # ties
    a[:, :, 5] = 7.0                                                 # constant
    got = _kernels.median_frames(torch.from_numpy(a).cuda(), 1.0, False, 0).cpu().numpy()
    np.testing.assert_array_equal(got, np.median(a, axis=1).astype(np.float32))


@pytest.mark.parametrize("nseg", [513, 700, 1023, 1024])
def test_median_kernel_many_columns_mid_length(rng, nseg):
    """513-1024 segments over at least 8192 columns: the call plans 256-thread
    blocks of eight columns instead of 512-thread blocks of sixteen.  Exact
    against numpy.median, including ties, signed zeros and extreme exponents,
    and equal to the same columns run as a small call (the other shape)."""
    rows, bins = 4, 2100
    a = rng.exponential(size=(rows, nseg, bins)).astype(np.float32)
    a[:, :, 5] = np.round(a[:, :, 5])
    a[:, :, 6] = 0.0
    a[:, ::2, 7] = -0.0
    a[:, :, 8] = a[:, :, 8] * 1e30
    a[:, :, 9] = a[:, :, 9] * 1e-30
    a[:, 0, 10] = 1e30
    got = _kernels.median_frames(torch.from_numpy(a).cuda(), 1.0, False, 0).cpu().numpy()
    np.testing.assert_array_equal(got, np.median(a, axis=1).astype(np.float32))
    small = _kernels.median_frames(torch.from_numpy(np.ascontiguousarray(a[:1, :, :64])).cuda(), 1.0,
                                   False, 0).cpu().numpy()
    np.testing.assert_array_equal(small, got[:1, :64])
