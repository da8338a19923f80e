"""Segmentation rules and constants shared by the spectral functions.

The integer rules (default lengths, overlap, padding, segment count), the
frequency and time grids and the scale factors are scipy.signal's: results
must match scipy bin for bin, so they are restated here from its documented
behaviour, and the texts of the errors and warnings are scipy's (and the
reference's, src/mlx_signal_processing/spectral.py:45-69, 215-310).  Window
values and every scale constant are computed in FP64 on the host.
"""

from __future__ import annotations

import collections
import math
import warnings

import numpy as np
import torch

from ._array import to_numpy
from .windows import window_values

DEFAULT_NPERSEG = 256
BOUNDARY_MODES = ("even", "odd", "constant", "zeros", None)
_DETREND_NAMES = {"constant": "constant", "c": "constant", "linear": "linear", "l": "linear"}


# ---------------------------------------------------------------------------
# windows and lengths
# ---------------------------------------------------------------------------


def resolve_window(window, nperseg, n_axis: int, stacklevel: int = 4):
    """(float64 window values, nperseg) for a window argument.

    A window name or tuple is computed at nperseg (256 by default), which is
    cut to the signal length with a warning when it is longer.  An array
    window fixes nperseg to its own length; it must be 1-D, no longer than
    the signal, and agree with an explicit nperseg."""
    if isinstance(window, (str, tuple)):
        length = DEFAULT_NPERSEG if nperseg is None else int(nperseg)
        if length > n_axis:
            warnings.warn(f"nperseg = {length:d} is greater than input length  = {n_axis:d}, "
                          f"using nperseg = {n_axis:d}", stacklevel=stacklevel)
            length = n_axis
        return window_values(window, length), length
    values = np.asarray(to_numpy(window), dtype=np.float64)
    if values.ndim != 1:
        raise ValueError("window must be 1-D")
    length = values.shape[0]
    if length > n_axis:
        raise ValueError("window is longer than input signal")
    if nperseg is not None and int(nperseg) != length:
        raise ValueError("value specified for nperseg is different from length of window")
    return values, length


def check_nperseg(nperseg):
    """nperseg as an int (None stays None); it must be positive."""
    if nperseg is None:
        return None
    nperseg = int(nperseg)
    if nperseg < 1:
        raise ValueError("nperseg must be a positive integer")
    return nperseg


def resolve_lengths(nperseg: int, noverlap, nfft):
    """(noverlap, nstep, nfft): overlap defaults to half a segment and must
    be shorter than it; nfft defaults to nperseg and may not be shorter."""
    nfft = nperseg if nfft is None else int(nfft)
    if nfft < nperseg:
        raise ValueError("nfft must be greater than or equal to nperseg.")
    noverlap = nperseg // 2 if noverlap is None else int(noverlap)
    if noverlap >= nperseg:
        raise ValueError("noverlap must be less than nperseg.")
    return noverlap, nperseg - noverlap, nfft


def istft_lengths(nf: int, onesided: bool, nperseg, noverlap, nfft):
    """(nperseg, nfft, nstep) of an inverse STFT over spectra of nf bins.

    The spectra imply a segment of 2 (nf - 1) samples (one-sided) or nf
    (two-sided), which is the default nperseg.  Without an explicit nfft the
    transform length is that implied length, except that a one-sided
    nperseg one longer than it (an odd segment, nf = (nperseg + 1) / 2)
    transforms unpadded.  Overlap and nfft follow resolve_lengths."""
    implied = 2 * (nf - 1) if onesided else nf
    seg = implied if nperseg is None else check_nperseg(nperseg)
    if nfft is None:
        nfft = seg if (onesided and seg == implied + 1) else implied
        noverlap, nstep, _ = resolve_lengths(seg, noverlap, max(nfft, seg))
        return seg, nfft, nstep
    noverlap, nstep, nfft = resolve_lengths(seg, noverlap, nfft)
    return seg, nfft, nstep


def istft_window(window, nperseg: int, scaling: str, fs: float):
    """(float64 window values, synthesis scale) of an inverse STFT: a named
    window is computed at nperseg, an array one must be 1-D of that length;
    the scale undoes the forward transform's: sum w for a spectrum, sqrt(fs
    sum w^2) for a density."""
    if isinstance(window, (str, tuple)):
        values = window_values(window, nperseg)
    else:
        values = np.asarray(to_numpy(window), dtype=np.float64)
        if values.ndim != 1:
            raise ValueError("window must be 1-D")
        if values.shape[0] != nperseg:
            raise ValueError(f"window must have length of {nperseg}")
    s1, s2 = window_sums(values)
    scales = {"spectrum": s1, "psd": math.sqrt(fs * s2)}
    if scaling not in scales:
        raise ValueError(f"Parameter {scaling=} not in ['spectrum', 'psd']!")
    return values, scales[scaling]


def segment_count(n: int, nperseg: int, nstep: int) -> int:
    return (n - nperseg) // nstep + 1


def padding_for(n: int, nperseg: int, nstep: int) -> int:
    """Zeros that make the last segment end at the end of the row (padded=True)."""
    return (-(n - nperseg) % nstep) % nperseg


def detrend_kind(detrend):
    """None (no detrend), "constant" or "linear"; other strings raise."""
    if detrend is False or detrend is None:
        return None
    if isinstance(detrend, str) and detrend in _DETREND_NAMES:
        return _DETREND_NAMES[detrend]
    raise ValueError("Trend type must be 'linear' or 'constant'.")


# ---------------------------------------------------------------------------
# scale constants
# ---------------------------------------------------------------------------

_sums_cache: dict = {}


def window_sums(values: np.ndarray):
    """(sum w, sum w^2) in FP64.  Remembered for the read-only arrays of the
    window cache; a caller's own (writable) array may change between calls."""
    hit = _sums_cache.get(id(values))
    if hit is not None and hit[0] is values:
        return hit[1]
    sums = (float(values.sum()), float(np.dot(values, values)))
    if not values.flags.writeable:
        if len(_sums_cache) > 256:
            _sums_cache.clear()
        _sums_cache[id(values)] = (values, sums)
    return sums


def spectral_scale(scaling: str, fs: float, values: np.ndarray) -> float:
    """1 / (fs sum w^2) for a density, 1 / (sum w)^2 for a power spectrum."""
    s1, s2 = window_sums(values)
    if scaling == "density":
        return 1.0 / (fs * s2)
    if scaling == "spectrum":
        return 1.0 / (s1 * s1)
    raise ValueError(f"Unknown scaling: {scaling!r}")


def onesided_factor(nfft: int) -> np.ndarray:
    """Per-bin doubling of a one-sided spectrum: every bin but DC and, for an
    even nfft, the Nyquist bin carries the power of its negative twin."""
    factor = np.ones(nfft // 2 + 1, dtype=np.float32)
    factor[1:(nfft + 1) // 2] = 2.0
    return factor


def median_bias(n: int) -> float:
    """Bias of the median of n periodograms of Gaussian noise: the alternating
    harmonic sum 1 - 1/2 + 1/3 - ... up to 1/(2m+1), m = (n - 1) // 2."""
    k = np.arange(1.0, 2 * ((n - 1) // 2) + 2)
    return float(np.sum(np.where(k % 2 == 1, 1.0, -1.0) / k))


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------


def freq_grid(nfft: int, fs: float, onesided: bool) -> np.ndarray:
    return np.fft.rfftfreq(nfft, 1 / fs) if onesided else np.fft.fftfreq(nfft, 1 / fs)


def time_grid(nperseg: int, n: int, nstep: int, fs: float, extended: bool) -> np.ndarray:
    """Segment centres of a row of n samples; a boundary extension of
    nperseg // 2 samples shifts them back by that much."""
    t = np.arange(nperseg / 2, n - nperseg / 2 + 1, nstep) / float(fs)
    return t - (nperseg / 2) / fs if extended else t


# ---------------------------------------------------------------------------
# device-resident constants
# ---------------------------------------------------------------------------


class DeviceConstants:
    """Small float32 constants (windows, grids, factors) kept on the device
    per (key, device), least recently used first out.  A steady-state call
    then issues no host-to-device copy."""

    def __init__(self, capacity: int = 256):
        self._items: "collections.OrderedDict" = collections.OrderedDict()
        self._ids: set = set()
        self._capacity = capacity

    def get(self, key, build, device) -> torch.Tensor:
        k = (key, str(device))
        t = self._items.get(k)
        if t is not None:
            self._items.move_to_end(k)
            return t
        t = torch.from_numpy(np.ascontiguousarray(build(), dtype=np.float32)).to(device)
        self._items[k] = t
        self._ids.add(id(t))
        if len(self._items) > self._capacity:
            _, old = self._items.popitem(last=False)
            self._ids.discard(id(old))
        return t

    def own(self, t: torch.Tensor) -> torch.Tensor:
        """``t`` if it is not one of the cached constants, else a copy the
        caller may modify."""
        return t.clone() if id(t) in self._ids else t


constants = DeviceConstants()


def window_on(device, window, values: np.ndarray) -> torch.Tensor:
    """The float32 window on the device (cached for named windows)."""
    if isinstance(window, (str, tuple)):
        return constants.get(("window", window, len(values)), lambda: values, device)
    return torch.from_numpy(values.astype(np.float32)).to(device)


def freqs_on(device, nfft: int, fs: float, onesided: bool) -> torch.Tensor:
    return constants.get(("freqs", nfft, float(fs), onesided),
                         lambda: freq_grid(nfft, fs, onesided), device)


def grids_on(device, nfft, fs, onesided, nperseg, n, nstep, extended) -> torch.Tensor:
    """Frequency grid followed by the time grid, one tensor (one side copy)."""
    def build():
        return np.concatenate([freq_grid(nfft, fs, onesided).astype(np.float32),
                               time_grid(nperseg, n, nstep, fs, extended).astype(np.float32)])

    return constants.get(("grids", nfft, float(fs), onesided, nperseg, n, nstep, extended),
                         build, device)


def times_on(device, nperseg, n, nstep, fs, extended) -> torch.Tensor:
    return constants.get(("times", nperseg, n, nstep, float(fs), extended),
                         lambda: time_grid(nperseg, n, nstep, fs, extended), device)


def factor_on(device, nfft: int) -> torch.Tensor:
    return constants.get(("onesided", nfft), lambda: onesided_factor(nfft), device)


def unit_ramp(n: int, device) -> torch.Tensor:
    """The centred ramp t - (n-1)/2 over n samples, scaled to unit norm."""
    def build():
        t = np.arange(n, dtype=np.float64) - (n - 1) / 2.0
        return t / math.sqrt(float(np.dot(t, t)))

    return constants.get(("ramp", n), build, device)
