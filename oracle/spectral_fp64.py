"""FP64 NumPy restatement of the reference's STFT/Welch path.

TEST INFRASTRUCTURE ONLY.  This module is the parity oracle: it may be
imported by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
leg, and never by the product package (mlx-signal_amd/), which must run its
HIP kernels or fail.

Every function restates the reference's algorithm (tabulai/mlx-signal,
src/mlx_signal_processing/spectral.py, which in turn mirrors
scipy.signal._spectral_py) in float64/complex128 arithmetic, with the
reference line it follows cited in its docstring.  Pinning (tests/test_oracle.py):

  * planner integers, freqs, t, scale: closed forms + scipy.signal in FP64;
  * windows: closed-form periodic cosine sums vs scipy.signal.get_window;
  * values of welch/csd/coherence/spectrogram/stft/periodogram: scipy.signal
    called on float64 inputs (the reference's own CPU route, spectral.py:433-450)
    over the reference tests' case lists, plus the committed golden fixtures;
  * the transform itself: an O(N^2) DFT on tiny inputs, a bin-centred tone's
    closed-form peak, Parseval, coherence(x, x) == 1, csd(x, x) == welch(x);
  * istft: scipy.signal.istft on complex128 input over the reference tests'
    cases, and the stft -> istft round trip returning the signal.
"""

from __future__ import annotations

import warnings

import numpy as np

__all__ = [
    "coherence",
    "csd",
    "dft",
    "get_window",
    "istft",
    "median_bias",
    "onesided_factor",
    "periodogram",
    "plan",
    "scipy_route",
    "spectral_helper",
    "spectrogram",
    "stft",
    "triage_segments",
    "welch",
]

# ---------------------------------------------------------------------------
# windows (reference: windows.py:19-23 -> scipy.signal.get_window, periodic)
# ---------------------------------------------------------------------------

_COSINE_SUMS = {
    "hann": (0.5, 0.5),
    "hanning": (0.5, 0.5),
    "hamming": (0.54, 0.46),
    "blackman": (0.42, 0.50, 0.08),
    "boxcar": (1.0,),
    "rect": (1.0,),
    "rectangular": (1.0,),
    "ones": (1.0,),
}


def _general_cosine_periodic(a, m: int) -> np.ndarray:
    """Periodic (fftbins=True) generalized cosine window of length m:
    w[n] = sum_k (-1)^k a_k cos(2 pi k n / m).  Closed form of
    scipy.signal.windows.general_cosine with sym=False (an m+1 point symmetric
    window truncated to m points)."""
    if m == 1:
        return np.ones(1)
    n = np.arange(m, dtype=np.float64)
    w = np.zeros(m)
    for k, ak in enumerate(a):
        w += (-1) ** k * ak * np.cos(2.0 * np.pi * k * n / m)
    return w


def _tukey_periodic(m: int, alpha: float) -> np.ndarray:
    """Periodic Tukey window: the m+1 point symmetric Tukey truncated to m."""
    if m == 1:
        return np.ones(1)
    M = m + 1
    if alpha <= 0:
        return np.ones(m)
    if alpha >= 1.0:
        return _general_cosine_periodic((0.5, 0.5), m)
    n = np.arange(M, dtype=np.float64)
    width = int(np.floor(alpha * (M - 1) / 2.0))
    n1 = n[0 : width + 1]
    n2 = n[width + 1 : M - width - 1]
    n3 = n[M - width - 1 :]
    w1 = 0.5 * (1 + np.cos(np.pi * (-1 + 2.0 * n1 / alpha / (M - 1))))
    w2 = np.ones(n2.shape)
    w3 = 0.5 * (1 + np.cos(np.pi * (-2.0 / alpha + 1 + 2.0 * n3 / alpha / (M - 1))))
    return np.concatenate((w1, w2, w3))[:m]


def get_window(window, nperseg: int) -> np.ndarray:
    """FP64 periodic window (reference windows.py:19-23).  Cosine sums and
    Tukey are restated in closed form; any other spec is delegated to
    scipy.signal.get_window, the reference's own source of window values."""
    name, param = (window, None)
    if isinstance(window, tuple):
        name = window[0]
        param = window[1] if len(window) > 1 else None
    if isinstance(name, str):
        key = name.lower()
        if key in _COSINE_SUMS and param is None:
            return _general_cosine_periodic(_COSINE_SUMS[key], nperseg)
        if key in ("tukey", "tuk") and param is not None:
            return _tukey_periodic(nperseg, float(param))
    import scipy.signal as sps

    return np.asarray(sps.get_window(window, nperseg, fftbins=True), dtype=np.float64)


# ---------------------------------------------------------------------------
# planner pieces
# ---------------------------------------------------------------------------


def triage_segments(window, nperseg, input_length):
    """(window values FP64, nperseg) exactly as spectral.py:45-69."""
    if isinstance(window, (str, tuple)):
        nperseg = 256 if nperseg is None else int(nperseg)
        if nperseg > input_length:
            warnings.warn(
                f"nperseg = {nperseg:d} is greater than input length "
                f" = {input_length:d}, using nperseg = {input_length:d}",
                stacklevel=3,
            )
            nperseg = input_length
        win = get_window(window, nperseg)
    else:
        win = np.asarray(window, dtype=np.float64)
        if win.ndim != 1:
            raise ValueError("window must be 1-D")
        if input_length < win.shape[0]:
            raise ValueError("window is longer than input signal")
        if nperseg is None:
            nperseg = win.shape[0]
        elif int(nperseg) != win.shape[0]:
            raise ValueError("value specified for nperseg is different from length of window")
        nperseg = int(nperseg)
    return win, nperseg


def onesided_factor(nfft: int) -> np.ndarray:
    """[1, 2, ..., 2, (1 if nfft even else 2)] (spectral.py:150-157)."""
    f = np.full(nfft // 2 + 1, 2.0)
    f[0] = 1.0
    if nfft % 2 == 0:
        f[-1] = 1.0
    return f


def median_bias(n: int) -> float:
    """1 + sum_{i=1}^{(n-1)//2} (1/(2i+1) - 1/(2i)) (spectral.py:160-162)."""
    ii_2 = 2 * np.arange(1.0, (n - 1) // 2 + 1)
    return float(1 + np.sum(1.0 / (ii_2 + 1) - 1.0 / ii_2))


def plan(n, window="hann", nperseg=None, noverlap=None, nfft=None, fs=1.0,
         scaling="density", return_onesided=True, is_complex=False,
         boundary=None, padded=False):
    """Integer segment plan + FP64 constants of spectral.py:247-357.

    Returns a dict with nperseg, noverlap, nstep, nfft, n (after boundary
    extension and padding), nadd, nseg, sides, bins, freqs, t, scale, win."""
    if nperseg is not None:
        nperseg = int(nperseg)
        if nperseg < 1:
            raise ValueError("nperseg must be a positive integer")
    win, nperseg = triage_segments(window, nperseg, input_length=n)
    nfft = nperseg if nfft is None else int(nfft)
    if nfft < nperseg:
        raise ValueError("nfft must be greater than or equal to nperseg.")
    noverlap = nperseg // 2 if noverlap is None else int(noverlap)
    if noverlap >= nperseg:
        raise ValueError("noverlap must be less than nperseg.")
    nstep = nperseg - noverlap
    if boundary is not None:
        n = n + 2 * (nperseg // 2)
    nadd = 0
    if padded:
        nadd = (-(n - nperseg) % nstep) % nperseg
        n = n + nadd
    sides = "twosided" if (is_complex or not return_onesided) else "onesided"
    freqs = np.fft.rfftfreq(nfft, 1 / fs) if sides == "onesided" else np.fft.fftfreq(nfft, 1 / fs)
    if scaling == "density":
        scale = 1.0 / (fs * float((win * win).sum()))
    elif scaling == "spectrum":
        scale = 1.0 / float(win.sum()) ** 2
    else:
        raise ValueError(f"Unknown scaling: {scaling!r}")
    nseg = (n - nperseg) // nstep + 1
    t = np.arange(nperseg / 2, n - nperseg / 2 + 1, nstep) / float(fs)
    if boundary is not None:
        t = t - (nperseg / 2) / fs
    return dict(nperseg=nperseg, noverlap=noverlap, nstep=nstep, nfft=nfft, n=n, nadd=nadd,
                nseg=nseg, sides=sides, bins=len(freqs), freqs=freqs, t=t, scale=scale, win=win)


# ---------------------------------------------------------------------------
# transform core
# ---------------------------------------------------------------------------


def dft(x, n=None):
    """O(N^2) FP64 DFT along the last axis (for pinning tiny cases)."""
    x = np.asarray(x, dtype=np.complex128)
    N = x.shape[-1] if n is None else n
    if N > x.shape[-1]:
        x = np.concatenate([x, np.zeros(x.shape[:-1] + (N - x.shape[-1],))], axis=-1)
    k = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(k, k) / N)
    return x[..., :N] @ W.T


def _detrend(d, kind):
    """Per-segment detrend (spectral.py:89-104): 'constant' removes the mean,
    'linear' the least-squares line (closed form, denom = n(n^2-1)/12)."""
    if kind is False or kind is None:
        return d
    if callable(kind):
        return kind(d)
    if not isinstance(kind, str):
        raise ValueError("Trend type must be 'linear' or 'constant'.")
    if kind in ("constant", "c"):
        return d - d.mean(axis=-1, keepdims=True)
    if kind in ("linear", "l"):
        n = d.shape[-1]
        denom = n * (n * n - 1) / 12.0
        if denom == 0.0:
            return d - d.mean(axis=-1, keepdims=True)
        t = np.arange(n, dtype=np.float64) - (n - 1) / 2.0
        m = d.mean(axis=-1, keepdims=True)
        slope = (d * (t / denom)).sum(axis=-1, keepdims=True)
        return d - m - slope * t
    raise ValueError("Trend type must be 'linear' or 'constant'.")


def _frames(x, nperseg, nstep):
    """Overlapping segments of the last axis (spectral.py:107-124)."""
    nseg = (x.shape[-1] - nperseg) // nstep + 1
    idx = np.arange(nperseg)[None, :] + nstep * np.arange(nseg)[:, None]
    return x[..., idx]


def _zero_ext(x, n):
    z = np.zeros(x.shape[:-1] + (n,), dtype=x.dtype)
    return np.concatenate([z, x, z], axis=-1)


def _even_ext(x, n):
    return np.concatenate([x[..., n:0:-1], x, x[..., -2 : -(n + 2) : -1]], axis=-1)


def _odd_ext(x, n):
    left = 2 * x[..., :1] - x[..., n:0:-1]
    right = 2 * x[..., -1:] - x[..., -2 : -(n + 2) : -1]
    return np.concatenate([left, x, right], axis=-1)


def _const_ext(x, n):
    left = np.repeat(x[..., :1], n, axis=-1)
    right = np.repeat(x[..., -1:], n, axis=-1)
    return np.concatenate([left, x, right], axis=-1)


_EXT = {"zeros": _zero_ext, "even": _even_ext, "odd": _odd_ext, "constant": _const_ext}


def _as64(x):
    x = np.asarray(x)
    return x.astype(np.complex128) if np.iscomplexobj(x) else x.astype(np.float64)


def spectral_helper(x, y=None, fs=1.0, window="hann", nperseg=None, noverlap=None, nfft=None,
                    detrend="constant", return_onesided=True, scaling="density", axis=-1,
                    mode="psd", boundary=None, padded=False):
    """FP64 restatement of _spectral_helper (spectral.py:186-368).

    Returns (freqs, t, result) with result[..., bins, nseg] (frequency axis at
    `axis`, segment axis last), onesided doubling applied for psd mode as
    scipy does.  y=None means the same-signal path."""
    if mode not in ("psd", "stft"):
        raise ValueError(f"Unknown value for mode {mode}")
    if boundary not in (None, "zeros", "even", "odd", "constant"):
        raise ValueError(f"Unknown boundary option '{boundary}'")
    same = y is None
    x = _as64(x)
    if not same:
        y = _as64(y)
        xo, yo = list(x.shape), list(y.shape)
        xo.pop(axis)
        yo.pop(axis)
        outer = np.broadcast_shapes(tuple(xo), tuple(yo))
    if same and x.size == 0:
        e = np.zeros(x.shape)
        return e, e, e
    if not same and (x.size == 0 or y.size == 0):
        e = np.moveaxis(np.zeros(outer + (min(x.shape[axis], y.shape[axis]),)), -1, axis)
        return e, e, e
    if x.ndim > 1 and axis != -1:
        x = np.moveaxis(x, axis, -1)
        if not same and y.ndim > 1:
            y = np.moveaxis(y, axis, -1)
    if not same and x.shape[-1] != y.shape[-1]:
        if x.shape[-1] < y.shape[-1]:
            x = np.concatenate([x, np.zeros(x.shape[:-1] + (y.shape[-1] - x.shape[-1],))], -1)
        else:
            y = np.concatenate([y, np.zeros(y.shape[:-1] + (x.shape[-1] - y.shape[-1],))], -1)
    is_complex = np.iscomplexobj(x) or (not same and np.iscomplexobj(y))
    p = plan(x.shape[-1], window, nperseg, noverlap, nfft, fs, scaling, return_onesided,
             is_complex, boundary=None, padded=False)
    nperseg, nstep, nfft, win = p["nperseg"], p["nstep"], p["nfft"], p["win"]
    if boundary is not None:
        x = _EXT[boundary](x, nperseg // 2)
        if not same:
            y = _EXT[boundary](y, nperseg // 2)
    if padded:
        nadd = (-(x.shape[-1] - nperseg) % nstep) % nperseg
        x = np.concatenate([x, np.zeros(x.shape[:-1] + (nadd,), dtype=x.dtype)], -1)
        if not same:
            y = np.concatenate([y, np.zeros(y.shape[:-1] + (nadd,), dtype=y.dtype)], -1)
    sides = p["sides"]
    freqs = p["freqs"]
    scale = p["scale"]
    if mode == "stft":
        scale = np.sqrt(scale)

    def fft_frames(s):
        f = _detrend(_frames(s, nperseg, nstep), detrend) * win
        if sides == "onesided":
            return np.fft.rfft(f.real, n=nfft, axis=-1)
        return np.fft.fft(f, n=nfft, axis=-1)

    fx = fft_frames(x)
    if not same:
        res = np.conj(fx) * fft_frames(y)
    elif mode == "psd":
        res = np.conj(fx) * fx
    else:
        res = fx
    res = res * scale
    if sides == "onesided" and mode == "psd":
        res = res * onesided_factor(nfft)
    t = np.arange(nperseg / 2, x.shape[-1] - nperseg / 2 + 1, nstep) / float(fs)
    if boundary is not None:
        t = t - (nperseg / 2) / fs
    if same and mode != "stft":
        res = res.real
    ax = axis - 1 if axis < 0 else axis
    res = np.moveaxis(res, -1, ax)
    return freqs, t, res


def csd(x, y, fs=1.0, window="hann", nperseg=None, noverlap=None, nfft=None,
        detrend="constant", return_onesided=True, scaling="density", axis=-1, average="mean"):
    """spectral.py:398-489 in FP64 (y=None: same-signal, i.e. welch)."""
    if average not in ("mean", "median"):
        raise ValueError(f"average must be 'mean' or 'median', got {average!r}")
    freqs, _, pxy = spectral_helper(x, y, fs, window, nperseg, noverlap, nfft, detrend,
                                    return_onesided, scaling, axis, mode="psd")
    if pxy.ndim >= 1 and pxy.size and pxy.shape[-1] > 1:
        if average == "median":
            bias = median_bias(pxy.shape[-1])
            if np.iscomplexobj(pxy):
                pxy = np.median(pxy.real, axis=-1) + 1j * np.median(pxy.imag, axis=-1)
            else:
                pxy = np.median(pxy, axis=-1)
            pxy = pxy / bias
        else:
            pxy = pxy.mean(axis=-1)
    elif pxy.size:
        pxy = pxy.reshape(pxy.shape[:-1])
    return freqs, pxy


def welch(x, fs=1.0, window="hann", nperseg=None, noverlap=None, nfft=None,
          detrend="constant", return_onesided=True, scaling="density", axis=-1, average="mean"):
    """spectral.py:492-526: csd(x, x) then the real part."""
    f, p = csd(x, None, fs, window, nperseg, noverlap, nfft, detrend, return_onesided, scaling,
               axis, average)
    return f, np.real(p)


def periodogram(x, fs=1.0, window="boxcar", nfft=None, detrend="constant",
                return_onesided=True, scaling="density", axis=-1):
    """spectral.py:529-578: welch with nperseg = n (or nfft), noverlap = 0."""
    x = _as64(x)
    if x.size == 0:
        return np.zeros(x.shape), np.zeros(x.shape)
    if window is None:
        window = "boxcar"
    n_ax = x.shape[axis]
    if nfft is None:
        nperseg = n_ax
    elif nfft == n_ax:
        nperseg = nfft
    elif nfft > n_ax:
        nperseg = n_ax
    else:
        x = np.take(x, np.arange(nfft), axis=axis)
        nperseg = nfft
        nfft = None
    return welch(x, fs, window, nperseg, 0, nfft, detrend, return_onesided, scaling, axis)


def coherence(x, y, fs=1.0, window="hann", nperseg=None, noverlap=None, nfft=None,
              detrend="constant", axis=-1):
    """spectral.py:581-632: |Pxy|^2 / Pxx / Pyy."""
    f, pxx = welch(x, fs, window, nperseg, noverlap, nfft, detrend, axis=axis)
    _, pyy = welch(y, fs, window, nperseg, noverlap, nfft, detrend, axis=axis)
    _, pxy = csd(x, y, fs, window, nperseg, noverlap, nfft, detrend, axis=axis)
    return f, np.abs(pxy) ** 2 / pxx / pyy


def spectrogram(x, fs=1.0, window=("tukey", 0.25), nperseg=None, noverlap=None, nfft=None,
                detrend="constant", return_onesided=True, scaling="density", axis=-1,
                mode="psd"):
    """spectral.py:635-691 (tukey(0.25) default, noverlap = nperseg // 8)."""
    x = _as64(x)
    window, nperseg = triage_segments(window, nperseg, input_length=x.shape[axis])
    if noverlap is None:
        noverlap = nperseg // 8
    hm = "psd" if mode == "psd" else "stft"
    f, t, s = spectral_helper(x, None, fs, window, nperseg, noverlap, nfft, detrend,
                              return_onesided, scaling, axis, mode=hm)
    if mode == "magnitude":
        s = np.abs(s)
    elif mode in ("angle", "phase"):
        s = np.angle(s)
        if mode == "phase":
            s = np.unwrap(s, axis=(axis - 1 if axis < 0 else axis))
    return f, t, s


def stft(x, fs=1.0, window="hann", nperseg=256, noverlap=None, nfft=None, detrend=False,
         return_onesided=True, boundary="zeros", padded=True, axis=-1, scaling="spectrum"):
    """spectral.py:694-738 (legacy scipy.signal.stft semantics)."""
    if scaling == "psd":
        scaling = "density"
    elif scaling != "spectrum":
        raise ValueError(f"Parameter {scaling=} not in ['spectrum', 'psd']!")
    return spectral_helper(x, None, fs, window, nperseg, noverlap, nfft, detrend,
                           return_onesided, scaling, axis, mode="stft", boundary=boundary,
                           padded=padded)


def _istft_axes(ndim: int, freq_axis: int, time_axis: int):
    """spectral.py:766-785: normalized axes and the permutation to (..., time, freq)."""
    if ndim < 2:
        raise ValueError("Input stft must be at least 2d!")
    if freq_axis == time_axis:
        raise ValueError("Must specify differing time and frequency axes!")
    freq_ax = freq_axis + ndim if freq_axis < 0 else freq_axis
    time_ax = time_axis + ndim if time_axis < 0 else time_axis
    if not (0 <= freq_ax < ndim and 0 <= time_ax < ndim):
        raise IndexError("axis out of range")
    if freq_ax == time_ax:
        raise ValueError("Must specify differing time and frequency axes!")
    outer = [i for i in range(ndim) if i not in (freq_ax, time_ax)]
    return freq_ax, time_ax, outer + [time_ax, freq_ax]


def istft_plan(nf, input_onesided=True, nperseg=None, noverlap=None, nfft=None):
    """Integer plan of istft (spectral.py:888-912): nperseg, nfft, nstep."""
    n_default = 2 * (nf - 1) if input_onesided else nf
    if nperseg is None:
        nperseg = n_default
    else:
        nperseg = int(nperseg)
        if nperseg < 1:
            raise ValueError("nperseg must be a positive integer")
    if nfft is None:
        nfft = nperseg if (input_onesided and nperseg == n_default + 1) else n_default
    elif nfft < nperseg:
        raise ValueError("nfft must be greater than or equal to nperseg.")
    else:
        nfft = int(nfft)
    noverlap = nperseg // 2 if noverlap is None else int(noverlap)
    if noverlap >= nperseg:
        raise ValueError("noverlap must be less than nperseg.")
    return nperseg, nfft, nperseg - noverlap


def istft(Zxx, fs=1.0, window="hann", nperseg=None, noverlap=None, nfft=None,
          input_onesided=True, boundary=True, time_axis=-1, freq_axis=-2, scaling="spectrum"):
    """spectral.py:824-992 in complex128: inverse FFT of every frame, times
    scale and the window, overlap-add, divide by the window-power sum where it
    exceeds 1e-10, boundary trim.  Returns (time, x) with the time vector the
    length of the output's time axis (the reference's documented deviation
    from scipy for N-d input)."""
    Z = np.asarray(Zxx).astype(np.complex128)
    freq_ax, time_ax, perm = _istft_axes(Z.ndim, int(freq_axis), int(time_axis))
    Z = np.transpose(Z, perm)
    nseg, nf = Z.shape[-2], Z.shape[-1]
    nperseg, nfft, nstep = istft_plan(nf, input_onesided, nperseg, noverlap, nfft)
    if isinstance(window, (str, tuple)):
        win = get_window(window, nperseg)
    else:
        win = np.asarray(window, dtype=np.float64)
        if win.ndim != 1:
            raise ValueError("window must be 1-D")
        if win.shape[0] != nperseg:
            raise ValueError(f"window must have length of {nperseg}")
    if scaling == "spectrum":
        scale = win.sum()
    elif scaling == "psd":
        scale = np.sqrt(fs * np.sum(win ** 2))
    else:
        raise ValueError(f"Parameter {scaling=} not in ['spectrum', 'psd']!")
    inv = np.fft.irfft(Z, n=nfft, axis=-1) if input_onesided else np.fft.ifft(Z, n=nfft, axis=-1)
    xsubs = inv[..., :nperseg] * scale * win
    outputlength = nperseg + (nseg - 1) * nstep
    x = np.zeros(Z.shape[:-2] + (outputlength,), dtype=xsubs.dtype)
    norm = np.zeros(outputlength)
    for s in range(nseg):
        x[..., s * nstep:s * nstep + nperseg] += xsubs[..., s, :]
        norm[s * nstep:s * nstep + nperseg] += win ** 2
    if boundary:
        x = x[..., nperseg // 2:-(nperseg // 2)]
        norm = norm[nperseg // 2:-(nperseg // 2)]
    if np.sum(norm > 1e-10) != len(norm):
        warnings.warn("NOLA condition failed, STFT may not be invertible."
                      + (" Possibly due to missing boundary" if not boundary else ""),
                      stacklevel=2)
    x = x / np.where(norm > 1e-10, norm, 1.0)
    if input_onesided:
        x = x.real
    out_time_axis = time_ax - (1 if freq_ax < time_ax else 0)
    if x.ndim > 1 and out_time_axis != x.ndim - 1:
        x = np.moveaxis(x, -1, out_time_axis)
    n_out = x.shape[out_time_axis if x.ndim > 1 else -1]
    return np.arange(n_out) / float(fs), x


def scipy_route(func: str, *args, **kwargs):
    """The reference's dispatch='scipy' CPU route (spectral.py:433-450):
    scipy.signal.<func> on the caller's arrays as given (float32 stays
    float32, so SciPy computes in complex64), results cast to float32 /
    complex64 (_array.py:108-123).  Timed as the CPU baseline."""
    import scipy.signal as sps

    out = getattr(sps, func)(*args, **kwargs)
    res = []
    for a in out:
        a = np.asarray(a)
        if a.dtype.kind == "f" and a.dtype.itemsize > 4:
            a = a.astype(np.float32)
        elif a.dtype.kind == "c" and a.dtype.itemsize > 8:
            a = a.astype(np.complex64)
        res.append(a)
    return tuple(res)
