"""Host-side contract of the drop-in API (CPU only): configuration, dispatch
routing, dtype policy, argument validation and empty-input shapes, following
the reference suite (tests/test_dispatch_dtypes.py, tests/test_edge_cases.py)."""

import warnings

import numpy as np
import pytest
import scipy.signal as sps
import torch

import mlx_signal_amd as msig
from _utils import assert_close, make_sig
from mlx_signal_amd import _array, _device, _plan, spectral


def test_set_config_is_atomic():
    before = msig.get_config()
    with pytest.raises(ValueError):
        msig.set_config(dispatch="scipy", float64="never")
    assert msig.get_config() == before
    with pytest.raises(ValueError):
        msig.set_config(gpu_min_size=-1)
    assert msig.get_config() == before
    with pytest.raises(TypeError):
        msig.set_config(nonsense=1)
    with pytest.raises(ValueError):
        msig.set_config(dispatch="cuda")


def test_config_context_restores():
    base = msig.get_config().dispatch
    with msig.config_context(dispatch="scipy"):
        assert msig.get_config().dispatch == "scipy"
    assert msig.get_config().dispatch == base


def test_mlx_dispatch_alias_accepted():
    with msig.config_context(dispatch="mlx"):
        assert msig.get_config().dispatch == "mlx"


def test_scipy_route_returns_float32_tensor(rng):
    x = rng.standard_normal(4096).astype(np.float32)
    with msig.config_context(dispatch="scipy"):
        f, p = msig.welch(x, nperseg=256)
    assert isinstance(p, torch.Tensor) and p.dtype == torch.float32
    assert_close(p, sps.welch(x, nperseg=256)[1])


def test_auto_small_input_routes_to_scipy(rng):
    x = rng.standard_normal(512).astype(np.float32)
    with msig.config_context(dispatch="auto", gpu_min_size=1 << 20):
        f, p = msig.welch(x, nperseg=128)
    assert isinstance(p, torch.Tensor) and p.dtype == torch.float32
    assert_close(p, sps.welch(x, nperseg=128)[1])


@pytest.mark.skipif(_device.gpu_available(), reason="checks the no-GPU behaviour")
def test_gpu_dispatch_without_gpu_raises(rng):
    x = rng.standard_normal(4096).astype(np.float32)
    with msig.config_context(dispatch="gpu"):
        with pytest.raises(RuntimeError, match="GPU"):
            msig.welch(x, nperseg=256)


def test_callable_detrend_falls_back_with_warning(rng):
    x = rng.standard_normal(1 << 16).astype(np.float32)
    with msig.config_context(dispatch="auto", gpu_min_size=0):
        if _device.gpu_available():
            with pytest.warns(msig.FallbackWarning):
                msig.welch(x, nperseg=256, detrend=lambda d: d - d.mean(-1, keepdims=True))
        else:
            msig.welch(x, nperseg=256, detrend=lambda d: d)


def test_callable_detrend_under_gpu_dispatch_raises(rng):
    x = rng.standard_normal(1 << 16).astype(np.float32)
    with msig.config_context(dispatch="gpu"):
        with pytest.raises(NotImplementedError):
            msig.welch(x, nperseg=256, detrend=lambda d: d)


def test_float64_strict_raises_on_scipy_path(rng):
    x = rng.standard_normal(100)
    with msig.config_context(dispatch="scipy", float64="strict"):
        with pytest.raises(TypeError, match="strict"):
            msig.welch(x, nperseg=32)


def test_dtype_policy_to_device():
    dev = torch.device("cpu")
    with msig.config_context(float64="downcast", warn_on_downcast=True):
        with pytest.warns(msig.DowncastWarning):
            t = _array.to_device(np.zeros(8), dev)
        assert t.dtype == torch.float32
        with pytest.warns(msig.DowncastWarning):
            t = _array.to_device(np.zeros(8, np.complex128), dev)
        assert t.dtype == torch.complex64
    with warnings.catch_warnings():
        warnings.simplefilter("error", msig.DowncastWarning)
        assert _array.to_device([1.0, 2.0], dev).dtype == torch.float32
        assert _array.to_device(np.arange(5, dtype=np.int32), dev).dtype == torch.float32
    with msig.config_context(float64="strict"):
        with pytest.raises(TypeError, match="strict"):
            _array.to_device(np.zeros(4), dev)
        with pytest.raises(TypeError):
            _array.check_strict(torch.zeros(3, dtype=torch.float64))


@pytest.mark.parametrize("func", ["welch", "csd"])
def test_empty_input_shapes(func, rng):
    x = np.array([], dtype=np.float32)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        if func == "welch":
            f_ref, p_ref = sps.welch(x)
            f, p = msig.welch(x)
        else:
            y = rng.standard_normal(0).astype(np.float32)
            f_ref, p_ref = sps.csd(x, y)
            f, p = msig.csd(x, y)
    assert tuple(f.shape) == f_ref.shape
    assert tuple(p.shape) == p_ref.shape


def test_argument_errors():
    x = np.zeros(4096, np.float32)
    with pytest.raises(ValueError, match="average"):
        msig.csd(x, x, average="mode")
    with pytest.raises(ValueError, match="scaling"):
        msig.stft(x, scaling="energy")
    with pytest.raises(ValueError, match="mode"):
        msig.spectrogram(x, mode="power")
    with pytest.raises(ValueError, match="mode"):
        spectral._estimate(x, x, mode="bad")
    with pytest.raises(ValueError, match="boundary"):
        spectral._estimate(x, x, boundary="wrap")
    with pytest.raises(ValueError, match="equal"):
        spectral._estimate(x, np.zeros(4096, np.float32), mode="stft")


def test_triage_segments_matches_reference_rules():
    w, n = _plan.resolve_window("hann", None, 1000)
    assert n == 256 and w.shape == (256,)
    with pytest.warns(UserWarning, match="nperseg"):
        w, n = _plan.resolve_window("hann", 512, 100)
    assert n == 100
    with pytest.raises(ValueError, match="1-D"):
        _plan.resolve_window(np.ones((2, 2)), None, 100)
    with pytest.raises(ValueError, match="longer"):
        _plan.resolve_window(np.ones(200), None, 100)
    with pytest.raises(ValueError, match="different"):
        _plan.resolve_window(np.ones(64), 32, 100)
    with pytest.raises(ValueError, match="positive"):
        _plan.check_nperseg(0)
    with pytest.raises(ValueError, match="nfft"):
        _plan.resolve_lengths(64, None, 32)
    with pytest.raises(ValueError, match="noverlap"):
        _plan.resolve_lengths(64, 64, None)
    assert _plan.resolve_lengths(64, None, None) == (32, 32, 64)
    np.testing.assert_array_equal(_plan.onesided_factor(8), [1, 2, 2, 2, 1])
    np.testing.assert_array_equal(_plan.onesided_factor(7), [1, 2, 2, 2])
    assert abs(_plan.median_bias(3) - 5 / 6) < 1e-15
    from scipy.signal._spectral_py import _median_bias as scipy_bias

    for n in (1, 2, 3, 4, 9, 10, 101, 4096):
        assert abs(_plan.median_bias(n) - scipy_bias(n)) < 1e-14, n
    for n, nps, step in ((1000, 256, 128), (1000, 100, 37), (256, 256, 1), (999, 64, 64)):
        nseg = _plan.segment_count(n, nps, step)
        assert nseg == len(range(0, n - nps + 1, step))
        pad = _plan.padding_for(n, nps, step)
        assert (n + pad - nps) % step == 0 and 0 <= pad < nps


def test_median_last_matches_numpy(rng):
    a = rng.standard_normal((3, 5, 8)).astype(np.float32)
    for n in (7, 8):
        t = torch.from_numpy(a[..., :n])
        np.testing.assert_allclose(spectral._median_over_last(t).numpy(), np.median(a[..., :n], axis=-1))


def test_unwrap_matches_numpy(rng):
    p = np.cumsum(rng.uniform(-3, 3, (4, 50)), axis=1).astype(np.float32)
    for ax in (0, 1):
        np.testing.assert_allclose(spectral._unwrap_phase(torch.from_numpy(p), dim=ax).numpy(),
                                   np.unwrap(p, axis=ax), rtol=1e-5, atol=1e-4)


def test_edge_extensions_match_scipy(rng):
    from scipy.signal._arraytools import const_ext, even_ext, odd_ext, zero_ext

    from mlx_signal_amd import _arraytools as at

    x = rng.standard_normal((2, 20)).astype(np.float32)
    t = torch.from_numpy(x)
    for ours, ref in ((at.even_ext, even_ext), (at.odd_ext, odd_ext),
                      (at.const_ext, const_ext), (at.zero_ext, zero_ext)):
        np.testing.assert_allclose(ours(t, 5).numpy(), ref(x, 5, axis=-1), rtol=1e-6)
    with pytest.raises(ValueError, match="too big"):
        at.even_ext(t, 20)


def test_composed_detrend_matches_oracle(rng):
    """The composed path's detrend/window against the oracle's FP64 formulas."""
    import oracle.spectral_fp64 as osp

    d = make_sig(rng, (3, 7, 64))
    w = np.hanning(64).astype(np.float32)
    for kind in (False, "constant", "linear"):
        got = spectral._window_segments(torch.from_numpy(d), kind, torch.from_numpy(w)).numpy()
        ref = osp._detrend(d.astype(np.float64), kind) * w
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5)
    with pytest.raises(ValueError):
        spectral._window_segments(torch.from_numpy(d), "quadratic", torch.from_numpy(w))


def test_istft_nola_helper_matches_oracle_rule():
    """The host NOLA verdict (FP64 window-power sum, cached per geometry)
    agrees with the oracle's explicit overlap-add of the window power."""
    from mlx_signal_amd.spectral import _nola_ok

    import oracle

    cases = [("hann", 256, 128, 10, True), ("hann", 256, 256, 10, True),
             ("hann", 256, 256, 10, False), ("boxcar", 64, 64, 5, False),
             ("hann", 64, 16, 1, True), ("blackman", 128, 100, 7, False)]
    for win_name, nps, nstep, nseg, boundary in cases:
        w = oracle.get_window(win_name, nps)
        outlen = nps + (nseg - 1) * nstep
        norm = np.zeros(outlen)
        for s in range(nseg):
            norm[s * nstep:s * nstep + nps] += w ** 2
        if boundary:
            norm = norm[nps // 2:-(nps // 2)]
        want = bool(np.sum(norm > 1e-10) == len(norm))
        assert _nola_ok(w, nps, nstep, nseg, boundary) == want, (win_name, nps, nstep, boundary)


def test_istft_scipy_route_on_cpu():
    """Without a GPU, auto dispatch runs scipy and returns tensors."""
    import scipy.signal as sps

    import oracle

    rng = np.random.default_rng(3)
    x = rng.standard_normal(2000).astype(np.float32)
    _, _, z = sps.stft(x, nperseg=128)
    with msig.config_context(dispatch="scipy"):
        t, xr = msig.istft(z.astype(np.complex64), nperseg=128)
    assert isinstance(xr, torch.Tensor) and isinstance(t, torch.Tensor)
    _, x_ref = oracle.istft(z, nperseg=128)
    np.testing.assert_allclose(xr.numpy(), x_ref, atol=1e-5 * np.abs(x_ref).max())


def test_window_sums_cached_only_for_read_only_windows():
    """The planner keeps (sum w, sum w^2) per named-window array (read-only,
    from the window cache) and recomputes them for a caller's own array,
    which may change between calls."""
    from mlx_signal_amd.windows import window_values

    w = window_values("hann", 256)
    assert not w.flags.writeable
    s1 = _plan.window_sums(w)
    assert s1 == (float(w.sum()), float((w * w).sum()))
    assert _plan.window_sums(w) is s1  # cached for the same read-only array
    mine = np.hanning(256)
    a = _plan.window_sums(mine)
    mine *= 2.0
    b = _plan.window_sums(mine)
    assert b[0] == pytest.approx(2 * a[0]) and b[1] == pytest.approx(4 * a[1])


def test_istft_lengths_match_oracle_plan():
    """The device istft's integer plan (nperseg, nfft, nstep) equals the
    oracle's restatement of the reference's (spectral.py:888-912) on every
    combination, errors included (same type and text)."""
    import re

    import oracle.spectral_fp64 as osp

    cases = 0
    for nf in (1, 2, 5, 9, 65, 129):
        for onesided in (True, False):
            for nperseg in (None, 0, 1, 7, 8, 16, 17, 128, 257):
                for noverlap in (None, 0, 3, 8, 200):
                    for nfft in (None, 4, 16, 256, 300):
                        args = (nf, onesided, nperseg, noverlap, nfft)
                        try:
                            want = osp.istft_plan(*args)
                        except ValueError as e:
                            with pytest.raises(ValueError, match=re.escape(str(e))):
                                _plan.istft_lengths(*args)
                            cases += 1
                            continue
                        assert _plan.istft_lengths(*args) == want, args
                        cases += 1
    assert cases == 6 * 2 * 9 * 5 * 5


def test_istft_window_scale_and_errors():
    import re

    for win in ("hann", ("tukey", 0.25), np.hanning(64)):
        n = 64
        vals, s = _plan.istft_window(win, n, "spectrum", 3.0)
        assert vals.dtype == np.float64 and vals.shape == (n,)
        np.testing.assert_allclose(s, vals.sum(), rtol=1e-15)
        _, sp = _plan.istft_window(win, n, "psd", 3.0)
        np.testing.assert_allclose(sp, np.sqrt(3.0 * np.sum(vals ** 2)), rtol=1e-14)
    with pytest.raises(ValueError, match="window must be 1-D"):
        _plan.istft_window(np.ones((2, 8)), 8, "spectrum", 1.0)
    with pytest.raises(ValueError, match="window must have length of 16"):
        _plan.istft_window(np.ones(8), 16, "spectrum", 1.0)
    with pytest.raises(ValueError, match=re.escape("Parameter scaling='x' not in ['spectrum', 'psd']!")):
        _plan.istft_window("hann", 16, "x", 1.0)
