"""HIP-path parity: every result against the FP64 oracle (itself pinned to
scipy.signal in FP64), with the bound |err| <= 1e-5 x peak|ref| for the fused
FP32 kernels and exact shapes / freqs.  Case lists follow the reference suite
(tests/test_spectral.py, tests/test_edge_cases.py, tests/test_metal_fft_workaround.py)."""

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
import oracle
from _utils import as_np, assert_close, assert_device_tensor, assert_peak_close, make_sig, peak_rel_err
from mlx_signal_amd import _kernels

pytestmark = pytest.mark.gpu

TOL = 1e-5  # per-bin error bound relative to the FP64 peak bin

WELCH_CASES = [
    {},
    {"nperseg": 128},
    {"nperseg": 128, "noverlap": 96},
    {"nperseg": 128, "window": "hamming"},
    {"nperseg": 100, "window": ("tukey", 0.25)},
    {"nperseg": 128, "detrend": "linear"},
    {"nperseg": 128, "detrend": False},
    {"nperseg": 128, "scaling": "spectrum"},
    {"nperseg": 128, "return_onesided": False},
    {"nperseg": 128, "average": "median"},
    {"nperseg": 128, "nfft": 256},
    {"nperseg": 129, "noverlap": 37, "fs": 250.0},
]


def _fused_count():
    return sum(v for k, v in _kernels.launches.items())


@pytest.mark.parametrize("n", [1024, 999, 4093])
@pytest.mark.parametrize("case", WELCH_CASES)
def test_welch_matches_oracle(rng, n, case):
    x = make_sig(rng, (n,))
    f_ref, p_ref = oracle.welch(x, **case)
    f, p = msig.welch(x, **case)
    assert_device_tensor(f)
    assert_device_tensor(p)
    assert p.device.type == "cuda"
    np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
    assert_peak_close(p, p_ref, TOL)


def test_welch_fused_path_taken(rng):
    x = make_sig(rng, (4, 1 << 14))
    before = _kernels.launches["mean_psd"]
    msig.welch(x, nperseg=1024)
    assert _kernels.launches["mean_psd"] == before + 1


def test_welch_array_window(rng):
    x = make_sig(rng, (2048,))
    import scipy.signal as sps

    w = sps.get_window("hann", 200)
    f_ref, p_ref = oracle.welch(x, window=w)
    f, p = msig.welch(x, window=w.astype(np.float32))
    assert_peak_close(p, p_ref, TOL)


@pytest.mark.parametrize("shape,axis", [((3, 2000), -1), ((2000, 3), 0), ((2, 3, 1500), -1),
                                        ((2, 1500, 3), 1)])
def test_welch_multidim(rng, shape, axis):
    x = make_sig(rng, shape)
    f_ref, p_ref = oracle.welch(x, nperseg=128, axis=axis)
    f, p = msig.welch(x, nperseg=128, axis=axis)
    assert tuple(p.shape) == p_ref.shape
    assert_peak_close(p, p_ref, TOL)


def test_welch_complex_input_switches_twosided(rng):
    x = make_sig(rng, (2048,), complex_=True)
    f_ref, p_ref = oracle.welch(x, nperseg=256)
    f, p = msig.welch(x, nperseg=256)
    np.testing.assert_array_equal(as_np(f), f_ref.astype(np.float32))
    assert_peak_close(p, p_ref, TOL)


def test_welch_short_signal_warns(rng):
    x = make_sig(rng, (100,))
    with pytest.warns(UserWarning, match="nperseg"):
        f_ref, p_ref = oracle.welch(x)
    with pytest.warns(UserWarning, match="nperseg"):
        f, p = msig.welch(x)
    assert_peak_close(p, p_ref, TOL)


@pytest.mark.parametrize("average", ["mean", "median"])
def test_csd_matches_oracle(rng, average):
    x = make_sig(rng, (4096,))
    y = make_sig(rng, (4096,))
    f_ref, p_ref = oracle.csd(x, y, nperseg=256, average=average)
    f, p = msig.csd(x, y, nperseg=256, average=average)
    assert p.dtype == torch.complex64
    assert_peak_close(p, p_ref, TOL)


def test_csd_broadcast(rng):
    x = make_sig(rng, (4096,))
    y = make_sig(rng, (3, 4096))
    _, p_ref = oracle.csd(x, y, nperseg=256)
    _, p = msig.csd(x, y, nperseg=256)
    assert tuple(p.shape) == p_ref.shape
    assert_peak_close(p, p_ref, TOL)


def test_csd_unequal_lengths(rng):
    x = make_sig(rng, (3000,))
    y = make_sig(rng, (4096,))
    _, p_ref = oracle.csd(x, y, nperseg=256)
    _, p = msig.csd(x, y, nperseg=256)
    assert_peak_close(p, p_ref, TOL)


def test_coherence_matches_oracle(rng):
    x = make_sig(rng, (8192,))
    y = 0.5 * x + 0.5 * make_sig(rng, (8192,))
    before = _kernels.launches["mean_coherence"]
    _, c_ref = oracle.coherence(x, y, nperseg=256)
    _, c = msig.coherence(x, y, nperseg=256)
    assert _kernels.launches["mean_coherence"] == before + 1
    assert_peak_close(c, c_ref, TOL)


@pytest.mark.parametrize("nperseg", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("detrend", [False, "constant"])
def test_grouped_means_match_per_frame_and_oracle(rng, nperseg, detrend):
    """Grouped mean kernels agree with the per-frame kernels and the oracle for
    every fused size; n leaves a partial last pair and a partial last run."""
    n = nperseg + (nperseg // 2) * 130
    x = make_sig(rng, (16, n))
    y = 0.5 * x + 0.5 * make_sig(rng, (16, n))
    _, pw = msig.welch(x, nperseg=nperseg, detrend=detrend)
    assert_peak_close(pw, oracle.welch(x, nperseg=nperseg, detrend=detrend)[1], TOL)
    _, pc = msig.csd(x, y, nperseg=nperseg, detrend=detrend)
    assert_peak_close(pc, oracle.csd(x, y, nperseg=nperseg, detrend=detrend)[1], TOL)
    _, co = msig.coherence(x, y, nperseg=nperseg, detrend=detrend)
    assert_peak_close(co, oracle.coherence(x, y, nperseg=nperseg, detrend=detrend)[1], TOL)

    dev = torch.device("cuda")
    xm = torch.from_numpy(x).to(dev)
    ym = torch.from_numpy(y).to(dev)
    import scipy.signal as sps

    win = torch.from_numpy(sps.get_window("hann", nperseg).astype(np.float32)).to(dev)
    hop = nperseg // 2
    p_grp = _kernels.rfft_power_mean(xm, win, nperseg, hop, detrend)
    p_frm = _kernels.rfft_frames(xm, win, nperseg, hop, detrend, power=True).double().mean(-2)
    assert_close(p_grp, as_np(p_frm), rtol=1e-5, atol_frac=1e-6)
    c_grp = _kernels.csd_mean(xm, ym, win, nperseg, hop, detrend)
    c_frm = _kernels.csd_frames(xm, ym, win, nperseg, hop, detrend).to(torch.complex128).mean(-2)
    assert_close(c_grp, as_np(c_frm), rtol=1e-5, atol_frac=1e-6)
    pxx_g, pyy_g, pxy_g = _kernels.coherence_mean(xm, ym, win, nperseg, hop, detrend)
    pxx_f, pyy_f, pxy_f = _kernels.coherence_frames(xm, ym, win, nperseg, hop, detrend)
    assert_close(pxx_g, as_np(pxx_f.double().mean(-2)), rtol=1e-5, atol_frac=1e-6)
    assert_close(pyy_g, as_np(pyy_f.double().mean(-2)), rtol=1e-5, atol_frac=1e-6)
    assert_close(pxy_g, as_np(pxy_f.to(torch.complex128).mean(-2)), rtol=1e-5, atol_frac=1e-6)


def _group_unit_counts(rows, nseg, nfft, target_blocks=1024, max_run=32):
    """Units walked by each FFT group of a row's last block, restating the
    launch geometry (256-thread blocks, nfft / 16 threads per group, one unit
    per segment pair); used only to pick shapes whose last block is uneven."""
    gpb = 256 // (nfft // 16)
    units = (nseg + 1) // 2
    bprt = max(1, -(-target_blocks // rows))
    run = max(1, min(max_run, -(-units // (gpb * bprt))))
    bpr = -(-units // (gpb * run))
    first = (bpr - 1) * gpb * run
    return [max(0, min(run, units - (first + g * run))) for g in range(gpb)]


@pytest.mark.parametrize("rows,nseg", [(4, 3001), (4, 3002), (3, 999), (1, 4099), (2, 2049)])
@pytest.mark.parametrize("detrend", [False, "constant"])
def test_uneven_last_block_nfft2048(rng, rows, nseg, detrend):
    """At nfft 2048 a block holds two FFT groups of two waves each, which
    synchronise with block barriers; the last block of a row gives them
    different unit counts (or one group none), so the kernels pad the shorter
    group with inactive units to keep both on the same barrier sequence.  The
    grouped means must match the per-frame kernels and the FP64 oracle, and
    repeated calls must give bitwise identical results (a barrier mismatch
    showed up as intermittent corruption of one unit's spectrum)."""
    N, hop = 2048, 1024
    counts = _group_unit_counts(rows, nseg, N)
    n = N + hop * (nseg - 1) + 77
    x = make_sig(rng, (rows, n))
    y = (0.5 * x + 0.5 * make_sig(rng, (rows, n))).astype(np.float32)
    dev = torch.device("cuda")
    xm, ym = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    import scipy.signal as sps

    win = torch.from_numpy(sps.get_window("hann", N).astype(np.float32)).to(dev)
    for h in (hop, 768):  # the hop == N/2 stream and the general one
        p_grp = _kernels.rfft_power_mean(xm, win, N, h, detrend)
        p_frm = _kernels.rfft_frames(xm, win, N, h, detrend, power=True).double().mean(-2)
        assert_close(p_grp, as_np(p_frm), rtol=1e-5, atol_frac=1e-6)
        c_grp = _kernels.csd_mean(xm, ym, win, N, h, detrend)
        c_frm = _kernels.csd_frames(xm, ym, win, N, h, detrend).to(torch.complex128).mean(-2)
        assert_close(c_grp, as_np(c_frm), rtol=1e-5, atol_frac=1e-6)
        for _ in range(6):
            assert torch.equal(_kernels.rfft_power_mean(xm, win, N, h, detrend), p_grp)
            assert torch.equal(_kernels.csd_mean(xm, ym, win, N, h, detrend), c_grp)
        pxx_g, pyy_g, pxy_g = _kernels.coherence_mean(xm, ym, win, N, h, detrend)
        pxx_f, pyy_f, pxy_f = _kernels.coherence_frames(xm, ym, win, N, h, detrend)
        assert_close(pxx_g, as_np(pxx_f.double().mean(-2)), rtol=1e-5, atol_frac=1e-6)
        assert_close(pyy_g, as_np(pyy_f.double().mean(-2)), rtol=1e-5, atol_frac=1e-6)
        assert_close(pxy_g, as_np(pxy_f.to(torch.complex128).mean(-2)), rtol=1e-5, atol_frac=1e-6)
    if rows * n <= 4_000_000:
        _, pw = msig.welch(x, nperseg=N, detrend=detrend)
        assert_peak_close(pw, oracle.welch(x, nperseg=N, detrend=detrend)[1], TOL)
    if not detrend:  # istft at nfft 2048: every group walks the same units
        _, _, z = msig.stft(xm, nperseg=N)
        _, xr = msig.istft(z, nperseg=N)
        for _ in range(4):
            assert torch.equal(msig.istft(z, nperseg=N)[1], xr)
        err = float((xr[:, :n] - xm).abs().max() / xm.abs().max())
        assert err < 1e-5, err
    assert counts[0] >= 1 and len(counts) == 2


def test_uneven_last_block_shapes_are_uneven():
    """The shapes above do give the two groups of a row's last block
    different unit counts (most of them; one is a control)."""
    uneven = [(r, s) for r, s in [(4, 3001), (4, 3002), (3, 999), (1, 4099), (2, 2049)]
              if len(set(_group_unit_counts(r, s, 2048))) > 1]
    assert len(uneven) >= 3, uneven


def test_grouped_single_group_and_nonlast_axis(rng):
    x_many = make_sig(rng, (2048, 320))  # nperseg=64, hop=32 -> nseg=9
    y_many = 0.5 * x_many + 0.5 * make_sig(rng, (2048, 320))
    assert_peak_close(msig.welch(x_many, nperseg=64)[1], oracle.welch(x_many, nperseg=64)[1], TOL)
    assert_peak_close(msig.coherence(x_many, y_many, nperseg=64)[1],
                      oracle.coherence(x_many, y_many, nperseg=64)[1], TOL)
    x_ax = make_sig(rng, (2, 4096, 3))
    y_ax = 0.5 * x_ax + 0.5 * make_sig(rng, (2, 4096, 3))
    assert_peak_close(msig.welch(x_ax, nperseg=128, axis=1)[1],
                      oracle.welch(x_ax, nperseg=128, axis=1)[1], TOL)
    assert_peak_close(msig.csd(x_ax, y_ax, nperseg=128, axis=1)[1],
                      oracle.csd(x_ax, y_ax, nperseg=128, axis=1)[1], TOL)
    assert_peak_close(msig.coherence(x_ax, y_ax, nperseg=128, axis=1)[1],
                      oracle.coherence(x_ax, y_ax, nperseg=128, axis=1)[1], TOL)


@pytest.mark.parametrize("mode", ["psd", "complex", "magnitude", "angle", "phase"])
def test_spectrogram_matches_oracle(rng, mode):
    x = make_sig(rng, (4096,))
    f_ref, t_ref, s_ref = oracle.spectrogram(x, fs=100.0, mode=mode)
    f, t, s = msig.spectrogram(x, fs=100.0, mode=mode)
    assert tuple(s.shape) == s_ref.shape
    np.testing.assert_allclose(as_np(t), t_ref, rtol=1e-6)
    if mode in ("angle", "phase"):
        mag = np.abs(oracle.spectrogram(x, fs=100.0, mode="complex")[2])
        keep = mag > 1e-4 * mag.max()
        diff = np.angle(np.exp(1j * (as_np(s)[keep] - s_ref[keep])))
        np.testing.assert_allclose(diff, np.zeros_like(diff), atol=2e-3)
    else:
        assert_peak_close(s, s_ref, TOL)


def test_spectrogram_fused_power_path(rng):
    x = make_sig(rng, (8, 1 << 14))
    before = _kernels.launches["frames_power"]
    _, _, s = msig.spectrogram(x, nperseg=1024, noverlap=512, window="hann")
    assert _kernels.launches["frames_power"] == before + 1
    assert_peak_close(s, oracle.spectrogram(x, nperseg=1024, noverlap=512, window="hann")[2], TOL)


@pytest.mark.parametrize("nperseg", [32, 256, 1024])
@pytest.mark.parametrize("odd", [False, True])
@pytest.mark.parametrize("detrend", [False, "constant"])
def test_spectrogram_half_hop_power_stream(rng, nperseg, odd, detrend):
    """Power frames at hop = N/2 run the full-unit stream (no bounds selects,
    carried half-block detrend sums) for every unit with both segments and
    the general fetch for the partner-less last unit of an odd count: odd
    and even segment counts, with and without detrend, on a signal with a
    large DC offset, several rows and many units per group; two-sided output
    keeps the general loop."""
    hop = nperseg // 2
    nseg = 37 * 16 + (1 if odd else 0)
    n = nperseg + hop * (nseg - 1) + hop // 3
    x = make_sig(rng, (3, n)) + 5.0
    kw = dict(nperseg=nperseg, noverlap=hop, window="hann", detrend=detrend)
    before = _kernels.launches["frames_power"]
    _, _, s = msig.spectrogram(x, **kw)
    assert _kernels.launches["frames_power"] == before + 1
    assert s.shape[-1] == nseg
    assert_peak_close(s, oracle.spectrogram(x, **kw)[2], TOL)
    kw2 = dict(kw, return_onesided=False, scaling="spectrum")
    assert_peak_close(msig.spectrogram(x, **kw2)[2], oracle.spectrogram(x, **kw2)[2], TOL)


def test_spectrogram_custom_params(rng):
    x = make_sig(rng, (4096,))
    kw = dict(nperseg=200, noverlap=50, nfft=256, scaling="spectrum", detrend="linear")
    assert_peak_close(msig.spectrogram(x, **kw)[2], oracle.spectrogram(x, **kw)[2], TOL)


STFT_CASES = [
    {},
    {"nperseg": 128},
    {"nperseg": 128, "noverlap": 100},
    {"nperseg": 128, "nfft": 256},
    {"boundary": "even"},
    {"boundary": "odd"},
    {"boundary": "constant"},
    {"boundary": None},
    {"padded": False},
    {"scaling": "psd"},
    {"detrend": "constant"},
    {"return_onesided": False},
]


@pytest.mark.parametrize("case", STFT_CASES)
def test_stft_matches_oracle(rng, case):
    x = make_sig(rng, (4000,))
    f_ref, t_ref, z_ref = oracle.stft(x, fs=8000.0, **case)
    f, t, z = msig.stft(x, fs=8000.0, **case)
    assert tuple(z.shape) == z_ref.shape
    np.testing.assert_allclose(as_np(f), f_ref, rtol=1e-6)
    np.testing.assert_allclose(as_np(t), t_ref, rtol=1e-6)
    assert_peak_close(z, z_ref, TOL)


def test_stft_multidim_axis0(rng):
    x = make_sig(rng, (3000, 4))
    z_ref = oracle.stft(x, axis=0)[2]
    z = msig.stft(x, axis=0)[2]
    assert tuple(z.shape) == z_ref.shape
    assert_peak_close(z, z_ref, TOL)


def test_stft_complex_input(rng):
    x = make_sig(rng, (2048,), complex_=True)
    assert_peak_close(msig.stft(x)[2], oracle.stft(x)[2], TOL)


@pytest.mark.parametrize("case", [{}, {"nfft": 4096}, {"nfft": 1000}, {"scaling": "spectrum"},
                                  {"detrend": False}, {"window": "hann"},
                                  {"return_onesided": False}])
def test_periodogram_matches_oracle(rng, case):
    x = make_sig(rng, (2048,))
    assert_peak_close(msig.periodogram(x, **case)[1], oracle.periodogram(x, **case)[1], TOL)


def test_parseval_periodogram(rng):
    x = make_sig(rng, (4096,))
    fs = 10.0
    _, p = msig.periodogram(x, fs=fs, detrend=False)
    total = float(as_np(p).astype(np.float64).sum()) * fs / len(x)
    np.testing.assert_allclose(total, np.mean(x.astype(np.float64) ** 2), rtol=1e-5)


@pytest.mark.parametrize("n,N,hop,detrend", [(5000, 256, 100, False), (5000, 256, 100, "constant"),
                                             (40000, 1024, 512, "constant"),
                                             (40000, 2048, 320, False), (3000, 64, 7, "constant")])
def test_fused_frames_match_oracle_random_window(rng, n, N, hop, detrend):
    """Fused per-frame rfft with a random window vs the FP64 frames of the oracle
    (reference: tests/test_metal_fft_workaround.py:97-131)."""
    x = rng.standard_normal((3, n)).astype(np.float32)
    w = rng.standard_normal(N).astype(np.float32)
    dev = torch.device("cuda")
    nseg = (n - N) // hop + 1
    idx = np.arange(N)[None, :] + hop * np.arange(nseg)[:, None]
    fr = x.astype(np.float64)[:, idx]
    if detrend == "constant":
        fr = fr - fr.mean(-1, keepdims=True)
    ref = np.fft.rfft(fr * w.astype(np.float64), axis=-1)
    z = _kernels.rfft_frames(torch.from_numpy(x).to(dev), torch.from_numpy(w).to(dev), N, hop,
                             detrend, power=False)
    assert_peak_close(z, ref, TOL)
    pw = _kernels.rfft_frames(torch.from_numpy(x).to(dev), torch.from_numpy(w).to(dev), N, hop,
                              detrend, power=True)
    assert_peak_close(pw, np.abs(ref) ** 2, TOL)


def test_nperseg_one_and_tiny(rng):
    x = rng.standard_normal(64).astype(np.float32)
    assert_peak_close(msig.welch(x, nperseg=1)[1], oracle.welch(x, nperseg=1)[1], TOL)
    assert_peak_close(msig.welch(x, nperseg=1, detrend="linear")[1],
                      oracle.welch(x, nperseg=1, detrend="linear")[1], TOL)
    y = rng.standard_normal(64).astype(np.float32)
    assert_peak_close(msig.csd(x, y, nperseg=1)[1], oracle.csd(x, y, nperseg=1)[1], TOL)
    z = rng.standard_normal(32).astype(np.float32)
    assert_peak_close(msig.stft(z, nperseg=1)[2], oracle.stft(z, nperseg=1)[2], TOL)


@pytest.mark.parametrize("func", ["welch", "csd"])
def test_empty_input_shapes(func, rng):
    x = np.array([], dtype=np.float32)
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        if func == "welch":
            import scipy.signal as sps

            f_ref, p_ref = sps.welch(x)
            f, p = msig.welch(x)
        else:
            import scipy.signal as sps

            y = rng.standard_normal(0).astype(np.float32)
            f_ref, p_ref = sps.csd(x, y)
            f, p = msig.csd(x, y)
    assert tuple(f.shape) == f_ref.shape
    assert tuple(p.shape) == p_ref.shape


def test_results_are_deterministic(rng):
    x = make_sig(rng, (8, 1 << 15))
    a = msig.welch(x, nperseg=1024)[1]
    b = msig.welch(x, nperseg=1024)[1]
    assert torch.equal(a, b)


def test_torch_device_input_passthrough(rng):
    x = torch.from_numpy(make_sig(rng, (1 << 16,))).cuda()
    f, p = msig.welch(x, nperseg=1024)
    assert p.device == x.device
    assert_peak_close(p, oracle.welch(x.cpu().numpy(), nperseg=1024)[1], TOL)


def test_kernel_timing_events(rng):
    from mlx_signal_amd import _kernels

    x = torch.from_numpy(make_sig(rng, (4, 1 << 16))).cuda()
    ref = msig.welch(x, nperseg=1024)[1]
    _kernels.kernel_timing_arm(2)
    for _ in range(3):  # the third call finds no free event pair and is not timed
        p = msig.welch(x, nperseg=1024)[1]
    tot, n = _kernels.kernel_timing_read()
    assert n == 2 and tot > 0.0
    assert torch.equal(p, ref)
    assert _kernels.kernel_timing_read() == (0.0, 0)


@pytest.mark.parametrize("nperseg,noverlap,cplx,detrend", [
    (256, 128, False, False),        # hop == N/2 stream
    (1024, 512, False, "constant"),
    (256, 200, False, "constant"),   # general stream
    (200, 150, False, False),        # nperseg < nfft
    (256, 128, True, False),         # complex input
])
def test_frames_virtual_zero_extension_equals_padded(rng, nperseg, noverlap, cplx, detrend):
    """The per-frame kernel reading stft's boundary zeros and padding in place
    (lead / nvalid) gives the same bits as the kernel on the explicitly
    zero-extended rows."""
    n = 9000 + nperseg // 3
    x = make_sig(rng, (3, n), complex_=cplx)
    dev = torch.device("cuda")
    xm = torch.from_numpy(x).to(dev)
    nfft = 256 if nperseg == 200 else nperseg
    nstep = nperseg - noverlap
    lead = nperseg // 2
    nadd = (-(n + 2 * lead - nperseg) % nstep) % nperseg
    n_ext = n + 2 * lead + nadd
    xp = torch.cat([torch.zeros(3, lead, dtype=xm.dtype, device=dev), xm,
                    torch.zeros(3, lead + nadd, dtype=xm.dtype, device=dev)], -1)
    import scipy.signal as sps

    win = torch.from_numpy(sps.get_window("hann", nperseg).astype(np.float32)).to(dev)
    onesided = not cplx
    (a,) = _kernels.spectral_frames("complex", xm, None, win, nperseg, nstep, nfft, detrend,
                                    onesided, False, 0.7, lead, n_ext)
    (b,) = _kernels.spectral_frames("complex", xp, None, win, nperseg, nstep, nfft, detrend,
                                    onesided, False, 0.7)
    assert a.shape == b.shape and torch.equal(a, b)
    if not cplx:  # only the complex frames (stft) read through the extension
        with pytest.raises(Exception, match="complex per-frame"):
            _kernels.spectral_frames("power", xm, None, win, nperseg, nstep, nfft, detrend,
                                     True, True, 0.7, lead, n_ext)


def test_stft_reads_boundary_zeros_in_place(rng):
    """stft(boundary='zeros', padded=True) runs one fused per-frame launch on
    the unextended rows (the kernel reads the zeros) and matches the oracle."""
    x = make_sig(rng, (4, 20000))
    xm = torch.from_numpy(x).cuda()
    before = _kernels.launches["frames_complex"]
    _, _, z = msig.stft(xm, nperseg=512)
    assert _kernels.launches["frames_complex"] == before + 1
    ref = oracle.stft(x, nperseg=512)[2]
    assert_peak_close(z, ref, TOL)


@pytest.mark.parametrize("fs", [1.0, 48000.0, 250.0, 44100.0, 1e-3, 7.3])
@pytest.mark.parametrize("nperseg", [16, 256, 1000, 1024, 4096])
def test_grids_are_fresh_and_exact(rng, fs, nperseg):
    """The frequency and time grids returned beside each result (copied from
    the cached constants by the call's own kernels on the fused paths) are
    scipy's, bit for bit, and owned by the caller: mutating them leaves the
    next call unaffected."""
    n = 3 * nperseg + 5
    x = make_sig(rng, (2, n))
    y = make_sig(rng, (2, n))
    fref = np.fft.rfftfreq(nperseg, 1 / fs).astype(np.float32)
    calls = [(lambda: msig.welch(x, fs=fs, nperseg=nperseg), oracle.welch(x, fs=fs, nperseg=nperseg)),
             (lambda: msig.csd(x, y, fs=fs, nperseg=nperseg), oracle.csd(x, y, fs=fs, nperseg=nperseg)),
             (lambda: msig.coherence(x, y, fs=fs, nperseg=nperseg),
              oracle.coherence(x, y, fs=fs, nperseg=nperseg)),
             (lambda: msig.welch(x, fs=fs, nperseg=nperseg, average="median"),
              oracle.welch(x, fs=fs, nperseg=nperseg, average="median")),
             (lambda: msig.spectrogram(x, fs=fs, nperseg=nperseg),
              oracle.spectrogram(x, fs=fs, nperseg=nperseg)),
             (lambda: msig.stft(x, fs=fs, nperseg=nperseg), oracle.stft(x, fs=fs, nperseg=nperseg))]
    for fn, ref in calls:
        got = fn()
        grids = got[:-1]
        for g, r in zip(grids, ref[:-1]):
            np.testing.assert_array_equal(as_np(g), np.asarray(r).astype(np.float32))
            g.fill_(-1.0)  # the caller owns it
        for g, r in zip(fn()[:-1], ref[:-1]):
            np.testing.assert_array_equal(as_np(g), np.asarray(r).astype(np.float32))
    np.testing.assert_array_equal(as_np(calls[0][0]()[0]), fref)
    xc = (x + 1j * y).astype(np.complex64)  # two-sided
    f, _ = msig.welch(xc, fs=fs, nperseg=nperseg)
    np.testing.assert_array_equal(as_np(f), np.fft.fftfreq(nperseg, 1 / fs).astype(np.float32))
    f, _ = msig.welch(x, fs=fs, nperseg=nperseg, return_onesided=False)
    np.testing.assert_array_equal(as_np(f), np.fft.fftfreq(nperseg, 1 / fs).astype(np.float32))
    _, _, z = msig.stft(x, fs=fs, nperseg=nperseg)
    t, _ = msig.istft(z, fs=fs, nperseg=nperseg)
    t_ref, _ = oracle.istft(z.cpu().numpy(), fs=fs, nperseg=nperseg)
    np.testing.assert_array_equal(as_np(t), t_ref.astype(np.float32))
    t.fill_(-1.0)
    np.testing.assert_array_equal(as_np(msig.istft(z, fs=fs, nperseg=nperseg)[0]),
                                  t_ref.astype(np.float32))


def test_means_repeat_bitwise_across_calls_streams_and_row_counts(rng):
    """Repeated calls, a call on another stream, and 5000 rows (then five of
    them) give the same bits and match the oracle (the cached workspace and
    its partials are reused call after call)."""
    dev = torch.device("cuda")
    x = make_sig(rng, (5000, 1200))
    y = make_sig(rng, (5000, 1200))
    xm, ym = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    for fn in (lambda: msig.welch(xm, nperseg=256)[1], lambda: msig.csd(xm, ym, nperseg=128)[1],
               lambda: msig.coherence(xm, ym, nperseg=512)[1]):
        first = fn().clone()
        for _ in range(3):
            assert torch.equal(fn(), first)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            other = fn()
        torch.cuda.current_stream().wait_stream(side)
        assert torch.equal(other, first)
    sel = [0, 1, 4095, 4096, 4999]
    p = msig.welch(xm, nperseg=256)[1]
    assert_peak_close(p[sel], oracle.welch(x[sel], nperseg=256)[1], TOL)
    _, p5 = msig.welch(xm[sel], nperseg=256)  # few rows after many: the buffer is reused
    assert_peak_close(p5, oracle.welch(x[sel], nperseg=256)[1], TOL)
