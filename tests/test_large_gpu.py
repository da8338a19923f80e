"""Sizes past 2^31 elements: the kernels index rows, segments and outputs with
64-bit offsets.  A row at a 2^31-element offset of a large input must give
exactly the bits of the same row run alone (the launch geometry of one row
is the same in both calls here, so the sums run in the same order).  The
inputs are generated on the device (12 GB and 8.6 GB); MI355X holds 288 GB."""

import pytest
import torch

import mlx_signal_amd as msig

pytestmark = pytest.mark.gpu


def test_welch_rows_past_2_31_elements():
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    n = 1 << 30  # 3 rows: row 2 starts at element 2^31
    x = torch.randn(3, n, device=dev, generator=g)
    from mlx_signal_amd import _kernels

    assert (_kernels.mean_geometry("psd", 3, n, 1024, 512, 1024)
            == _kernels.mean_geometry("psd", 1, n, 1024, 512, 1024))
    _, p = msig.welch(x, nperseg=1024)
    for r in (0, 2):
        _, p_row = msig.welch(x[r:r + 1].clone(), nperseg=1024)
        assert torch.equal(p[r], p_row[0]), r
    # and the level of N(0,1) noise: one-sided density 2 sigma^2 / fs = 2 at fs = 1
    level = float(p[2, 2:-1].mean())
    assert abs(level - 2.0) / 2.0 < 2e-3, level
    del x, p
    torch.cuda.empty_cache()


def test_spectrogram_output_past_2_31_elements():
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    rows, n = 16, 1 << 27  # 16 x (2^20 - 1) frames x 129 bins = 2.16e9 outputs
    x = torch.randn(rows, n, device=dev, generator=g)
    f, t, s = msig.spectrogram(x, nperseg=256, noverlap=128)
    assert s.shape == (rows, 129, (n - 256) // 128 + 1)
    assert s.numel() > (1 << 31)
    for r in (0, rows - 1):
        _, _, s_row = msig.spectrogram(x[r:r + 1].clone(), nperseg=256, noverlap=128)
        assert torch.equal(s[r], s_row[0]), r
    del x, s
    torch.cuda.empty_cache()


def test_csd_and_istft_past_2_31_elements():
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    n = 1 << 30
    x = torch.randn(3, n, device=dev, generator=g)
    y = 0.5 * x + torch.randn(3, n, device=dev, generator=g)
    _, c = msig.csd(x, y, nperseg=512)
    _, c_row = msig.csd(x[2:3].clone(), y[2:3].clone(), nperseg=512)
    assert torch.equal(c[2], c_row[0])
    del x, y, c, c_row
    torch.cuda.empty_cache()
    rows, n = 16, 1 << 27
    x = torch.randn(rows, n, device=dev, generator=g)
    _, _, z = msig.stft(x, nperseg=256)  # (16, 129, 2^20 + 1) complex: 2.2e9 values
    assert z.numel() > (1 << 31)
    _, xr = msig.istft(z, nperseg=256)
    _, xr_row = msig.istft(z[rows - 1:rows].contiguous(), nperseg=256)
    assert torch.equal(xr[rows - 1], xr_row[0])
    err = float((xr[rows - 1, :n] - x[rows - 1]).abs().max() / x[rows - 1].abs().max())
    assert err < 1e-5, err
    del x, z, xr, xr_row
    torch.cuda.empty_cache()


def test_median_frames_past_2_31_elements():
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(6)
    rows, n = 1 << 14, 1 << 18  # per-frame power: 2^14 x 2047 x 129 = 4.3e9 values
    x = torch.randn(rows, n, device=dev, generator=g)
    _, p = msig.welch(x, nperseg=256, average="median")
    for r in (0, rows - 1):
        _, p_row = msig.welch(x[r:r + 1].clone(), nperseg=256, average="median")
        assert torch.equal(p[r], p_row[0]), r
    del x, p
    torch.cuda.empty_cache()
