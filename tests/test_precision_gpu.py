"""Split-precision matrix-core form (precision="fp16_split" / "bf16_split"):
parity with the FP64 oracle, with the bound stated per precision, on the
256-point welch configuration and its edge cases; the FP32 kernels keep
every call the form does not cover."""

import ctypes

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
import oracle
from _utils import make_sig, peak_rel_err
from mlx_signal_amd import _hip, _kernels

pytestmark = pytest.mark.gpu

# per-bin error relative to the FP64 peak bin.  FP16 hi + lo carries ~22
# significant bits per operand: the FP32 path's 1e-5 bound holds.  BF16 hi + lo
# carries ~16: the stated bound is 2e-5.
TOLS = {"fp16_split": 1e-5, "bf16_split": 2e-5}


def _welch(x, precision, **kw):
    with msig.config_context(precision=precision):
        return msig.welch(x, **kw)


@pytest.mark.parametrize("precision", ["fp16_split", "bf16_split"])
@pytest.mark.parametrize("case", [
    dict(nperseg=256),                                   # noverlap 128: even nseg
    dict(nperseg=256, noverlap=100),                     # odd hop, odd nseg below
    dict(nperseg=256, detrend=False),
    dict(nperseg=256, scaling="spectrum", window="blackman"),
    dict(nperseg=256, return_onesided=False),
    dict(nperseg=256, noverlap=0, fs=48000.0),
])
@pytest.mark.parametrize("n", [1 << 14, 9999])
def test_precision_form_matches_oracle(rng, precision, case, n):
    x = make_sig(rng, (3, n))
    before = _kernels.launches["mean_psd_" + precision]
    f, p = _welch(x, precision, **case)
    assert _kernels.launches["mean_psd_" + precision] == before + 1
    f_ref, p_ref = oracle.welch(x, **case)
    np.testing.assert_array_equal(f.cpu().numpy(), f_ref.astype(np.float32))
    err = peak_rel_err(p, p_ref)
    assert err <= TOLS[precision], f"{precision} {case} n={n}: {err:.3e}"


@pytest.mark.parametrize("precision", ["fp16_split", "bf16_split"])
@pytest.mark.parametrize("amp", [1e-6, 3e4])
def test_precision_form_amplitude_range(rng, precision, amp):
    """FP16's five exponent bits: the per-pair power-of-two scale keeps tiny
    and large inputs in range."""
    x = (make_sig(rng, (2, 1 << 13)) * amp).astype(np.float32)
    _, p = _welch(x, precision, nperseg=256)
    p_ref = oracle.welch(x, nperseg=256)[1]
    assert np.isfinite(p.cpu().numpy()).all()
    assert peak_rel_err(p, p_ref) <= TOLS[precision]


@pytest.mark.parametrize("precision", ["fp16_split", "bf16_split"])
def test_precision_form_segment_count_edges(rng, precision):
    # one segment, two segments, three (odd: the last has no partner)
    for n in (256, 256 + 128, 256 + 256):
        x = make_sig(rng, (2, n))
        _, p = _welch(x, precision, nperseg=256)
        assert peak_rel_err(p, oracle.welch(x, nperseg=256)[1]) <= TOLS[precision], n


def test_precision_form_constant_input_detrends_to_zero(rng):
    x = np.full((2, 4096), 3.25, np.float32)
    for precision in ("fp16_split", "bf16_split"):
        _, p = _welch(x, precision, nperseg=256)
        assert float(p.abs().max()) <= 1e-12


def test_precision_form_full_size_properties(rng):
    """256 channels x 2^16 (the configuration's size): FP16 and BF16 agree
    with the FP32 kernels to their bounds on every channel, and with the
    oracle on a sample of channels."""
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.randn(256, 1 << 16, device="cuda", generator=g)
    t = torch.arange(1 << 16, device="cuda", dtype=torch.float32)
    x += torch.sin(2 * np.pi * 0.1 * t)
    _, p32 = _welch(x, "fp32", nperseg=256)
    ref = oracle.welch(x[:4].cpu().numpy(), nperseg=256)[1]
    for precision in ("fp16_split", "bf16_split"):
        _, p = _welch(x, precision, nperseg=256)
        assert peak_rel_err(p[:4], ref) <= TOLS[precision]
        peak = p32.abs().amax(dim=-1, keepdim=True)
        assert float(((p - p32).abs() / peak).max()) <= 2 * TOLS[precision]


def test_other_shapes_keep_fp32_kernels(rng):
    x = make_sig(rng, (2, 1 << 13))
    with msig.config_context(precision="fp16_split"):
        before = dict(_kernels.launches)
        msig.welch(x, nperseg=512)
        msig.csd(x, x[::-1].copy(), nperseg=256)
    assert _kernels.launches["mean_psd"] == before.get("mean_psd", 0) + 1
    assert _kernels.launches["mean_csd"] == before.get("mean_csd", 0) + 1
    assert _kernels.launches["mean_psd_fp16_split"] == before.get("mean_psd_fp16_split", 0)


def test_abi_rejects_precision_flags_outside_the_form(rng):
    lib = _hip.load()
    sz = ctypes.c_size_t(0)
    nseg = (4096 - 512) // 256 + 1
    job = _hip.MsigJob(_hip.PSD_MEAN, _hip.F_ONESIDED | _hip.F_MFMA_F16, 1, 4096, 4096, 4096,
                       512, 256, 512, nseg, 1.0)
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_ENOTSUP
    job = _hip.MsigJob(_hip.CSD_MEAN, _hip.F_ONESIDED | _hip.F_MFMA_BF16, 1, 4096, 4096, 4096,
                       256, 128, 256, (4096 - 256) // 128 + 1, 1.0)
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_ENOTSUP
    job.kind = _hip.PSD_MEAN
    job.flags = _hip.F_ONESIDED | _hip.F_MFMA_F16 | _hip.F_MFMA_BF16
    assert lib.msig_workspace_bytes(ctypes.byref(job), ctypes.byref(sz)) == _hip.MSIG_EINVAL
