"""Shared helpers: tolerance checks against FP64 references."""

import numpy as np
import torch


def as_np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def make_sig(rng, shape, complex_=False):
    """N(0,1) plus a 0.5-amplitude tone at 0.12 cycles/sample
    (the reference tests' make_sig, tests/test_spectral.py:11-19)."""
    x = rng.standard_normal(shape).astype(np.float32)
    n = shape[-1] if isinstance(shape, tuple) else shape
    t = np.arange(n, dtype=np.float32)
    x = x + 0.5 * np.sin(2 * np.pi * 0.12 * t).astype(np.float32)
    if complex_:
        x = (x + 1j * rng.standard_normal(shape)).astype(np.complex64)
    return x


def peak_rel_err(actual, desired) -> float:
    """max |a - d| / max |d| (the parity metric: error relative to the peak bin)."""
    a = as_np(actual).astype(np.complex128)
    d = np.asarray(desired).astype(np.complex128)
    assert a.shape == d.shape, f"shape mismatch: {a.shape} vs {d.shape}"
    if d.size == 0:
        return 0.0
    scale = np.max(np.abs(d))
    if not np.isfinite(scale) or scale == 0.0:
        scale = 1.0
    return float(np.max(np.abs(a - d)) / scale)


def assert_peak_close(actual, desired, tol=1e-5):
    """Every bin within tol x the peak |reference| bin (FP64 reference)."""
    e = peak_rel_err(actual, desired)
    assert e <= tol, f"max error {e:.3e} x peak exceeds {tol:.1e}"


def assert_close(actual, desired, rtol=1e-4, atol_frac=1e-5):
    """|err| <= rtol |ref| + atol_frac max|ref| (the reference suite's helper,
    tests/_utils.py:13-26)."""
    a = as_np(actual)
    d = np.asarray(desired)
    assert a.shape == d.shape, f"shape mismatch: {a.shape} vs {d.shape}"
    scale = np.max(np.abs(d)) if d.size else 1.0
    if not np.isfinite(scale) or scale == 0.0:
        scale = 1.0
    np.testing.assert_allclose(a, d.astype(a.dtype), rtol=rtol, atol=atol_frac * scale)


def assert_device_tensor(x, complex_ok=False):
    assert isinstance(x, torch.Tensor), f"expected torch.Tensor, got {type(x)}"
    if complex_ok:
        assert x.dtype in (torch.float32, torch.complex64), x.dtype
    else:
        assert x.dtype == torch.float32, x.dtype
