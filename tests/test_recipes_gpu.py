"""Cached launch recipes of the fused means: a repeated welch / csd /
coherence call on device tensors of the same signature skips the planner;
its results must be bit-identical to the planned call's, for every axis and
layout the cache accepts, and a different argument must not hit another
call's recipe."""

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
import oracle
from _utils import assert_peak_close
from mlx_signal_amd import spectral

pytestmark = pytest.mark.gpu


def _inputs(shape, seed, layout):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn(*shape, device="cuda", generator=g)
    y = 0.5 * x + torch.randn(*shape, device="cuda", generator=g)
    if layout == "transposed":
        x, y = x.t(), y.t()
    return x, y


@pytest.mark.parametrize("shape,axis,layout", [((3, 5000), -1, "plain"), ((5000,), -1, "plain"),
                                               ((4000, 3), 0, "plain"), ((3, 4000), 0, "transposed"),
                                               ((2, 3000, 3), 1, "plain"), ((2, 3, 3000), -1, "plain")])
@pytest.mark.parametrize("func", ["welch", "csd", "coherence"])
def test_recipe_repeats_are_bitwise_equal(func, shape, axis, layout):
    x, y = _inputs(shape, 7, layout)
    if x.shape[axis] < 1000:
        axis = -1
    kw = dict(fs=100.0, nperseg=256, axis=axis)
    call = {"welch": lambda: msig.welch(x, **kw), "csd": lambda: msig.csd(x, y, **kw),
            "coherence": lambda: msig.coherence(x, y, **kw)}[func]
    f0, p0 = call()
    hits = spectral.recipe_hits
    for _ in range(3):
        f1, p1 = call()
        assert torch.equal(f1, f0) and torch.equal(p1, p0)
        assert f1.data_ptr() != f0.data_ptr()  # fresh tensors per call
    # layouts whose analysed axis is contiguous are cached; the others plan
    # (and copy) on every call
    cacheable = x.stride(axis) == 1
    assert spectral.recipe_hits == hits + (3 if cacheable else 0)
    xh = x.cpu().numpy()
    yh = y.cpu().numpy()
    ref = {"welch": lambda: oracle.welch(xh, **kw)[1], "csd": lambda: oracle.csd(xh, yh, **kw)[1],
           "coherence": lambda: oracle.coherence(xh, yh, **kw)[1]}[func]()
    assert tuple(p0.shape) == ref.shape
    assert_peak_close(p0, ref, 1e-5)


def test_recipe_keys_separate_arguments():
    x, _ = _inputs((2, 8192), 3, "plain")
    a = msig.welch(x, nperseg=512)[1]
    b = msig.welch(x, nperseg=512, scaling="spectrum")[1]
    c = msig.welch(x, nperseg=1024)[1]
    d = msig.welch(x, nperseg=512, window="blackman")[1]
    a2 = msig.welch(x, nperseg=512)[1]
    assert torch.equal(a, a2)
    assert a.shape != c.shape
    assert not torch.allclose(a, b) and not torch.allclose(a, d)
    with msig.config_context(precision="fp16_split"):
        e = msig.welch(x, nperseg=256)[1]
    f = msig.welch(x, nperseg=256)[1]
    assert not torch.equal(e, f)  # the precision setting is part of the key
    np.testing.assert_allclose(e.cpu().numpy(), f.cpu().numpy(), rtol=1e-4, atol=1e-6)


def test_recipe_caller_owns_results():
    x, _ = _inputs((2, 8192), 5, "plain")
    f0, p0 = msig.welch(x, nperseg=512)
    f0.zero_()
    p0.zero_()
    f1, p1 = msig.welch(x, nperseg=512)
    assert f1[1] > 0 and p1.abs().sum() > 0
