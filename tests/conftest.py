import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "mlx-signal_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import mlx_signal_amd  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libmsig.so")


@pytest.fixture(autouse=True)
def _pin_gpu_path(request):
    """GPU tests exercise the HIP path (dispatch='gpu' raises instead of
    silently routing to scipy); CPU tests run under 'auto'."""
    if request.node.get_closest_marker("gpu"):
        ctx = mlx_signal_amd.config_context(dispatch="gpu", warn_on_downcast=False)
    else:
        ctx = mlx_signal_amd.config_context(dispatch="auto", warn_on_downcast=False)
    with ctx:
        yield


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
