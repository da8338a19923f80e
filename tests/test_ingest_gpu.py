"""Pipelined NumPy upload (mlx_signal_amd._ingest): bit-identical to a plain
host-to-device copy for every dtype the boundary accepts, including the
downcasts done inside the staging copy, and the same welch result as a
device-resident input."""

import warnings

import numpy as np
import pytest
import torch

import mlx_signal_amd as msig
from mlx_signal_amd import _ingest
from mlx_signal_amd._array import to_device

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def test_float32_upload_bitwise(rng):
    n = _ingest.CHUNK_BYTES // 4 * 2 + 12345           # two full chunks and a tail
    x = rng.standard_normal(n).astype(np.float32)
    got = to_device(x, DEV)
    assert torch.equal(got.cpu(), torch.from_numpy(x))


@pytest.mark.parametrize("src,dst", [(np.float64, np.float32), (np.complex128, np.complex64),
                                     (np.complex64, np.complex64), (np.float16, np.float32)])
def test_upload_with_downcast(rng, src, dst):
    shape = (3, (_ingest.MIN_BYTES // 8) // 3 + 999)
    x = rng.standard_normal(shape)
    if np.dtype(src).kind == "c":
        x = x + 1j * rng.standard_normal(shape)
    x = x.astype(src)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        got = to_device(x, DEV)
    assert got.shape == x.shape and got.is_contiguous()
    np.testing.assert_array_equal(got.cpu().numpy(), x.astype(dst))


def test_strided_view_upload(rng):
    base = rng.standard_normal((4, 2 * (_ingest.MIN_BYTES // 16) + 3)).astype(np.float32)
    x = base[:, ::2]                                   # non-contiguous view
    got = to_device(x, DEV)
    np.testing.assert_array_equal(got.cpu().numpy(), x)


def test_back_to_back_uploads_reuse_staging(rng):
    xs = [rng.standard_normal(_ingest.MIN_BYTES // 4 + i).astype(np.float32) for i in range(3)]
    outs = [to_device(x, DEV) for x in xs]
    for x, o in zip(xs, outs):
        assert torch.equal(o.cpu(), torch.from_numpy(x))


def test_welch_numpy_equals_device_input(rng):
    x = rng.standard_normal((8, 1 << 20)).astype(np.float32)   # 32 MiB: the pipelined path
    f1, p1 = msig.welch(x, nperseg=1024)
    f2, p2 = msig.welch(torch.from_numpy(x).to(DEV), nperseg=1024)
    assert torch.equal(p1, p2) and torch.equal(f1, f2)
