"""Ragged batches of Welch / CSD / coherence requests in one submission.

No counterpart in the reference, which launches one Metal kernel per call
(src/mlx_signal_processing/_stft_metal.py:530-557, 703-736).  Here many
requests of uneven channel count, length and nperseg are planned once
(``msig_batch_create``), and each run is kernels only: one persistent launch
walks the ragged work queue of every welch and csd request at hop = nfft/2
(all FFT sizes), one per (nfft, kind) class covers any other request, and one
finalize launch completes every request (``msig_batch_run``).  A run allocates nothing and
copies nothing from the host, so it can be captured in a HIP graph
(:meth:`SpectralBatch.graph`) and replayed.

Each request is computed with the same block geometry it gets when planned
alone under the same tuning, so its result is bitwise identical alone and in
any mix.

Requests follow the scipy signatures of :func:`welch`, :func:`csd` and
:func:`coherence` for the mean forms the fused kernels cover: real float32
input, power-of-two nfft in [16, 4096], constant or no detrend, mean average.
"""

from __future__ import annotations

import ctypes
import dataclasses

import torch

from . import _hip, _plan
from ._device import compute_device

__all__ = ["SpectralRequest", "SpectralBatch", "tuning_struct"]


@dataclasses.dataclass
class SpectralRequest:
    """One welch / csd / coherence call: the same arguments as the functions."""

    kind: str  # "welch", "csd" or "coherence"
    x: torch.Tensor
    y: torch.Tensor | None = None
    fs: float = 1.0
    window: object = "hann"
    nperseg: int | None = None
    noverlap: int | None = None
    nfft: int | None = None
    detrend: object = "constant"
    return_onesided: bool = True
    scaling: str = "density"


def tuning_struct(tuning, per_class_launches: bool = False) -> _hip.MsigTuning | None:
    """{nfft: {"max_run", "target_blocks", "blocks_per_cu", "radix_schedule"}}
    -> MsigTuning (None for the defaults)."""
    if not tuning and not per_class_launches:
        return None
    t = _hip.MsigTuning()
    t.per_class_launches = int(bool(per_class_launches))
    for nfft, knobs in (tuning or {}).items():
        l2 = int(nfft).bit_length() - 1
        t.max_run[l2] = int(knobs.get("max_run", 0))
        t.target_blocks[l2] = int(knobs.get("target_blocks", 0))
        t.blocks_per_cu[l2] = int(knobs.get("blocks_per_cu", 0))
        t.radix_schedule[l2] = int(knobs.get("radix_schedule", 0))
    return t


_KINDS = {"welch": _hip.PSD_MEAN, "csd": _hip.CSD_MEAN, "coherence": _hip.COHERENCE}


class SpectralBatch:
    """A planned ragged batch.  ``run()`` launches it on the current stream;
    ``outputs[i]`` is request i's result tensor ((..., bins)), rewritten by
    every run; ``freqs[i]`` its frequency grid (the batch's own copy).

    Welch and csd requests with noverlap = nperseg/2 = nfft/2 all run in one
    persistent launch; ``per_class_launches=True`` gives them one launch per
    (nfft, kind) class instead (the results are the same bit for bit).  Runs
    of one batch must not overlap: they share its workspace and work counter."""

    def __init__(self, requests, tuning=None, per_class_launches: bool = False):
        self.device = compute_device()
        lib = _hip.load()
        n = len(requests)
        if n == 0:
            raise ValueError("a batch needs at least one request")
        jobs = (_hip.MsigBatchJob * n)()
        self._keep = []
        self.outputs = []
        self._grids = []  # cached device grids, shared with other calls
        self._freqs = None
        self.samples = 0
        self.bytes = 0
        for i, r in enumerate(requests):
            jobs[i] = self._plan(r)
        self.jobs = jobs  # the planned msig_batch_job array (kept alive with its tensors)
        self._tuning = tuning_struct(tuning, per_class_launches)
        handle = ctypes.c_void_p()
        tp = ctypes.byref(self._tuning) if self._tuning is not None else None
        _hip.check(lib.msig_batch_create(jobs, n, tp, ctypes.byref(handle)), "batch plan")
        self._handle = handle
        sz = ctypes.c_size_t(0)
        _hip.check(lib.msig_batch_workspace_bytes(handle, ctypes.byref(sz)), "batch workspace")
        self.workspace = torch.empty(max(1, sz.value), dtype=torch.uint8, device=self.device)
        # descriptor block: packed into pinned host memory, copied to the
        # device asynchronously on the current stream (no host-side sync)
        _hip.check(lib.msig_batch_descriptor_bytes(handle, ctypes.byref(sz)), "batch descriptors")
        self.descriptors = torch.empty(max(256, sz.value), dtype=torch.uint8, device=self.device)
        host = torch.empty(self.descriptors.numel(), dtype=torch.uint8, pin_memory=True)
        _hip.check(lib.msig_batch_pack(handle, _hip.ptr(host), host.numel(),
                                       _hip.ptr(self.descriptors)), "batch pack")
        self.descriptors.copy_(host, non_blocking=True)
        self._host_desc = host  # pinned source must outlive the copy
        self.launches = lib.msig_batch_launches(handle)

    def _plan(self, r: SpectralRequest) -> _hip.MsigBatchJob:
        if r.kind not in _KINDS:
            raise ValueError(f"kind must be one of {sorted(_KINDS)}, got {r.kind!r}")
        x = r.x
        if not isinstance(x, torch.Tensor) or x.device != self.device or x.dtype != torch.float32:
            raise ValueError("batch inputs must be float32 tensors on the compute device")
        pair = r.kind != "welch"
        y = r.y
        if pair and (y is None or y.shape != x.shape or y.dtype != x.dtype or y.device != x.device):
            raise ValueError(f"{r.kind} needs y with x's shape, dtype and device")
        n = x.shape[-1]
        batch_shape = tuple(x.shape[:-1])
        x2 = x.reshape(-1, n).contiguous()
        y2 = y.reshape(-1, n).contiguous() if pair else None
        win_np, nperseg = _plan.resolve_window(r.window, _plan.check_nperseg(r.nperseg), n)
        _, nstep, nfft = _plan.resolve_lengths(nperseg, r.noverlap, r.nfft)
        detrend = r.detrend
        const = isinstance(detrend, str) and detrend in ("constant", "c")
        if not (const or detrend is False or detrend is None):
            raise ValueError("batched requests take detrend='constant' or no detrend")
        onesided = bool(r.return_onesided)
        if not _hip.load().msig_eligible(nperseg, nfft, int(onesided), int(const), 0):
            raise ValueError(f"nfft={nfft} is outside the fused power-of-two range")
        if r.kind == "coherence":
            scale, onesided = 1.0, True
        else:
            scale = _plan.spectral_scale(r.scaling, r.fs, win_np)
        bins = nfft // 2 + 1 if onesided else nfft
        rows = x2.shape[0]
        nseg = _plan.segment_count(n, nperseg, nstep)
        dtype = torch.complex64 if r.kind == "csd" else torch.float32
        out = torch.empty(rows, bins, dtype=dtype, device=self.device)
        win = _plan.window_on(self.device, r.window, win_np)
        self._keep += [x2, y2, win, out]
        self.outputs.append(out.reshape(*batch_shape, bins))
        self._grids.append(_plan.freqs_on(self.device, nfft, r.fs, onesided))
        self.samples += rows * n
        self.bytes += (2 if pair else 1) * 4 * rows * n + out.element_size() * rows * bins
        flags = (_hip.F_ONESIDED if onesided else 0) | (_hip.F_DETREND_CONST if const else 0)
        if onesided and r.kind != "coherence":
            flags |= _hip.F_APPLY_FACTOR
        job = _hip.MsigJob(_KINDS[r.kind], flags, rows, n, x2.stride(0),
                           y2.stride(0) if pair else n, nperseg, nstep, nfft, nseg, scale)
        return _hip.MsigBatchJob(job, x2.data_ptr(), y2.data_ptr() if pair else None,
                                 win.data_ptr(), out.data_ptr(), None, None)

    @property
    def freqs(self):
        """Each request's frequency grid, as tensors this batch owns (copies
        of the shared cached grids, made on first access)."""
        if self._freqs is None:
            self._freqs = [g.clone() for g in self._grids]
        return self._freqs

    def run(self):
        """Launch the batch on the current stream of its device."""
        lib = _hip.load()
        stream = ctypes.c_void_p(_hip.stream_of_device(self.device))
        rc = lib.msig_batch_run(self._handle, _hip.ptr(self.descriptors), _hip.ptr(self.workspace),
                                self.workspace.numel(), stream)
        _hip.check(rc, "batch run")
        return self.outputs

    def graph(self) -> torch.cuda.CUDAGraph:
        """Capture one run in a HIP graph (replay with ``g.replay()``)."""
        self.run()  # warm: first launches load code objects outside capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            _hip.load().msig_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
