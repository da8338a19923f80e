"""Fused frame+detrend+window+FFT(+power / cross / mean) operators on gfx950.

Counterpart of the reference's fused-kernel module
(src/mlx_signal_processing/_stft_metal.py): the same operator names and
argument meanings (``eligible``, ``rfft_frames``, ``rfft_power_mean``,
``csd_frames``, ``coherence_frames``, ``csd_mean``, ``coherence_mean``), each
one call into libmsig.so on the tensors' current HIP stream.  The generic
``spectral_mean`` / ``spectral_frames`` carry the extra options the planner
uses (raw window + scale applied in FP64, in-kernel onesided factor,
nfft > nperseg zero padding, two-sided and complex input).

Tensors passed in must live on a ROCm device; outputs are allocated there.
"""

from __future__ import annotations

import collections
import ctypes

import torch

from . import _hip

#: launches per kind name, so tests can assert the fused path ran
launches: collections.Counter = collections.Counter()

_MEAN_KINDS = {
    "psd": _hip.PSD_MEAN,
    "csd": _hip.CSD_MEAN,
    "coh_trio": _hip.COH_TRIO_MEAN,
    "coherence": _hip.COHERENCE,
}
_FRAME_KINDS = {
    "complex": _hip.FRAMES_COMPLEX,
    "power": _hip.FRAMES_POWER,
    "csd": _hip.CSD_FRAMES,
    "coh": _hip.COH_FRAMES,
}


def _is_const(detrend) -> bool:
    return isinstance(detrend, str) and detrend in ("constant", "c")


def eligible(x, nperseg: int, nfft: int, sides: str, detrend) -> bool:
    """True if the fused kernels compute this windowed FFT
    (reference: _stft_metal.py:150-160; widened to nfft >= nperseg powers of
    two in [16, 4096], two-sided output and complex input)."""
    if not isinstance(x, torch.Tensor) or x.device.type != "cuda":
        return False
    if x.dtype not in (torch.float32, torch.complex64):
        return False
    if not (detrend is False or detrend is None or _is_const(detrend)):
        return False
    cplx = x.dtype == torch.complex64
    if cplx and sides == "onesided":
        return False
    lib = _hip.load()
    return bool(lib.msig_eligible(int(nperseg), int(nfft), int(sides == "onesided"),
                                  int(_is_const(detrend)), int(cplx)))


# One partial-sum workspace per (device, stream), grown on demand and reused:
# calls on one stream are ordered, calls on different streams use different
# buffers.
_WORKSPACES: dict = {}


def workspace_and_stream(dev: torch.device, nbytes: int):
    """(block-partials buffer, raw handle of the current stream).  The buffer
    is cached per (device, current stream): calls on one stream run in order,
    so they can share it; other streams get their own."""
    stream = _hip.stream_of_device(dev)
    key = (dev.index, stream)
    buf = _WORKSPACES.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=dev)
        _WORKSPACES[key] = buf
    return buf, stream


def _workspace(dev: torch.device, nbytes: int) -> torch.Tensor:
    return workspace_and_stream(dev, nbytes)[0]


def row_major(t: torch.Tensor) -> torch.Tensor:
    """A (rows, n) tensor whose rows are contiguous and do not overlap (a
    copy only when the view is not)."""
    if t.stride(-1) != 1 or (t.dim() == 2 and t.stride(0) < t.shape[1]):
        t = t.contiguous()
    return t


_rows = row_major


def _job(kind, x2, y2, nperseg, nstep, nfft, detrend, onesided, apply_factor, scale,
         lead=0, n_ext=None):
    """Job descriptor for rows x2 (and y2).  lead / n_ext: the segments are cut
    from each row extended by `lead` zeros in front and zeros after it up to
    n_ext samples in all (per-frame kinds; read in the kernel, not copied)."""
    rows, stored = x2.shape
    n = stored if n_ext is None else int(n_ext)
    cplx = x2.dtype == torch.complex64
    flags = (_hip.F_COMPLEX_INPUT if cplx else 0) | (_hip.F_ONESIDED if onesided else 0)
    flags |= _hip.F_APPLY_FACTOR if apply_factor else 0
    flags |= _hip.F_DETREND_CONST if _is_const(detrend) else 0
    nseg = (n - nperseg) // nstep + 1
    ld_y = y2.stride(0) if y2 is not None else stored
    ext = lead != 0 or n != stored
    return _hip.MsigJob(kind, flags, rows, n, x2.stride(0), ld_y, nperseg, nstep, nfft, nseg,
                        float(scale), int(lead), stored if ext else 0)


def _check_inputs(x2, y2, win, nperseg):
    if x2.dim() != 2 or x2.device.type != "cuda":
        raise ValueError("fused kernels take (rows, n) tensors on a ROCm device")
    if y2 is not None and (y2.shape != x2.shape or y2.dtype != x2.dtype or y2.device != x2.device):
        raise ValueError("x and y must have the same shape, dtype and device")
    if win.dtype != torch.float32 or win.numel() != nperseg or not win.is_contiguous():
        raise ValueError("window must be a contiguous float32 tensor of length nperseg")
    if win.device != x2.device:
        raise ValueError("window must live on the input's device")


_PRECISION_FLAGS = {"fp16_split": _hip.F_MFMA_F16, "bf16_split": _hip.F_MFMA_BF16}


def mfma_eligible(kind: str, x2, nperseg: int, nfft: int) -> bool:
    """Shapes the split-precision matrix-core form covers: real-input welch
    means with nperseg == nfft == 256."""
    return (kind == "psd" and x2.dtype == torch.float32 and nperseg == _hip.MFMA_NFFT
            and nfft == _hip.MFMA_NFFT)


def mean_geometry(kind: str, rows: int, n: int, nperseg: int, hop: int, nfft: int,
                  detrend=False, cplx: bool = False) -> dict:
    """The launch plan the library picks for a mean job on the current
    device: units per row, FFT groups per block, units per group (run) and
    blocks per row.  Diagnostic (msig_mean_plan)."""
    lib = _hip.load()
    flags = _hip.F_ONESIDED | (_hip.F_COMPLEX_INPUT if cplx else 0)
    flags |= _hip.F_DETREND_CONST if _is_const(detrend) else 0
    nseg = (n - nperseg) // hop + 1
    job = _hip.MsigJob(_MEAN_KINDS[kind], flags, rows, n, n, n, nperseg, hop, nfft, nseg, 1.0,
                       0, 0)
    out = (ctypes.c_int32 * 4)()
    _hip.check(lib.msig_mean_plan(ctypes.byref(job), out), "mean_plan")
    return {"units": out[0], "gpb": out[1], "run": out[2], "bpr": out[3]}


# Workspace bytes per (kind, flags, rows, nseg, nfft): the block partials'
# size is a function of the launch geometry, which depends on nothing else.
_WS_BYTES: dict = {}


def _side(side, dev):
    """(src, dst) float32 device tensors of one length, or None: the side copy
    a call's kernels make (cached constants returned beside a result)."""
    if side is None:
        return 0, 0, 0
    src, dst = side
    for t in (src, dst):
        if t.dtype != torch.float32 or t.device != dev or not t.is_contiguous():
            raise ValueError(f"side copy buffers must be contiguous float32 tensors on {dev}")
    if src.numel() != dst.numel():
        raise ValueError("side copy buffers differ in length")
    return src.data_ptr(), dst.data_ptr(), src.numel()


def mean_job(kind: str, x2, y2, win, nperseg, nstep, nfft, detrend, onesided=True,
             apply_factor=False, scale=1.0, precision="fp32"):
    """(job descriptor, workspace bytes, launch-counter name) of a fused mean
    over row-major (rows, n) inputs; checks the arguments as spectral_mean."""
    _check_inputs(x2, y2, win, nperseg)
    code = _MEAN_KINDS[kind]
    if code != _hip.PSD_MEAN and y2 is None:
        raise ValueError(f"{kind} needs two signals")
    if precision != "fp32" and precision not in _PRECISION_FLAGS:
        raise ValueError(f"unknown precision {precision!r}")
    job = _job(code, x2, y2, nperseg, nstep, nfft, detrend, onesided, apply_factor, scale)
    mfma = precision != "fp32" and mfma_eligible(kind, x2, nperseg, nfft)
    if mfma:
        job.flags |= _PRECISION_FLAGS[precision]
    key = (job.kind, job.flags, job.rows, job.nseg, job.nfft)  # what the partials depend on
    nbytes = _WS_BYTES.get(key)
    if nbytes is None:
        ws_bytes = ctypes.c_size_t(0)
        _hip.check(_hip.load().msig_workspace_bytes(ctypes.byref(job), ctypes.byref(ws_bytes)),
                   kind)
        nbytes = _WS_BYTES[key] = ws_bytes.value
        if len(_WS_BYTES) > 4096:
            _WS_BYTES.clear()
    return job, nbytes, "mean_" + kind + ("_" + precision if mfma else "")


def spectral_mean(kind: str, x2, y2, win, nperseg, nstep, nfft, detrend, onesided=True,
                  apply_factor=False, scale=1.0, precision="fp32", side=None):
    """Mean over segments of |X|^2 ("psd"), conj(X) Y ("csd"), the
    (Pxx, Pyy, Pxy) trio ("coh_trio") or |Pxy|^2/(Pxx Pyy) ("coherence"),
    times ``scale`` (and the onesided factor when ``apply_factor``).
    x2 / y2: (rows, n).  Returns a tuple of (rows, bins) tensors.

    ``side``: optional (src, dst) float32 device tensors; the call's kernels
    also copy src into dst (the caller's cached frequency grid), so
    returning it takes no copy launch of its own.

    ``precision`` "fp16_split" / "bf16_split" runs the matrix-core form where
    :func:`mfma_eligible` holds and the FP32 kernels elsewhere."""
    x2 = _rows(x2)
    y2 = None if y2 is None else _rows(y2)
    code = _MEAN_KINDS[kind]
    job, nbytes, counter = mean_job(kind, x2, y2, win, nperseg, nstep, nfft, detrend, onesided,
                                    apply_factor, scale, precision)
    lib = _hip.load()
    cplx = x2.dtype == torch.complex64
    bins = nfft // 2 + 1 if (onesided and not cplx) else nfft
    rows = x2.shape[0]
    dev = x2.device
    ws = _workspace(dev, nbytes)
    f32 = dict(dtype=torch.float32, device=dev)
    c64 = dict(dtype=torch.complex64, device=dev)
    if code == _hip.PSD_MEAN or code == _hip.COHERENCE:
        outs = (torch.empty(rows, bins, **f32),)
    elif code == _hip.CSD_MEAN:
        outs = (torch.empty(rows, bins, **c64),)
    else:
        outs = (torch.empty(rows, bins, **f32), torch.empty(rows, bins, **f32),
                torch.empty(rows, bins, **c64))
    o = list(outs) + [None] * (3 - len(outs))
    job.side_src, job.side_dst, job.side_len = _side(side, dev)
    rc = lib.msig_spectral_mean(ctypes.byref(job), _hip.ptr(x2), _hip.ptr(y2), _hip.ptr(win),
                                _hip.ptr(o[0]), _hip.ptr(o[1]), _hip.ptr(o[2]), _hip.ptr(ws),
                                ws.numel(), _hip.stream_of(x2))
    _hip.check(rc, f"spectral_mean({kind})")
    launches[counter] += 1
    return outs


def spectral_frames(kind: str, x2, y2, win, nperseg, nstep, nfft, detrend, onesided=True,
                    apply_factor=False, scale=1.0, lead=0, n_ext=None, side=None):
    """Per-segment X ("complex", times sqrt(scale)), |X|^2 ("power"),
    conj(X) Y ("csd") or the trio ("coh"), times ``scale``.  Returns
    (rows, nseg, bins) tensors.  lead / n_ext: segments are cut from each row
    extended by `lead` zeros in front and zeros after it to n_ext samples
    (the kernel reads the zeros; nothing is copied).  ``side``: an optional
    (src, dst) side copy, as for :func:`spectral_mean`."""
    x2 = _rows(x2)
    y2 = None if y2 is None else _rows(y2)
    _check_inputs(x2, y2, win, nperseg)
    code = _FRAME_KINDS[kind]
    if code in (_hip.CSD_FRAMES, _hip.COH_FRAMES) and y2 is None:
        raise ValueError(f"{kind} frames need two signals")
    lib = _hip.load()
    job = _job(code, x2, y2, nperseg, nstep, nfft, detrend, onesided, apply_factor, scale,
               lead, n_ext)
    cplx = x2.dtype == torch.complex64
    bins = nfft // 2 + 1 if (onesided and not cplx) else nfft
    rows = x2.shape[0]
    nseg = job.nseg
    dev = x2.device
    job.side_src, job.side_dst, job.side_len = _side(side, dev)
    f32 = dict(dtype=torch.float32, device=dev)
    c64 = dict(dtype=torch.complex64, device=dev)
    if code == _hip.FRAMES_POWER:
        outs = (torch.empty(rows, nseg, bins, **f32),)
    elif code in (_hip.FRAMES_COMPLEX, _hip.CSD_FRAMES):
        outs = (torch.empty(rows, nseg, bins, **c64),)
    else:
        outs = (torch.empty(rows, nseg, bins, **f32), torch.empty(rows, nseg, bins, **f32),
                torch.empty(rows, nseg, bins, **c64))
    o = list(outs) + [None] * (3 - len(outs))
    rc = lib.msig_spectral_frames(ctypes.byref(job), _hip.ptr(x2), _hip.ptr(y2), _hip.ptr(win),
                                  _hip.ptr(o[0]), _hip.ptr(o[1]), _hip.ptr(o[2]),
                                  _hip.stream_of(x2))
    _hip.check(rc, f"spectral_frames({kind})")
    launches["frames_" + kind] += 1
    return outs


# ---------------------------------------------------------------------------
# operator-level functions with the reference's signatures; each is one call
# of the matching named C entry point
# ---------------------------------------------------------------------------


def _flat(x):
    n = x.shape[-1]
    return x.reshape(-1, n).contiguous(), x.shape[:-1], n


def _ws(kind, rows, n, nperseg, hop, dev):
    sz = ctypes.c_size_t(0)
    lib = _hip.load()
    _hip.check(lib.msig_mean_workspace_bytes(kind, rows, n, nperseg, hop, ctypes.byref(sz)),
               "workspace")
    ws = _workspace(dev, sz.value)
    return ws, ws.numel()


def rfft_frames(x, win, nperseg: int, hop: int, detrend, power: bool):
    """(..., nseg, nperseg//2+1) windowed (optionally mean-detrended) rfft of
    every segment, complex64, or float32 |X|^2 when ``power``
    (reference: _stft_metal.py:192-220).  ``win`` carries any scale."""
    x2, batch, n = _flat(x)
    _check_inputs(x2, None, win, nperseg)
    nseg = (n - nperseg) // hop + 1
    m1 = nperseg // 2 + 1
    out = torch.empty(x2.shape[0], nseg, m1, device=x2.device,
                      dtype=torch.float32 if power else torch.complex64)
    lib = _hip.load()
    rc = lib.msig_rfft_frames(_hip.ptr(x2), x2.shape[0], n, _hip.ptr(win), nperseg, hop,
                              int(_is_const(detrend)), int(bool(power)), _hip.ptr(out),
                              _hip.stream_of(x2))
    _hip.check(rc, "rfft_frames")
    launches["op_rfft_frames"] += 1
    return out.reshape(*batch, nseg, m1)


def rfft_power_mean(x, win, nperseg: int, hop: int, detrend):
    """(..., nperseg//2+1) mean windowed |rfft|^2 over segments, without the
    onesided factor (reference: _stft_metal.py:530-557)."""
    x2, batch, n = _flat(x)
    _check_inputs(x2, None, win, nperseg)
    m1 = nperseg // 2 + 1
    out = torch.empty(x2.shape[0], m1, device=x2.device, dtype=torch.float32)
    ws, wsb = _ws(_hip.PSD_MEAN, x2.shape[0], n, nperseg, hop, x2.device)
    lib = _hip.load()
    rc = lib.msig_rfft_power_mean(_hip.ptr(x2), x2.shape[0], n, _hip.ptr(win), nperseg, hop,
                                  int(_is_const(detrend)), _hip.ptr(out), _hip.ptr(ws), wsb,
                                  _hip.stream_of(x2))
    _hip.check(rc, "rfft_power_mean")
    launches["op_rfft_power_mean"] += 1
    return out.reshape(*batch, m1)


def csd_frames(x, y, win, nperseg: int, hop: int, detrend):
    """conj(F(x_seg)) * F(y_seg) for all segments (reference: _stft_metal.py:411-413)."""
    x2, batch, n = _flat(x)
    y2, _, _ = _flat(y)
    _check_inputs(x2, y2, win, nperseg)
    nseg = (n - nperseg) // hop + 1
    m1 = nperseg // 2 + 1
    out = torch.empty(x2.shape[0], nseg, m1, device=x2.device, dtype=torch.complex64)
    lib = _hip.load()
    rc = lib.msig_csd_frames(_hip.ptr(x2), _hip.ptr(y2), x2.shape[0], n, _hip.ptr(win), nperseg,
                             hop, int(_is_const(detrend)), _hip.ptr(out), _hip.stream_of(x2))
    _hip.check(rc, "csd_frames")
    launches["op_csd_frames"] += 1
    return out.reshape(*batch, nseg, m1)


def coherence_frames(x, y, win, nperseg: int, hop: int, detrend):
    """(|Fx|^2, |Fy|^2, conj(Fx) Fy) per segment (reference: _stft_metal.py:416-418)."""
    x2, batch, n = _flat(x)
    y2, _, _ = _flat(y)
    _check_inputs(x2, y2, win, nperseg)
    nseg = (n - nperseg) // hop + 1
    m1 = nperseg // 2 + 1
    rows = x2.shape[0]
    pxx = torch.empty(rows, nseg, m1, device=x2.device, dtype=torch.float32)
    pyy = torch.empty_like(pxx)
    pxy = torch.empty(rows, nseg, m1, device=x2.device, dtype=torch.complex64)
    lib = _hip.load()
    rc = lib.msig_coherence_frames(_hip.ptr(x2), _hip.ptr(y2), rows, n, _hip.ptr(win), nperseg,
                                   hop, int(_is_const(detrend)), _hip.ptr(pxx), _hip.ptr(pyy),
                                   _hip.ptr(pxy), _hip.stream_of(x2))
    _hip.check(rc, "coherence_frames")
    launches["op_coherence_frames"] += 1
    shp = (*batch, nseg, m1)
    return pxx.reshape(shp), pyy.reshape(shp), pxy.reshape(shp)


def csd_mean(x, y, win, nperseg: int, hop: int, detrend):
    """Mean conj(F(x_seg)) * F(y_seg) over segments (reference: _stft_metal.py:739-741)."""
    x2, batch, n = _flat(x)
    y2, _, _ = _flat(y)
    _check_inputs(x2, y2, win, nperseg)
    m1 = nperseg // 2 + 1
    out = torch.empty(x2.shape[0], m1, device=x2.device, dtype=torch.complex64)
    ws, wsb = _ws(_hip.CSD_MEAN, x2.shape[0], n, nperseg, hop, x2.device)
    lib = _hip.load()
    rc = lib.msig_csd_mean(_hip.ptr(x2), _hip.ptr(y2), x2.shape[0], n, _hip.ptr(win), nperseg,
                           hop, int(_is_const(detrend)), _hip.ptr(out), _hip.ptr(ws), wsb,
                           _hip.stream_of(x2))
    _hip.check(rc, "csd_mean")
    launches["op_csd_mean"] += 1
    return out.reshape(*batch, m1)


def coherence_mean(x, y, win, nperseg: int, hop: int, detrend):
    """Mean (|Fx|^2, |Fy|^2, conj(Fx) Fy) over segments (reference: _stft_metal.py:744-746)."""
    x2, batch, n = _flat(x)
    y2, _, _ = _flat(y)
    _check_inputs(x2, y2, win, nperseg)
    m1 = nperseg // 2 + 1
    rows = x2.shape[0]
    pxx = torch.empty(rows, m1, device=x2.device, dtype=torch.float32)
    pyy = torch.empty_like(pxx)
    pxy = torch.empty(rows, m1, device=x2.device, dtype=torch.complex64)
    ws, wsb = _ws(_hip.COH_TRIO_MEAN, rows, n, nperseg, hop, x2.device)
    lib = _hip.load()
    rc = lib.msig_coherence_mean(_hip.ptr(x2), _hip.ptr(y2), rows, n, _hip.ptr(win), nperseg,
                                 hop, int(_is_const(detrend)), _hip.ptr(pxx), _hip.ptr(pyy),
                                 _hip.ptr(pxy), _hip.ptr(ws), wsb, _hip.stream_of(x2))
    _hip.check(rc, "coherence_mean")
    launches["op_coherence_mean"] += 1
    return pxx.reshape(*batch, m1), pyy.reshape(*batch, m1), pxy.reshape(*batch, m1)


def kernel_timing_arm(max_launches: int) -> None:
    """Time the fused transform kernel of the next ``max_launches`` mean /
    frames calls on the current device: each launch carries a HIP event pair
    stamped with its dispatch's start and end (instrumentation)."""
    _hip.check(_hip.load().msig_kernel_timing_arm(int(max_launches)), "kernel timing")


def kernel_timing_read() -> tuple[float, int]:
    """(summed kernel milliseconds, launches timed) since the last arm; disarms."""
    tot = ctypes.c_double(0.0)
    n = ctypes.c_int32(0)
    _hip.check(_hip.load().msig_kernel_timing_read(ctypes.byref(tot), ctypes.byref(n)),
               "kernel timing")
    return tot.value, n.value


def istft_frames(z3, weff, nperseg: int):
    """Inverse transform of one-sided spectra, frame by frame, times ``weff``
    (reference: the inverse kernel of istft_frames, _stft_metal.py:817-870).
    z3: (rows, nseg, nperseg/2 + 1) complex64; returns (rows, nseg, nperseg)
    float32 frames."""
    if z3.dtype != torch.complex64 or z3.dim() != 3 or z3.shape[-1] != nperseg // 2 + 1:
        raise ValueError("istft_frames takes (rows, nseg, nperseg/2+1) complex64")
    z3 = z3.contiguous()
    rows, nseg, _ = z3.shape
    frames = torch.empty(rows, nseg, nperseg, dtype=torch.float32, device=z3.device)
    rc = _hip.load().msig_istft_frames(_hip.ptr(z3), rows, nseg, nperseg, _hip.ptr(weff),
                                      _hip.ptr(frames), _hip.stream_of(z3))
    _hip.check(rc, "istft_frames")
    launches["istft_frames"] += 1
    return frames


def overlap_add(frames, win, nperseg: int, hop: int, first: int, count: int):
    """Gather overlap-add with the window-power divide (reference: the OLA
    kernel of istft_frames, _stft_metal.py:793-815).  frames: (rows, nseg,
    nperseg) float32 or complex64; returns (rows, count) of the same dtype,
    samples first .. first + count - 1 of the overlap-added signal."""
    cplx = frames.dtype == torch.complex64
    fr = torch.view_as_real(frames.contiguous()) if cplx else frames.contiguous()
    rows, nseg = frames.shape[0], frames.shape[1]
    ch = 2 if cplx else 1
    out = torch.empty((rows, count, 2) if cplx else (rows, count), dtype=torch.float32,
                      device=frames.device)
    if out.numel() == 0:
        return torch.view_as_complex(out) if cplx else out
    rc = _hip.load().msig_overlap_add(_hip.ptr(fr), rows, nseg, nperseg, hop, ch, _hip.ptr(win),
                                     first, count, _hip.ptr(out), _hip.stream_of(fr))
    _hip.check(rc, "overlap_add")
    launches["overlap_add"] += 1
    return torch.view_as_complex(out) if cplx else out


def istft_ola(z3, weff, win, nperseg: int, first: int, count: int, side=None):
    """istft for hop = nperseg/2 with the overlap-add fused: the result of
    ``overlap_add(istft_frames(z3, weff, nperseg), win, nperseg, nperseg // 2,
    first, count)`` to a few ulps (the window-power divide is folded into
    per-sample weights) without the frames in memory (reference: the two
    kernels of istft_frames, _stft_metal.py:753-815).  z3: (rows, nseg,
    nperseg/2 + 1) complex64; returns (rows, count) float32.  ``side``: an
    optional (src, dst) side copy, as for :func:`spectral_mean`."""
    if z3.dtype != torch.complex64 or z3.dim() != 3 or z3.shape[-1] != nperseg // 2 + 1:
        raise ValueError("istft_ola takes (rows, nseg, nperseg/2+1) complex64")
    z3 = z3.contiguous()
    rows, nseg, _ = z3.shape
    out = torch.empty(rows, count, dtype=torch.float32, device=z3.device)
    if out.numel() == 0:
        if side is not None:
            side[1].copy_(side[0])
        return out
    src, dst, n = _side(side, z3.device)
    rc = _hip.load().msig_istft_ola(_hip.ptr(z3), rows, nseg, nperseg, _hip.ptr(weff),
                                   _hip.ptr(win), first, count, _hip.ptr(out),
                                   ctypes.c_void_p(src), ctypes.c_void_p(dst), n,
                                   _hip.stream_of(z3))
    _hip.check(rc, "istft_ola")
    launches["istft_ola"] += 1
    return out


MEDIAN_MAX_NSEG = 16000


def median_frames(p, scale: float, apply_factor: bool, nfft: int):
    """Median over the segment axis (numpy semantics) of (rows, nseg, bins)
    float32 or complex64 frames (real and imaginary parts separately), times
    ``scale`` and, with ``apply_factor``, the onesided factor of each bin
    (reference: the median average of csd, spectral.py:469-477).  Returns
    (rows, bins) of the input's dtype."""
    rows, nseg, bins = p.shape
    cplx = p.dtype == torch.complex64
    src = torch.view_as_real(p.contiguous()).reshape(rows, nseg, 2 * bins) if cplx else p.contiguous()
    cols = src.shape[-1]
    out = torch.empty(rows, cols, dtype=torch.float32, device=p.device)
    rc = _hip.load().msig_median_frames(_hip.ptr(src), rows, nseg, cols, 2 if cplx else 1,
                                       float(scale), int(bool(apply_factor)), int(nfft),
                                       _hip.ptr(out), _hip.stream_of(src))
    _hip.check(rc, "median_frames")
    launches["median_frames"] += 1
    return torch.view_as_complex(out.reshape(rows, bins, 2)) if cplx else out
