"""Autotuner for batched and served spectral jobs.

For each FFT size it times candidate launch shapes of the batch kernels on a
representative ragged batch (several channel counts and lengths, welch and
csd), with HIP events on the device, and keeps the fastest.  The knobs are
the tile shape of a job's decomposition:

* ``target_blocks`` - blocks one job's geometry aims for (how finely a job's
  segments are split across blocks);
* ``max_run``       - segment pairs one thread group sums before its partial
  is written (run length of the register accumulation);
* ``blocks_per_cu`` - resident blocks per CU of a per-class persistent
  launch;
* ``radix_schedule`` - order of the FFT passes: 0 runs the radix-16 passes
  first and the remainder radix last, 1 the remainder radix first (a
  separate kernel only where nfft is not a power of 16).

Batches run welch and csd at hop = nfft/2 in one persistent launch, which
has radix schedule 0 and its own occupancy, so for them only the tile shape
(``target_blocks``, ``max_run``) is tuned.  Tables for batches planned with
one launch per class (``per_class_launches=True``) tune all four knobs and
are cached apart.

The timing loop is the library's own (``msig_batch_autotune``): it plans
every candidate over the probe jobs, warms the clocks, and times runs between
HIP events on the stream, so any host of the C ABI tunes the same way.  The
chosen table is cached per device name in a JSON file so later processes
reuse it.  Jobs are planned with their size's tuned knobs both alone and in a
batch, so tuning never breaks the alone-equals-in-mix property.
"""

from __future__ import annotations

import ctypes
import itertools
import json
import os
from pathlib import Path

import torch

from . import _hip
from .batch import SpectralBatch, SpectralRequest, tuning_struct

__all__ = ["autotune", "cached_tuning", "candidates_for", "load_tuning", "save_tuning",
           "time_candidates", "CANDIDATES"]

CANDIDATES = [
    dict(target_blocks=tb, max_run=mr, blocks_per_cu=bpc)
    for tb, mr, bpc in itertools.product((256, 1024, 4096), (8, 32), (2, 4, 8))
]
# the single launch's knobs: tile shape only
SINGLE_CANDIDATES = [
    dict(target_blocks=tb, max_run=mr) for tb, mr in itertools.product((256, 1024, 4096), (8, 32))
]


def _schedules(nfft: int):
    """Radix schedules that differ for this size (nfft = 16^m r, r > 1)."""
    r = int(nfft)
    while r >= 16:
        r //= 16
    return (0, 1) if r > 1 else (0,)


def candidates_for(nfft: int, base=None, per_class_launches: bool = True):
    """Tile-shape candidates crossed with the size's distinct radix schedules
    (per-class launches), or the single launch's tile shapes at schedule 0."""
    if not per_class_launches:
        return [dict(c, radix_schedule=0) for c in (base or SINGLE_CANDIDATES)]
    return [dict(c, radix_schedule=sch) for c in (base or CANDIDATES) for sch in _schedules(nfft)]


def _cache_path() -> Path:
    root = os.environ.get("MSIG_TUNING_CACHE")
    if root:
        return Path(root)
    return Path(os.path.expanduser("~/.cache/mlx_signal_amd/tuning.json"))


def _device_key(device, per_class_launches: bool = False) -> str:
    props = torch.cuda.get_device_properties(device)
    mode = "per-class" if per_class_launches else "single"
    return f"{props.name}|{getattr(props, 'gcnArchName', '')}|{props.multi_processor_count}|{mode}"


def load_tuning(device=None, per_class_launches: bool = False):
    p = _cache_path()
    if not p.exists():
        return None
    try:
        table = json.loads(p.read_text()).get(
            _device_key(device or torch.cuda.current_device(), per_class_launches))
    except (ValueError, OSError):
        return None
    if not table:
        return None
    return {int(k): v for k, v in table.items()}


def save_tuning(tuning, device=None, per_class_launches: bool = False) -> Path:
    p = _cache_path()
    p.parent.mkdir(parents=True, exist_ok=True)
    data = {}
    if p.exists():
        try:
            data = json.loads(p.read_text())
        except (ValueError, OSError):
            data = {}
    data[_device_key(device or torch.cuda.current_device(), per_class_launches)] = {
        str(k): v for k, v in tuning.items()}
    p.write_text(json.dumps(data, indent=1) + "\n")
    return p


def _probe_requests(nfft, dev, gen):
    reqs = []
    for ch, n, kind in ((1, 1 << 14, "welch"), (8, 1 << 17, "csd"), (64, 1 << 18, "welch"),
                        (4, 1 << 20, "welch"), (32, 1 << 15, "csd")):
        n = max(n, 2 * nfft)
        x = torch.randn(ch, n, device=dev, generator=gen)
        y = torch.randn(ch, n, device=dev, generator=gen) if kind == "csd" else None
        reqs.append(SpectralRequest(kind, x, y, nperseg=nfft))
    return reqs


def time_candidates(requests, nfft, candidates, reps=5, per_class_launches: bool = True):
    """Median ms per batch run of `requests` (all of FFT size nfft) under each
    candidate knob dict, timed by the library (msig_batch_autotune), with the
    given launch mode."""
    lib = _hip.load()
    b = SpectralBatch(requests)
    base = tuning_struct(None, per_class_launches)
    try:
        n = len(candidates)
        tiles = (_hip.MsigTile * n)()
        for i, c in enumerate(candidates):
            tiles[i] = _hip.MsigTile(int(c.get("max_run", 0)), int(c.get("target_blocks", 0)),
                                     int(c.get("blocks_per_cu", 0)), int(c.get("radix_schedule", 0)))
        ms = (ctypes.c_float * n)()
        best = ctypes.c_int32(-1)
        stream = ctypes.c_void_p(torch.cuda.current_stream(b.device).cuda_stream)
        rc = lib.msig_batch_autotune(b.jobs, len(requests),
                                     ctypes.byref(base) if base is not None else None, tiles, n,
                                     int(reps), ms, ctypes.byref(best), stream)
        _hip.check(rc, "batch autotune")
        return [float(v) for v in ms], int(best.value)
    finally:
        b.close()


def autotune(nffts=(256, 1024, 4096), reps=5, seed=0, candidates=None, save=True,
             per_class_launches: bool = False):
    """Time every candidate for every nfft; return {nfft: best knobs} for
    batches of the given launch mode."""
    dev = torch.device("cuda", torch.cuda.current_device())
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    best = {}
    report = {}
    for nfft in nffts:
        reqs = _probe_requests(nfft, dev, gen)
        cands = candidates_for(nfft, candidates, per_class_launches)
        ms, i = time_candidates(reqs, nfft, cands, reps, per_class_launches)
        best[nfft] = dict(cands[i])
        order = sorted(range(len(cands)), key=lambda k: ms[k])
        report[nfft] = [(round(ms[k], 4), cands[k]) for k in order[:3]]
        del reqs
    if save:
        save_tuning(best, dev, per_class_launches)
    autotune.last_report = report
    return best


def cached_tuning(nffts=(256, 1024, 4096), per_class_launches: bool = False):
    """The cached table for this device and launch mode, tuning (and caching)
    it if absent."""
    t = load_tuning(per_class_launches=per_class_launches)
    if t is None or any(n not in t for n in nffts):
        t = autotune(nffts, per_class_launches=per_class_launches)
    return t
