"""Continuous-batching request loop for spectral estimation on one GPU.

No counterpart in the reference (a library of synchronous calls).  Requests
(welch / csd / coherence, see :class:`SpectralRequest`) arrive over time;
the loop admits them while an HBM budget and a bounded number of in-flight
batches allow, turns everything admitted since the last launch into one
ragged batch (:class:`SpectralBatch`: one persistent launch per FFT size plus
one finalize), and launches it on the next of a few HIP streams, so batches
formed while others are still running overlap on the device.  Completion
is observed with stream events; finished batches return their memory to the
budget and their requests' results to the caller.

Every request is computed with the geometry it would get alone under the
same tuning, so its result is bitwise identical alone and in the mix.
"""

from __future__ import annotations

import collections
import dataclasses
import time

import torch

from .batch import SpectralBatch, SpectralRequest

__all__ = ["ServedRequest", "SpectralServer", "request_bytes"]


def request_bytes(r: SpectralRequest) -> int:
    """Device bytes a request holds while in flight: its inputs, outputs and
    a workspace bound (the block partials never exceed the input size)."""
    n_in = r.x.numel() * (2 if r.y is not None else 1)
    nfft = r.nfft or r.nperseg or 256
    rows = max(1, r.x.numel() // r.x.shape[-1])
    return 4 * n_in + 8 * rows * (nfft // 2 + 1) + 4 * n_in // 2


@dataclasses.dataclass
class ServedRequest:
    request: SpectralRequest
    arrival: float                 # seconds, on the server's clock
    admitted: float | None = None
    done: float | None = None
    result: torch.Tensor | None = None
    freqs: torch.Tensor | None = None
    nbytes: int = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _InFlight:
    def __init__(self, batch, served, event, stream, nbytes):
        self.batch = batch
        self.served = served
        self.event = event
        self.stream = stream
        self.nbytes = nbytes


class SpectralServer:
    """Admit by HBM budget, batch continuously, overlap batches on streams."""

    def __init__(self, hbm_budget: int = 64 << 30, streams: int = 3, max_inflight: int = 3,
                 max_batch: int = 256, tuning=None, device=None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.budget = int(hbm_budget)
        # default-priority streams only
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(streams)]
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        self.tuning = tuning
        self.queue: collections.deque = collections.deque()
        self.inflight: list = []
        self.bytes_in_flight = 0
        self.completed: list = []
        self.batches = 0
        self._next_stream = 0
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, request: SpectralRequest, arrival: float | None = None) -> ServedRequest:
        nb = request_bytes(request)
        if nb > self.budget:
            raise ValueError(f"request needs {nb} bytes, more than the HBM budget {self.budget}")
        s = ServedRequest(request, self.now() if arrival is None else arrival, nbytes=nb)
        self.queue.append(s)
        return s

    def _launch(self):
        if not self.queue or len(self.inflight) >= self.max_inflight:
            return False
        take, nb = [], 0
        while self.queue and len(take) < self.max_batch:
            s = self.queue[0]
            if self.bytes_in_flight + nb + s.nbytes > self.budget:
                break
            take.append(self.queue.popleft())
            nb += s.nbytes
        if not take:
            return False
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        t = self.now()
        for s in take:
            s.admitted = t
        # inputs were written on the default stream; order the batch after them
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            b = SpectralBatch([s.request for s in take], tuning=self.tuning)
            b.run()
            ev = torch.cuda.Event()
            ev.record(stream)
        self.inflight.append(_InFlight(b, take, ev, stream, nb))
        self.bytes_in_flight += nb
        self.batches += 1
        return True

    def _reap(self):
        still = []
        for f in self.inflight:
            if f.event.query():
                t = self.now()
                for s, out, fr in zip(f.served, f.batch.outputs, f.batch.freqs):
                    s.done = t
                    s.result = out
                    s.freqs = fr
                    self.completed.append(s)
                # results are read on other streams later: keep their memory
                # tied to this batch's stream until then
                for out in f.batch.outputs:
                    out.record_stream(torch.cuda.current_stream(self.device))
                self.bytes_in_flight -= f.nbytes
                f.batch.close()
            else:
                still.append(f)
        self.inflight = still

    def step(self) -> None:
        """Reap finished batches, then launch what the budget admits."""
        self._reap()
        while self._launch():
            pass

    def drain(self) -> None:
        while self.queue or self.inflight:
            self.step()

    def serve(self, arrivals):
        """Run an arrival schedule [(t_seconds, SpectralRequest), ...] in real
        time; returns the ServedRequest records in arrival order."""
        self.t0 = time.perf_counter()
        pending = collections.deque(sorted(arrivals, key=lambda a: a[0]))
        records = []
        while pending or self.queue or self.inflight:
            now = self.now()
            while pending and pending[0][0] <= now:
                t, r = pending.popleft()
                records.append(self.submit(r, arrival=t))
            self.step()
        return records


def latency_stats(records) -> dict:
    lat = sorted(r.latency for r in records)
    if not lat:
        return {}

    def pct(p):
        i = min(len(lat) - 1, max(0, int(round(p / 100.0 * (len(lat) - 1)))))
        return lat[i]

    span = max(r.done for r in records) - min(r.arrival for r in records)
    return {"requests": len(lat), "p50_ms": pct(50) * 1e3, "p99_ms": pct(99) * 1e3,
            "mean_ms": sum(lat) / len(lat) * 1e3, "max_ms": lat[-1] * 1e3,
            "throughput_req_s": len(lat) / span if span > 0 else float("nan")}
