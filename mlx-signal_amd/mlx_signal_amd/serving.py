"""Continuous-batching request loop for spectral estimation on one GPU.

No counterpart in the reference (a library of synchronous calls).  Requests
(welch / csd / coherence, see :class:`SpectralRequest`) arrive over time;
the loop admits them while an HBM budget and a bounded number of in-flight
batches allow, turns queued requests into ragged batches
(:class:`SpectralBatch`) bounded by estimated GPU time, and launches them on
a few HIP streams: small requests on express streams, large ones on a bulk
stream, so batches formed while others are still running overlap on the
device and a small request does not wait behind a large one.  Completion
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
    io_bytes: int = 0

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
        self.express = True


class SpectralServer:
    """Admit by HBM budget, batch continuously, overlap batches on streams.

    Two lanes share the GPU.  Requests whose estimated GPU time (bytes
    moved / ``rate_gbps``) is at most ``express_ms`` go to the express lane:
    batches of them, formed in arrival order and bounded by ``batch_ms`` of
    estimated time and ``max_batch`` requests, run on their own streams, so a
    small request never waits behind a large one's kernels.  Larger requests
    go to the bulk lane in arrival order, alone or grouped up to ``batch_ms``,
    on one stream.  ``max_inflight`` bounds the batches in flight per lane."""

    def __init__(self, hbm_budget: int = 64 << 30, streams: int = 3, max_inflight: int = 3,
                 max_batch: int = 256, tuning=None, device=None, batch_ms: float = 1.0,
                 express_ms: float = 0.25, rate_gbps: float = 2000.0):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.budget = int(hbm_budget)
        # default-priority streams only: the first carries the bulk lane
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(2, streams))]
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        self.tuning = tuning
        self.batch_ms = batch_ms
        self.express_ms = express_ms
        self.rate = rate_gbps * 1e6  # bytes per ms
        self._lanes = {True: collections.deque(), False: collections.deque()}  # express, bulk
        self.inflight: list = []
        self.bytes_in_flight = 0
        self.completed: list = []
        self.batches = 0
        self._next_stream = 0
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    @property
    def queue(self) -> list:
        """Requests waiting for admission (express lane first)."""
        return list(self._lanes[True]) + list(self._lanes[False])

    def est_ms(self, s) -> float:
        """Estimated GPU time of a request: its input and output bytes at the
        kernels' rate."""
        return s.io_bytes / self.rate

    def submit(self, request: SpectralRequest, arrival: float | None = None) -> ServedRequest:
        nb = request_bytes(request)
        if nb > self.budget:
            raise ValueError(f"request needs {nb} bytes, more than the HBM budget {self.budget}")
        s = ServedRequest(request, self.now() if arrival is None else arrival, nbytes=nb)
        s.io_bytes = 4 * request.x.numel() * (2 if request.y is not None else 1)
        self._lanes[self.est_ms(s) <= self.express_ms].append(s)
        return s

    def _lane_inflight(self, express: bool) -> int:
        return sum(1 for f in self.inflight if f.express == express)

    def _take(self, express: bool):
        """The next batch of one lane from the queue (arrival order within
        the lane), within the budget, batch_ms and max_batch."""
        take, nb, ms = [], 0, 0.0
        for s in self._lanes[express]:
            t = self.est_ms(s)
            if take and (ms + t > self.batch_ms or len(take) >= self.max_batch):
                break
            if self.bytes_in_flight + nb + s.nbytes > self.budget:
                break  # keep the lane's order: wait for memory
            take.append(s)
            nb += s.nbytes
            ms += t
        return take, nb

    def _launch(self, express: bool) -> bool:
        if self._lane_inflight(express) >= self.max_inflight:
            return False
        take, nb = self._take(express)
        if not take:
            return False
        lane = self._lanes[express]
        for _ in take:
            lane.popleft()
        if express:
            stream = self.streams[1 + self._next_stream % (len(self.streams) - 1)]
            self._next_stream += 1
        else:
            stream = self.streams[0]
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
        f = _InFlight(b, take, ev, stream, nb)
        f.express = express
        self.inflight.append(f)
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
        """Reap finished batches, then launch what the budget admits, express
        lane first."""
        self._reap()
        while self._launch(True):
            pass
        while self._launch(False):
            pass

    def _waiting(self) -> bool:
        return bool(self._lanes[True] or self._lanes[False])

    def drain(self) -> None:
        while self._waiting() or self.inflight:
            self.step()

    def serve(self, arrivals):
        """Run an arrival schedule [(t_seconds, SpectralRequest), ...] in real
        time; returns the ServedRequest records in arrival order."""
        self.t0 = time.perf_counter()
        pending = collections.deque(sorted(arrivals, key=lambda a: a[0]))
        records = []
        while pending or self._waiting() or self.inflight:
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
