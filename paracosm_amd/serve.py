"""Continuous-batching serving loop over one engine with standing queries.

Update requests (each a small batch of edge updates) arrive over time.  The
server keeps the graph, the indexes and the standing queries on the device
and runs a loop:

* every request that has arrived and fits is admitted into the next device
  batch, in arrival order: up to ``max_batch`` updates, and only while the
  device bytes the engine holds plus the batch's working set stay inside
  ``hbm_budget`` (requests that do not fit wait for the next batch);
* the batch is uploaded, searched and read back; each of its requests
  completes when the batch does.

Because admitted requests are concatenated in arrival order, every update
is answered exactly as if the stream were applied one update at a time --
the same result alone and in the mix (tests/test_gpu_serve.py).  Latency of
an update = completion time of its batch - arrival time of its request.
"""
from __future__ import annotations

import dataclasses
import time
from typing import List, Optional, Sequence

import numpy as np

from .engine import UPDATE_DTYPE, Matching

REC_BYTES = UPDATE_DTYPE.itemsize


@dataclasses.dataclass
class Request:
    arrival: float                 # seconds from the start of the run
    updates: np.ndarray            # csm_update records
    done: float = float("nan")     # completion time (s from start)
    counts: Optional[np.ndarray] = None   # [n, num_queries] exact counts
    safe: Optional[np.ndarray] = None


@dataclasses.dataclass
class ServeStats:
    requests: int
    updates: int
    batches: int
    elapsed: float
    updates_per_s: float
    matches: int
    p50_ms: float
    p99_ms: float
    max_ms: float
    mean_batch: float
    refused_for_memory: int


def poisson_requests(updates: np.ndarray, offered_per_s: float, seed: int = 0,
                     max_request: int = 4096) -> List[Request]:
    """Cut a stream into requests with log-uniform sizes in [1, max_request]
    arriving as a Poisson process whose update rate is ``offered_per_s``."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x53455256]))
    out: List[Request] = []
    t = 0.0
    i = 0
    n = len(updates)
    mean_size = (max_request - 1) / np.log(max_request) if max_request > 1 else 1.0
    rate = offered_per_s / mean_size  # requests per second
    while i < n:
        size = int(np.exp(rng.uniform(0.0, np.log(max_request))))
        size = max(1, min(size, n - i))
        t += rng.exponential(1.0 / rate)
        out.append(Request(arrival=t, updates=updates[i:i + size]))
        i += size
    return out


class Server:
    def __init__(self, engine: Matching, *, max_batch: int = 65534, hbm_budget: int = 0):
        self.m = engine
        self.nq = engine.num_queries
        self.max_batch = max_batch
        self.hbm_budget = hbm_budget
        # device and pinned buffers through the engine's own HIP runtime
        self.d_upd = engine.device_buffer(max_batch * REC_BYTES)
        self.d_cnt = engine.device_buffer(max_batch * self.nq * 8)
        self.d_safe = engine.device_buffer(max_batch)
        self.h_upd = engine.host_buffer(max_batch * REC_BYTES)
        self.h_cnt = engine.host_buffer(max_batch * self.nq * 8)
        self.h_safe = engine.host_buffer(max_batch)
        self.refused = 0

    def close(self) -> None:
        for b in (self.h_upd, self.h_cnt, self.h_safe, self.d_upd, self.d_cnt, self.d_safe):
            b.close()

    def working_set(self, n: int) -> int:
        """Device bytes a batch of n updates needs beyond the engine's own."""
        return n * (REC_BYTES + 8 * self.nq + 1)

    def fits(self, n: int) -> bool:
        if not self.hbm_budget:
            return True
        return self.m.device_bytes() + self.working_set(n) <= self.hbm_budget

    def run_batch(self, reqs: Sequence[Request], t0: float) -> None:
        ups = np.concatenate([r.updates for r in reqs]) if len(reqs) > 1 else reqs[0].updates
        n = len(ups)
        nb = n * REC_BYTES
        self.h_upd.array[:nb] = np.ascontiguousarray(ups, UPDATE_DTYPE).view(np.uint8)
        self.d_upd.upload(self.h_upd.array[:nb])
        self.m.apply_batch_device(self.d_upd.ptr, n, self.d_cnt.ptr, self.d_safe.ptr)
        self.d_cnt.download(self.h_cnt.array[: n * self.nq * 8])
        self.d_safe.download(self.h_safe.array[:n])
        done = time.perf_counter() - t0
        cnt = self.h_cnt.array[: n * self.nq * 8].view(np.uint64).reshape(n, self.nq).copy()
        safe = self.h_safe.array[:n].copy()
        o = 0
        for r in reqs:
            k = len(r.updates)
            r.counts = cnt[o:o + k]
            r.safe = safe[o:o + k]
            r.done = done
            o += k

    def serve(self, requests: List[Request]) -> ServeStats:
        """Serve requests in arrival order in real time (requests must be
        sorted by arrival)."""
        t0 = time.perf_counter()
        i = 0
        batches = 0
        n_req = len(requests)
        while i < n_req:
            now = time.perf_counter() - t0
            if requests[i].arrival > now:
                time.sleep(min(requests[i].arrival - now, 0.05))
                continue
            batch: List[Request] = []
            size = 0
            while i < n_req and requests[i].arrival <= now:
                k = len(requests[i].updates)
                if batch and size + k > self.max_batch:
                    break
                if not self.fits(size + k):
                    self.refused += 1
                    if not batch:
                        raise MemoryError("a single request exceeds the HBM budget")
                    break
                batch.append(requests[i])
                size += k
                i += 1
            self.run_batch(batch, t0)
            batches += 1
        elapsed = time.perf_counter() - t0
        lat = []
        matches = 0
        for r in requests:
            lat.extend([r.done - r.arrival] * len(r.updates))
            matches += int(r.counts.sum())
        lat_ms = 1e3 * np.asarray(lat)
        n_upd = sum(len(r.updates) for r in requests)
        return ServeStats(requests=n_req, updates=n_upd, batches=batches, elapsed=elapsed,
                          updates_per_s=n_upd / elapsed, matches=matches,
                          p50_ms=float(np.percentile(lat_ms, 50)),
                          p99_ms=float(np.percentile(lat_ms, 99)),
                          max_ms=float(lat_ms.max()), mean_batch=n_upd / max(batches, 1),
                          refused_for_memory=self.refused)
