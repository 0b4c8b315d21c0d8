"""Continuous-batching serving loop over one engine with standing queries.

Update requests (each a small batch of edge updates) arrive over time.  The
server keeps the graph, the indexes and the standing queries on the device
and overlaps the stages of consecutive batches:

* one batch at a time is *open*: the admission thread (the caller's) appends
  every request that arrives, in arrival order, to it, and starts the
  upload of those records on the engine's transfer stream at once, while
  the device is still searching earlier batches.  A request is appended
  only while the open batch holds at most ``max_batch`` updates and the
  engine's device bytes plus the working sets of every batch in flight stay
  inside ``hbm_budget``; one that does not fit waits for a later batch;
* the device thread closes the open batch the moment it is free (so a batch
  holds everything that arrived while the one before it ran), makes the
  engine stream wait for the batch's uploads, runs it (graph update, index
  update, search), and starts the download of its counts on the transfer
  stream, where it overlaps the next batch's search;
* back on the admission thread, each finished batch's download is waited
  for and its requests completed.

Batches run on the engine one after another in arrival order, and every
update sees exactly the graph before it, so every update is answered as if
the stream were applied one update at a time -- the same result alone and in
the mix (tests/test_gpu_serve.py).  Latency of an update = completion time of
its batch - arrival time of its request.
"""
from __future__ import annotations

import dataclasses
import queue
import threading
import time
from typing import List, Optional, Sequence

import numpy as np

from .engine import MARKS, STREAM_ENGINE, STREAM_TRANSFER, UPDATE_DTYPE, Matching

REC_BYTES = UPDATE_DTYPE.itemsize


@dataclasses.dataclass
class Request:
    arrival: float                 # seconds from the start of the run
    updates: np.ndarray            # csm_update records
    done: float = float("nan")     # completion time (s from start)
    counts: Optional[np.ndarray] = None   # [n, num_queries] exact counts
    safe: Optional[np.ndarray] = None
    status: Optional[np.ndarray] = None   # [n] USTAT_* bits (time limit, result cap)


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
    overlapped_uploads: int = 0    # uploads started while another batch was on the device
    partial_updates: int = 0       # updates whose search the engine's time limit stopped
    # per query: batches in which the time limit dropped part of its search
    query_partial_batches: Optional[List[int]] = None


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


class _Slot:
    """Pinned staging and device buffers of one batch in flight, and its
    three markers (uploaded, searched, downloaded)."""

    def __init__(self, m: Matching, idx: int, max_batch: int, nq: int):
        self.idx = idx
        self.cap = max_batch
        self.d_upd = m.device_buffer(max_batch * REC_BYTES)
        self.d_cnt = m.device_buffer(max_batch * nq * 8)
        self.d_safe = m.device_buffer(max_batch)
        self.h_upd = m.host_buffer(max_batch * REC_BYTES)
        self.h_cnt = m.host_buffer(max_batch * nq * 8)
        self.h_safe = m.host_buffer(max_batch)
        self.mk_up, self.mk_done, self.mk_dl = 3 * idx, 3 * idx + 1, 3 * idx + 2
        self.reset()

    def reset(self) -> None:
        self.reqs: List[Request] = []
        self.n = 0
        self.error: Optional[BaseException] = None
        self.ustatus: Optional[np.ndarray] = None
        self.qstatus: Optional[np.ndarray] = None

    def close(self) -> None:
        for b in (self.h_upd, self.h_cnt, self.h_safe, self.d_upd, self.d_cnt, self.d_safe):
            b.close()


class Server:
    """hbm_budget: device bytes the engine plus the batches in flight may
    hold; 0 = 90 % of the device's memory.  slots: batches in flight at once
    (one open, one on the device, one downloading)."""

    def __init__(self, engine: Matching, *, max_batch: int = 65534, hbm_budget: int = 0,
                 slots: int = 3):
        assert 2 <= slots and 3 * slots <= MARKS
        self.m = engine
        self.nq = engine.num_queries
        self.max_batch = max_batch
        self.slots = [_Slot(engine, i, max_batch, self.nq) for i in range(slots)]
        if not hbm_budget:
            hbm_budget = int(0.9 * engine.mem_info()[1])
        self.hbm_budget = hbm_budget
        self.refused = 0

    def close(self) -> None:
        for s in self.slots:
            s.close()
        self.slots = []

    def working_set(self, n: int) -> int:
        """Device bytes a batch of n updates needs beyond the engine's own."""
        return n * (REC_BYTES + 8 * self.nq + 1)

    def fits(self, n: int, in_flight: int = 0) -> bool:
        return (self.m.device_bytes() + self.working_set(in_flight) + self.working_set(n)
                <= self.hbm_budget)

    # -- the three stages
    def _append(self, slot: _Slot, reqs: Sequence[Request]) -> None:
        """Stage requests behind the open batch's records and start their
        upload; the upload marker is recorded again after it, so waiting on
        it covers every record of the batch."""
        ups = np.concatenate([r.updates for r in reqs]) if len(reqs) > 1 else reqs[0].updates
        o, nb = slot.n * REC_BYTES, len(ups) * REC_BYTES
        slot.h_upd.array[o:o + nb] = np.ascontiguousarray(ups, UPDATE_DTYPE).view(np.uint8)
        self.m.copy_async(slot.d_upd.ptr + o, slot.h_upd.ptr + o, nb)
        self.m.mark(STREAM_TRANSFER, slot.mk_up)
        slot.reqs.extend(reqs)
        slot.n += len(ups)

    def _take(self, cond: threading.Condition) -> Optional[_Slot]:
        """Close the open batch once it holds updates (None: stopping)."""
        with cond:
            while not self._stop and (self._open is None or self._open.n == 0):
                cond.wait()
            slot = self._open
            if slot is None or slot.n == 0:
                return None
            self._open = None
            return slot

    def _device_loop(self, cond: threading.Condition, done: "queue.Queue") -> None:
        m = self.m
        while True:
            slot = self._take(cond)
            if slot is None:
                done.put(None)
                return
            try:
                m.wait_mark(STREAM_ENGINE, slot.mk_up)
                # a time limit (csm_opts.time_limit_s) is a normal outcome
                # here: the batch is applied and its stopped updates flagged
                m.apply_batch_device(slot.d_upd.ptr, slot.n, slot.d_cnt.ptr, slot.d_safe.ptr,
                                     partial_ok=True)
                slot.ustatus = m.update_status(slot.n)
                slot.qstatus = m.query_status()
                # the batch call has synchronised its results: download them
                # on the transfer stream while the next batch searches
                m.mark(STREAM_ENGINE, slot.mk_done)
                m.wait_mark(STREAM_TRANSFER, slot.mk_done)
                m.copy_async(slot.h_cnt.ptr, slot.d_cnt.ptr, slot.n * self.nq * 8)
                m.copy_async(slot.h_safe.ptr, slot.d_safe.ptr, slot.n)
                m.mark(STREAM_TRANSFER, slot.mk_dl)
            except BaseException as exc:   # handed to the caller's thread
                slot.error = exc
            done.put(slot)

    def _complete(self, slot: _Slot, t0: float) -> None:
        self.m.sync_mark(slot.mk_dl)
        t = time.perf_counter() - t0
        n = slot.n
        cnt = slot.h_cnt.array[: n * self.nq * 8].view(np.uint64).reshape(n, self.nq).copy()
        safe = slot.h_safe.array[:n].copy()
        o = 0
        for r in slot.reqs:
            k = len(r.updates)
            r.counts = cnt[o:o + k]
            r.safe = safe[o:o + k]
            r.status = slot.ustatus[o:o + k]
            r.done = t
            o += k
        self._qpartial += (slot.qstatus & 2) != 0

    def serve(self, requests: List[Request]) -> ServeStats:
        """Serve requests in arrival order in real time (requests must be
        sorted by arrival)."""
        for r in requests:
            if len(r.updates) > self.max_batch:
                raise ValueError(f"a request of {len(r.updates)} updates exceeds "
                                 f"max_batch={self.max_batch}")
        cond = threading.Condition()
        done: "queue.Queue" = queue.Queue()
        free = list(self.slots)
        self._qpartial = np.zeros(self.nq, np.int64)
        self._open: Optional[_Slot] = None
        self._stop = False
        in_flight: dict = {}       # slot idx -> updates (open, on the device, downloading)
        dev = threading.Thread(target=self._device_loop, args=(cond, done), daemon=True)
        t0 = time.perf_counter()
        dev.start()
        i = 0
        batches = 0
        overlapped = 0
        n_req = len(requests)
        error: Optional[BaseException] = None

        def reap(block: bool, timeout: float = 0.05) -> None:
            nonlocal error
            try:
                slot = done.get(block=block, timeout=timeout if block else None)
            except queue.Empty:
                return
            if slot is None:
                return
            if slot.error is not None:
                error = error or slot.error
            else:
                self._complete(slot, t0)
            in_flight.pop(slot.idx, None)
            free.append(slot)

        try:
            while i < n_req and error is None:
                reap(False)
                now = time.perf_counter() - t0
                if requests[i].arrival > now:
                    wait = min(requests[i].arrival - now, 0.05)
                    if in_flight:
                        reap(True, wait)
                    else:
                        time.sleep(wait)
                    continue
                added: List[Request] = []
                with cond:
                    slot = self._open
                    if slot is None and free:
                        slot = free.pop()
                        slot.reset()
                        self._open = slot
                        in_flight[slot.idx] = 0
                        batches += 1
                    if slot is not None:
                        busy = sum(v for k, v in in_flight.items() if k != slot.idx)
                        size = slot.n
                        while i < n_req and requests[i].arrival <= now:
                            k = len(requests[i].updates)
                            if size + k > self.max_batch:
                                break
                            if not self.fits(size + k, busy):
                                self.refused += 1
                                if size == 0 and busy == 0:
                                    raise MemoryError("a single request exceeds the HBM budget")
                                break
                            added.append(requests[i])
                            size += k
                            i += 1
                        if added:
                            self._append(slot, added)
                            in_flight[slot.idx] = slot.n
                            overlapped += len(in_flight) > 1
                            cond.notify()
                if not added:
                    # the open batch is full, or memory is short: poll until
                    # the device takes it or a batch finishes; no slot: wait
                    # for a batch to finish
                    reap(True, 0.001 if self._open is not None else 0.05)
            with cond:
                self._stop = True
                cond.notify()
            while in_flight and error is None:
                reap(True)
        finally:
            with cond:
                self._stop = True
                cond.notify()
            dev.join()
            while True:            # complete what the device thread finished
                try:
                    slot = done.get_nowait()
                except queue.Empty:
                    break
                if slot is not None and slot.idx in in_flight:
                    if slot.error is None:
                        self._complete(slot, t0)
                    else:
                        error = error or slot.error
                    in_flight.pop(slot.idx, None)
                    free.append(slot)
            self._open = None
        if error is not None:
            raise error
        elapsed = time.perf_counter() - t0
        lat = []
        matches = 0
        partial = 0
        for r in requests:
            lat.extend([r.done - r.arrival] * len(r.updates))
            matches += int(r.counts.sum())
            partial += int(((r.status & 1) != 0).sum())
        lat_ms = 1e3 * np.asarray(lat)
        n_upd = sum(len(r.updates) for r in requests)
        return ServeStats(requests=n_req, updates=n_upd, batches=batches, elapsed=elapsed,
                          updates_per_s=n_upd / elapsed, matches=matches,
                          p50_ms=float(np.percentile(lat_ms, 50)),
                          p99_ms=float(np.percentile(lat_ms, 99)),
                          max_ms=float(lat_ms.max()), mean_batch=n_upd / max(batches, 1),
                          refused_for_memory=self.refused, overlapped_uploads=overlapped,
                          partial_updates=partial,
                          query_partial_batches=self._qpartial.tolist())
