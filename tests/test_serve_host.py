"""The serving loop's host logic on the CPU, over a stand-in engine.

The stand-in has the engine calls the server makes (buffers, copies,
markers, the device batch call) and answers update i of the stream with
(its stream position, its first endpoint), after a delay per batch: an
update answered out of order, twice or never, a batch over max_batch or
over the HBM budget, all show here without a GPU.  The GPU tests
(test_gpu_serve.py) run the same server on the real engine against the
oracle.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from paracosm_amd.engine import UPDATE_DTYPE, make_updates
from paracosm_amd.serve import Request, Server, poisson_requests


class _Buf:
    def __init__(self, nbytes):
        self.array = np.zeros(max(nbytes, 1), np.uint8)
        self.ptr = self.array.ctypes.data
        self.closed = False

    def close(self):
        self.closed = True


class FakeEngine:
    num_queries = 2

    def __init__(self, batch_delay=0.0, per_update_delay=0.0, base_bytes=1000):
        self.batch_delay = batch_delay
        self.per_update_delay = per_update_delay
        self.base_bytes = base_bytes
        self.applied = 0
        self.batches = []
        self.running = threading.Lock()
        self.bufs = {}

    def _buf(self, nbytes):
        b = _Buf(nbytes)
        self.bufs[b.ptr] = b
        return b

    device_buffer = host_buffer = _buf

    def mem_info(self):
        return 1 << 40, 1 << 40

    def device_bytes(self):
        return self.base_bytes

    def copy_async(self, dst, src, nbytes):
        ctypes.memmove(dst, src, nbytes)

    def mark(self, stream, slot):
        pass

    wait_mark = mark

    def sync_mark(self, slot):
        pass

    def apply_batch_device(self, d_upd, n, d_cnt, d_safe, partial_ok=False):
        assert self.running.acquire(blocking=False), "two batches on the device at once"
        try:
            recs = np.frombuffer((ctypes.c_uint8 * (n * UPDATE_DTYPE.itemsize)).from_address(d_upd),
                                 UPDATE_DTYPE)
            cnt = np.frombuffer((ctypes.c_uint8 * (n * 16)).from_address(d_cnt), np.uint64)
            cnt = cnt.reshape(n, 2)
            cnt[:, 0] = self.applied + np.arange(n)
            cnt[:, 1] = recs["id1"]
            ctypes.memset(d_safe, 1, n)
            self.applied += n
            self.batches.append(n)
            time.sleep(self.batch_delay + self.per_update_delay * n)
        finally:
            self.running.release()
        return False

    def update_status(self, n):
        return np.zeros(n, np.uint8)

    def query_status(self):
        return np.zeros(self.num_queries, np.uint32)


def _stream(n, seed=0):
    rng = np.random.default_rng(seed)
    v1 = rng.integers(0, 1 << 20, n)
    return make_updates(np.ones(n, bool), v1, v1 + 1, np.zeros(n))


def _check(reqs, ups):
    got = np.concatenate([r.counts for r in reqs])
    np.testing.assert_array_equal(got[:, 0], np.arange(len(ups)))
    np.testing.assert_array_equal(got[:, 1], ups["id1"])
    for r in reqs:
        assert r.done >= r.arrival and len(r.safe) == len(r.updates)


@pytest.mark.parametrize("offered", [2e3, 2e5, 1e9])
def test_every_update_answered_once_in_arrival_order(offered):
    ups = _stream(3000)
    eng = FakeEngine(batch_delay=0.002)
    reqs = poisson_requests(ups, offered, seed=3, max_request=64)
    srv = Server(eng, max_batch=512)
    st = srv.serve(reqs)
    _check(reqs, ups)
    assert sum(eng.batches) == len(ups) and max(eng.batches) <= 512
    assert st.batches == len(eng.batches)
    srv.close()
    assert all(b.closed for b in eng.bufs.values())


def test_open_batch_takes_arrivals_until_the_device_is_free():
    """A slow device and a steady stream: each batch holds what arrived
    while the batch before it ran, so batches stay few and latency stays
    near two batch times, not three."""
    ups = _stream(4000, seed=1)
    eng = FakeEngine(batch_delay=0.05)
    reqs = poisson_requests(ups, 2e4, seed=4, max_request=16)   # about 0.2 s of arrivals
    st = Server(eng, max_batch=65534).serve(reqs)
    _check(reqs, ups)
    print("batches", eng.batches, st)
    assert len(eng.batches) <= 12
    assert st.p99_ms < 3 * 1e3 * eng.batch_delay + 60


def test_hbm_budget_bounds_every_batch_in_flight():
    ups = _stream(600, seed=2)
    eng = FakeEngine(batch_delay=0.001)
    reqs = poisson_requests(ups, 1e9, seed=5, max_request=8)
    srv = Server(eng, max_batch=4096)
    srv.hbm_budget = eng.base_bytes + srv.working_set(40)
    st = srv.serve(reqs)
    _check(reqs, ups)
    assert st.refused_for_memory > 0
    # the open batch and the batches on the device share the 40-update budget
    assert max(eng.batches) <= 40
    srv.hbm_budget = eng.base_bytes + srv.working_set(1)
    with pytest.raises(MemoryError):
        srv.serve([Request(arrival=0.0, updates=ups[:5])])


def test_request_over_max_batch_is_refused():
    ups = _stream(100)
    srv = Server(FakeEngine(), max_batch=64)
    with pytest.raises(ValueError):
        srv.serve([Request(arrival=0.0, updates=ups)])


def test_device_error_reaches_the_caller():
    class Failing(FakeEngine):
        def apply_batch_device(self, *a, **k):
            raise RuntimeError("device batch failed")

    ups = _stream(200)
    reqs = poisson_requests(ups, 1e9, seed=6, max_request=16)
    with pytest.raises(RuntimeError, match="device batch failed"):
        Server(Failing(), max_batch=64).serve(reqs)
