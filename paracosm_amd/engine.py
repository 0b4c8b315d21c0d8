"""Python host mirror of the reference plugin interface over the C-ABI.

`Matching` mirrors ``class matching`` (ParaCOSM/CSM/matching/matching.h:9-66):
Preprocessing / InitialMatching / AddEdge / RemoveEdge / AddVertex /
RemoveVertex / Classify / GetNum*Results / GetMemoryCost, with the same
argument meaning.  Errors that make the reference print and ``exit(-1)``
(deletion of an absent edge, unknown algorithm) raise :class:`CsmError`
instead.  Everything runs in ``libcsm_mi355x.so`` on the GPU; there is no CPU
fallback, and importing without the built library fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CSM_LIB selects another build of the engine library (A/B measurements)
LIB_PATH = os.environ.get("CSM_LIB") or os.path.join(_HERE, "libcsm_mi355x.so")

CSM_OK = 0
CSM_E_ARG = -1
CSM_E_UNKNOWN_ALGO = -2
CSM_E_DELETE_MISSING = -3
CSM_E_DEVICE = -4
CSM_E_CAPACITY = -5
CSM_E_UNSUPPORTED = -6
CSM_E_TIME_LIMIT = -7
CSM_E_ADMISSION = -8

STREAM_ENGINE, STREAM_TRANSFER = 0, 1   # csm_mark / csm_wait_mark streams
MARKS = 16

USTAT_TIME_LIMIT = 1   # the time limit stopped part of the update's search
USTAT_CAPPED = 2       # a count of the update reached max_results
USTAT_CERTIFIED = 4    # certified before its search to change no count

ALGORITHMS = ("parallel_graphflow", "parallel_symbi", "parallel_turboflux",
              "graphflow", "symbi", "turboflux")


class CsmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Graph(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_uint32), ("vlabels", ctypes.c_void_p),
                ("num_edges", ctypes.c_size_t), ("src", ctypes.c_void_p),
                ("dst", ctypes.c_void_p), ("elabels", ctypes.c_void_p)]


class _Opts(ctypes.Structure):
    _fields_ = [("max_results", ctypes.c_uint64), ("print_prep", ctypes.c_uint32),
                ("print_enum", ctypes.c_uint32), ("homo", ctypes.c_uint32),
                ("threads", ctypes.c_uint32), ("auto_tuning", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32), ("want_hash", ctypes.c_uint32),
                ("time_limit_s", ctypes.c_double), ("device_budget", ctypes.c_uint64),
                ("split_chunk", ctypes.c_uint32), ("device", ctypes.c_uint32),
                ("match_budget", ctypes.c_uint64), ("spill_budget", ctypes.c_uint32),
                ("reference_orders", ctypes.c_uint32), ("want_work", ctypes.c_uint32),
                ("no_graphs", ctypes.c_uint32), ("enum_cap", ctypes.c_uint32)]


UPDATE_DTYPE = np.dtype([("type", np.uint8), ("is_add", np.uint8), ("pad", np.uint16),
                         ("id1", np.uint32), ("id2", np.uint32), ("label", np.uint32)])

_lib = None


def lib():
    """The engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m paracosm_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
        P = ctypes.POINTER
        sig = {
            "csm_create": (vp, [ctypes.c_char_p, P(_Graph), P(_Graph), P(_Opts), P(ctypes.c_int)]),
            "csm_create_from_records": (vp, [ctypes.c_char_p, P(_Graph), vp, sz, P(_Opts),
                                             P(ctypes.c_int)]),
            "csm_parse_text_to_device": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, vp, sz,
                                                        P(vp), P(sz), P(u64)]),
            "csm_free_records": (None, [ctypes.c_int, vp]),
            "csm_register_query": (ctypes.c_int, [vp, P(_Graph), P(u32)]),
            "csm_num_queries": (u32, [vp]),
            "csm_preprocess": (ctypes.c_int, [vp]),
            "csm_initial": (ctypes.c_int, [vp, P(u64)]),
            "csm_classify": (ctypes.c_int, [vp, u32, u32, u32, P(ctypes.c_int)]),
            "csm_add_edge": (ctypes.c_int, [vp, u32, u32, u32, P(u64)]),
            "csm_remove_edge": (ctypes.c_int, [vp, u32, u32, P(u64)]),
            "csm_add_vertex": (ctypes.c_int, [vp, u32, u32]),
            "csm_remove_vertex": (ctypes.c_int, [vp, u32]),
            "csm_apply_batch": (ctypes.c_int, [vp, vp, sz, vp, vp, vp, vp]),
            "csm_apply_batch_device": (ctypes.c_int, [vp, vp, sz, vp, vp, vp]),
            "csm_sync": (ctypes.c_int, [vp]),
            "csm_query_status": (ctypes.c_int, [vp, vp, sz]),
            "csm_update_status": (ctypes.c_int, [vp, vp, sz]),
            "csm_set_time_limit": (ctypes.c_int, [vp, ctypes.c_double]),
            "csm_update_work": (ctypes.c_int, [vp, vp, sz]),
            "csm_embeddings": (ctypes.c_int, [vp, vp, sz, P(sz), P(u64), P(u32)]),
            "csm_totals": (ctypes.c_int, [vp, P(u64), P(u64), P(u64)]),
            "csm_memory": (ctypes.c_int, [vp, P(sz), P(sz)]),
            "csm_device_bytes": (ctypes.c_int, [vp, P(u64)]),
            "csm_stats": (ctypes.c_int, [vp, P(u64), sz]),
            "csm_index_dump": (ctypes.c_int, [vp, u32, u32, vp, sz]),
            "csm_query_orders": (ctypes.c_int, [vp, u32, vp, sz, vp, ctypes.POINTER(ctypes.c_size_t)]),
            "csm_device_alloc": (ctypes.c_int, [vp, sz, P(vp)]),
            "csm_device_free": (ctypes.c_int, [vp, vp]),
            "csm_host_alloc": (ctypes.c_int, [vp, sz, P(vp)]),
            "csm_host_free": (ctypes.c_int, [vp, vp]),
            "csm_copy_to_device": (ctypes.c_int, [vp, vp, vp, sz]),
            "csm_copy_to_host": (ctypes.c_int, [vp, vp, vp, sz]),
            "csm_copy_device": (ctypes.c_int, [vp, vp, vp, sz]),
            "csm_stream": (vp, [vp]),
            "csm_copy_async": (ctypes.c_int, [vp, vp, vp, sz]),
            "csm_mark": (ctypes.c_int, [vp, ctypes.c_int, u32]),
            "csm_wait_mark": (ctypes.c_int, [vp, ctypes.c_int, u32]),
            "csm_sync_mark": (ctypes.c_int, [vp, u32]),
            "csm_mem_info": (ctypes.c_int, [vp, P(u64), P(u64)]),
            "csm_parse_text": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, vp, sz, vp, sz, P(sz),
                                              P(u64)]),
            "csm_parse_text_device": (ctypes.c_int, [vp, ctypes.c_int, vp, sz, vp, sz, P(sz),
                                                     P(u64)]),
            "csm_parse_text_device_static": (ctypes.c_int, [vp, ctypes.c_int, vp, sz, vp, sz,
                                                            P(sz), P(u64)]),
            "csm_last_error": (ctypes.c_char_p, [vp]),
            "csm_destroy": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> Sequence[str]:
    return ("csm_create", "csm_create_from_records", "csm_parse_text_to_device",
            "csm_free_records", "csm_register_query", "csm_num_queries", "csm_preprocess",
            "csm_initial", "csm_classify", "csm_add_edge", "csm_remove_edge",
            "csm_add_vertex", "csm_remove_vertex", "csm_apply_batch",
            "csm_apply_batch_device", "csm_sync", "csm_query_status", "csm_update_status",
            "csm_set_time_limit", "csm_update_work", "csm_embeddings", "csm_totals",
            "csm_memory",
            "csm_device_bytes", "csm_stats", "csm_index_dump", "csm_query_orders",
            "csm_device_alloc",
            "csm_device_free", "csm_host_alloc", "csm_host_free", "csm_copy_to_device",
            "csm_copy_to_host", "csm_copy_device", "csm_stream", "csm_copy_async", "csm_mark",
            "csm_wait_mark", "csm_sync_mark", "csm_mem_info", "csm_parse_text",
            "csm_parse_text_device", "csm_parse_text_device_static", "csm_last_error",
            "csm_destroy")


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class _GraphArgs:
    """Keeps the arrays alive while the C side reads them."""

    def __init__(self, vlabels, src, dst, elabels):
        self.vl, self.s, self.d, self.l = _u32(vlabels), _u32(src), _u32(dst), _u32(elabels)
        self.g = _Graph(len(self.vl), self.vl.ctypes.data, len(self.s), self.s.ctypes.data,
                        self.d.ctypes.data, self.l.ctypes.data)


def make_updates(add, v1, v2, label, vertex=None) -> np.ndarray:
    """Structured array of csm_update records (InsertUnit layout)."""
    n = len(v1)
    u = np.zeros(n, dtype=UPDATE_DTYPE)
    u["type"] = ord("e")
    if vertex is not None:
        u["type"][np.asarray(vertex, bool)] = ord("v")
    u["is_add"] = np.asarray(add, np.uint8)
    u["id1"] = _u32(v1)
    u["id2"] = _u32(v2)
    u["label"] = _u32(label)
    return u


def _opt_values(max_results: int = 0, print_prep: bool = False, print_enum: bool = False,
                homo: bool = False, threads: int = 0, auto_tuning: int = 1, max_batch: int = 0,
                want_hash: bool = False, device_budget: int = 0, split_chunk: int = 0,
                device: int = 0, match_budget: int = 0, spill_budget: int = 0,
                reference_orders: bool = False, time_limit_s: float = 0.0,
                want_work: bool = False, no_graphs: bool = False, enum_cap: int = 0) -> dict:
    """csm_opts fields from Matching's keyword arguments."""
    return dict(max_results=max_results, print_prep=int(print_prep), print_enum=int(print_enum),
                homo=int(homo), threads=threads, auto_tuning=auto_tuning, max_batch=max_batch,
                want_hash=int(want_hash), time_limit_s=time_limit_s, device_budget=device_budget,
                split_chunk=split_chunk, device=device, match_budget=match_budget,
                spill_budget=spill_budget, reference_orders=int(reference_orders),
                want_work=int(want_work), no_graphs=int(no_graphs), enum_cap=enum_cap)


class Matching:
    """One device engine: a data graph plus standing queries."""

    @classmethod
    def from_text(cls, algo: str, query, text: bytes, device: int = 0, **opts) -> "Matching":
        """An engine whose data graph is a graph file's text (the reference's
        Graph::LoadFromFile input), parsed on the device and built into the
        CSR there (csm_parse_text_to_device + csm_create_from_records)."""
        L = lib()
        data = bytes(text)
        d_rec, n, bad = ctypes.c_void_p(0), ctypes.c_size_t(0), ctypes.c_uint64(0)
        rc = L.csm_parse_text_to_device(device, PARSE_GRAPH, data if data else None, len(data),
                                        ctypes.byref(d_rec), ctypes.byref(n), ctypes.byref(bad))
        if rc == -1 and bad.value:
            raise ParseError(rc, int(bad.value))
        if rc:
            raise CsmError(rc, "csm_parse_text_to_device failed")
        self = cls.__new__(cls)
        try:
            self._q = _GraphArgs(query.vlabels, query.src, query.dst, query.elabels)
            o = _Opts(**{k: v for k, v in _opt_values(device=device, **opts).items()})
            err = ctypes.c_int(0)
            self.h = L.csm_create_from_records(algo.encode(), ctypes.byref(self._q.g), d_rec,
                                               n.value, ctypes.byref(o), ctypes.byref(err))
        finally:
            L.csm_free_records(device, d_rec)
        if not self.h:
            if err.value == CSM_E_UNKNOWN_ALGO:
                raise CsmError(err.value, "Unknown algorithm")
            raise CsmError(err.value, f"csm_create_from_records failed for {algo!r}")
        self.algo = algo
        self.want_hash = bool(opts.get("want_hash", False))
        return self

    def __init__(self, algo: str, query, vlabels, src, dst, elabels, *,
                 max_results: int = 0, print_prep: bool = False, print_enum: bool = False,
                 homo: bool = False, threads: int = 0, auto_tuning: int = 1,
                 max_batch: int = 0, want_hash: bool = False, device_budget: int = 0,
                 split_chunk: int = 0, device: int = 0, match_budget: int = 0,
                 spill_budget: int = 0, reference_orders: bool = False,
                 time_limit_s: float = 0.0, want_work: bool = False,
                 no_graphs: bool = False, enum_cap: int = 0):
        L = lib()
        self._q = _GraphArgs(query.vlabels, query.src, query.dst, query.elabels)
        d = _GraphArgs(vlabels, src, dst, elabels)
        o = _Opts(max_results, int(print_prep), int(print_enum), int(homo), threads,
                  auto_tuning, max_batch, int(want_hash), time_limit_s, device_budget, split_chunk,
                  device, match_budget, spill_budget, int(reference_orders), int(want_work),
                  int(no_graphs), enum_cap)
        err = ctypes.c_int(0)
        self.h = L.csm_create(algo.encode(), ctypes.byref(self._q.g), ctypes.byref(d.g),
                              ctypes.byref(o), ctypes.byref(err))
        if not self.h:
            if err.value == CSM_E_UNKNOWN_ALGO:
                raise CsmError(err.value, "Unknown algorithm")
            raise CsmError(err.value, f"csm_create failed for {algo!r}")
        self.algo = algo
        self.want_hash = want_hash
        self._query_k = [len(query.vlabels)]

    # -- helpers
    def _check(self, rc: int) -> None:
        if rc != CSM_OK:
            raise CsmError(rc, lib().csm_last_error(self.h).decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().csm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference interface (matching.h:44-62)
    def Preprocessing(self) -> None:
        self._check(lib().csm_preprocess(self.h))

    def InitialMatching(self) -> int:
        n = ctypes.c_uint64(0)
        self._check(lib().csm_initial(self.h, ctypes.byref(n)))
        return int(n.value)

    def AddEdge(self, v1: int, v2: int, label: int) -> int:
        n = ctypes.c_uint64(0)
        self._check(lib().csm_add_edge(self.h, v1, v2, label, ctypes.byref(n)))
        return int(n.value)

    def RemoveEdge(self, v1: int, v2: int) -> int:
        n = ctypes.c_uint64(0)
        self._check(lib().csm_remove_edge(self.h, v1, v2, ctypes.byref(n)))
        return int(n.value)

    def AddVertex(self, vid: int, label: int) -> None:
        self._check(lib().csm_add_vertex(self.h, vid, label))

    def RemoveVertex(self, vid: int) -> None:
        self._check(lib().csm_remove_vertex(self.h, vid))

    def Classify(self, v1: int, v2: int, label: int) -> bool:
        s = ctypes.c_int(0)
        self._check(lib().csm_classify(self.h, v1, v2, label, ctypes.byref(s)))
        return bool(s.value)

    def GetNumInitialResults(self) -> int:
        return self.totals()[0]

    def GetNumPositiveResults(self) -> int:
        return self.totals()[1]

    def GetNumNegativeResults(self) -> int:
        return self.totals()[2]

    def GetMemoryCost(self):
        e, v = ctypes.c_size_t(0), ctypes.c_size_t(0)
        self._check(lib().csm_memory(self.h, ctypes.byref(e), ctypes.byref(v)))
        return int(e.value), int(v.value)

    # -- extensions
    def register_query(self, query) -> int:
        g = _GraphArgs(query.vlabels, query.src, query.dst, query.elabels)
        qid = ctypes.c_uint32(0)
        self._check(lib().csm_register_query(self.h, ctypes.byref(g.g), ctypes.byref(qid)))
        self._query_k.append(len(query.vlabels))
        return int(qid.value)

    @property
    def num_queries(self) -> int:
        return int(lib().csm_num_queries(self.h))

    def query_status(self) -> np.ndarray:
        """Per query of the last batch: bit 1 = stopped at the match budget,
        bit 2 = the time limit dropped part of its search."""
        out = np.zeros(max(1, self.num_queries), np.uint8)
        self._check(lib().csm_query_status(self.h, out.ctypes.data, out.size))
        return out[: self.num_queries]

    def update_status(self, n: int) -> np.ndarray:
        """USTAT_* bits per update of the last batch call (n entries)."""
        out = np.zeros(max(1, n), np.uint8)
        self._check(lib().csm_update_status(self.h, out.ctypes.data, n))
        return out[:n]

    def update_work(self, n: int) -> np.ndarray:
        """Candidate entries tested per update of the last device batch
        (the engine needs want_work=True)."""
        out = np.zeros(max(1, n), np.uint64)
        self._check(lib().csm_update_work(self.h, out.ctypes.data, n))
        return out[:n]

    def embeddings(self):
        """(rows, n_found) of the last searching call (print_enum=True):
        rows is [n, 2 + max k] uint32 -- update index, query id, then the data
        vertex of each query vertex (0xFFFFFFFF past the query's size)."""
        L = lib()
        n, found, w = ctypes.c_size_t(0), ctypes.c_uint64(0), ctypes.c_uint32(0)
        self._check(L.csm_embeddings(self.h, None, 0, ctypes.byref(n), ctypes.byref(found),
                                     ctypes.byref(w)))
        rows = np.zeros((max(1, n.value), w.value), np.uint32)
        self._check(L.csm_embeddings(self.h, rows.ctypes.data, n.value, ctypes.byref(n),
                                     ctypes.byref(found), ctypes.byref(w)))
        return rows[:n.value], int(found.value)

    def set_time_limit(self, seconds: float) -> None:
        self._check(lib().csm_set_time_limit(self.h, float(seconds)))

    def totals(self):
        a, b, c = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(lib().csm_totals(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def stats(self):
        out = (ctypes.c_uint64 * 20)()
        self._check(lib().csm_stats(self.h, out, 20))
        return {"tasks": out[0], "candidates": out[1], "alg_bytes": out[2], "batches": out[3],
                "repacks": out[4], "search_ns": out[5], "search_launches": out[6],
                "max_rounds": out[7], "split_chunk": out[8], "spill_budget": out[9],
                "tuning": out[10], "graph_batches": out[11],
                # (split_chunk, spill_budget) per query size class: up to 8,
                # 12, 16 and 32 vertices
                "size_class_settings": [(out[12 + 2 * c], out[13 + 2 * c]) for c in range(4)]}

    def device_bytes(self) -> int:
        b = ctypes.c_uint64(0)
        self._check(lib().csm_device_bytes(self.h, ctypes.byref(b)))
        return int(b.value)

    def index_dump(self, qid: int, what: int, nv: int, rows: int) -> np.ndarray:
        """SymBi index rows of query `qid` (see csm_index_dump)."""
        out = np.zeros(max(1, rows * nv), np.uint32)
        self._check(lib().csm_index_dump(self.h, qid, what, out.ctypes.data, out.size))
        return out if what == 0 else out.reshape(rows, nv)

    def query_orders(self, qid: int = 0):
        """(orders [n, k], paths [k]) of query `qid`: its matching orders, one
        per query edge, and TurboFlux's explicit-path counts (see
        csm_query_orders; zeros unless parallel_turboflux with
        reference_orders)."""
        n = ctypes.c_size_t(0)
        self._check(lib().csm_query_orders(self.h, qid, None, 0, None, ctypes.byref(n)))
        k = self._query_k[qid]
        out = np.zeros(max(1, n.value * k), np.uint32)
        paths = np.zeros(k, np.uint32)
        self._check(lib().csm_query_orders(self.h, qid, out.ctypes.data, out.size,
                                           paths.ctypes.data, ctypes.byref(n)))
        return out[: n.value * k].reshape(n.value, k), paths

    # -- device buffers on the engine's own HIP runtime (no second runtime)
    def device_buffer(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def host_buffer(self, nbytes: int) -> "HostBuffer":
        return HostBuffer(self, nbytes)

    def stream(self) -> int:
        return int(lib().csm_stream(self.h) or 0)

    # -- overlapping copies (transfer stream + markers; see csm_copy_async)
    def copy_async(self, dst: int, src: int, nbytes: int) -> None:
        rc = lib().csm_copy_async(self.h, dst, src, nbytes)
        if rc:
            raise CsmError(rc, "csm_copy_async failed")

    def mark(self, stream: int, slot: int) -> None:
        rc = lib().csm_mark(self.h, stream, slot)
        if rc:
            raise CsmError(rc, "csm_mark failed")

    def wait_mark(self, stream: int, slot: int) -> None:
        rc = lib().csm_wait_mark(self.h, stream, slot)
        if rc:
            raise CsmError(rc, "csm_wait_mark failed")

    def sync_mark(self, slot: int) -> None:
        rc = lib().csm_sync_mark(self.h, slot)
        if rc:
            raise CsmError(rc, "csm_sync_mark failed")

    def mem_info(self):
        """(free, total) device bytes on the engine's device."""
        f, t = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(lib().csm_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)))
        return int(f.value), int(t.value)

    def _check_partial(self, rc: int, partial_ok: bool) -> bool:
        """True if the call hit the time limit (allowed by partial_ok)."""
        if rc == CSM_E_TIME_LIMIT and partial_ok:
            return True
        self._check(rc)
        return False

    def apply_batch(self, updates: np.ndarray, partial_ok: bool = False):
        """Host arrays in, (pos, neg, safe, hash) out; pos/neg/hash are
        [n, num_queries] uint64, safe is [n] uint8.  With partial_ok, a call
        stopped by the time limit returns its (partial) results instead of
        raising; update_status(n) says which updates they concern."""
        u = np.ascontiguousarray(updates, dtype=UPDATE_DTYPE)
        n, nq = len(u), self.num_queries
        pos = np.zeros((n, nq), np.uint64)
        neg = np.zeros((n, nq), np.uint64)
        safe = np.zeros(n, np.uint8)
        hsh = np.zeros((n, nq), np.uint64)
        rc = lib().csm_apply_batch(self.h, u.ctypes.data, n, pos.ctypes.data,
                                   neg.ctypes.data, safe.ctypes.data,
                                   hsh.ctypes.data if self.want_hash else None)
        self._check_partial(rc, partial_ok)
        return pos, neg, safe, (hsh if self.want_hash else None)

    def apply_batch_device(self, d_updates_ptr: int, n: int, d_counts_ptr: int,
                           d_safe_ptr: int = 0, d_hash_ptr: int = 0,
                           partial_ok: bool = False) -> bool:
        """Device pointers (e.g. DeviceBuffer.ptr) in HBM; results
        stay on the device.  Synchronous with respect to the engine's stream's
        control read-back."""
        if n < 0:
            raise ValueError(f"negative batch size {n}")
        rc = lib().csm_apply_batch_device(self.h, d_updates_ptr, n, d_counts_ptr,
                                          d_safe_ptr or None, d_hash_ptr or None)
        return self._check_partial(rc, partial_ok)

    def sync(self) -> None:
        self._check(lib().csm_sync(self.h))

    def parse_text_device(self, d_text_ptr: int, nbytes: int, mode: int, d_out_ptr: int,
                          cap: int, static: bool = False) -> int:
        """Text in HBM to device records (see csm_parse_text_device); returns
        the record count.  static=True: the allocation-free entry point
        (csm_parse_text_device_static)."""
        n, bad = ctypes.c_size_t(0), ctypes.c_uint64(0)
        f = lib().csm_parse_text_device_static if static else lib().csm_parse_text_device
        rc = f(self.h, mode, d_text_ptr, nbytes, d_out_ptr, cap, ctypes.byref(n), ctypes.byref(bad))
        if rc == -1 and bad.value:
            raise ParseError(rc, int(bad.value))
        self._check(rc)
        return int(n.value)


class DeviceBuffer:
    """HBM buffer allocated through the engine (freed with the engine or close())."""

    def __init__(self, m: Matching, nbytes: int):
        self.m = m
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p(0)
        m._check(lib().csm_device_alloc(m.h, self.nbytes, ctypes.byref(p)))
        self.ptr = int(p.value)

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        self.m._check(lib().csm_copy_to_device(self.m.h, self.ptr + offset, a.ctypes.data, a.nbytes))

    def download(self, out: np.ndarray, offset: int = 0) -> np.ndarray:
        assert out.flags["C_CONTIGUOUS"] and offset + out.nbytes <= self.nbytes
        self.m._check(lib().csm_copy_to_host(self.m.h, out.ctypes.data, self.ptr + offset,
                                             out.nbytes))
        return out

    def copy_within(self, dst_offset: int, src_offset: int, nbytes: int) -> None:
        """Device-to-device copy inside the buffer (one copy kernel)."""
        assert max(dst_offset, src_offset) + nbytes <= self.nbytes
        self.m._check(lib().csm_copy_device(self.m.h, self.ptr + dst_offset,
                                            self.ptr + src_offset, nbytes))

    def close(self) -> None:
        if self.ptr and self.m.h:
            lib().csm_device_free(self.m.h, self.ptr)
        self.ptr = 0


class HostBuffer:
    """Pinned host buffer allocated through the engine's runtime, viewable
    as a numpy array."""

    def __init__(self, m: Matching, nbytes: int):
        self.m = m
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p(0)
        m._check(lib().csm_host_alloc(m.h, self.nbytes, ctypes.byref(p)))
        self.ptr = int(p.value)
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self.ptr))

    def close(self) -> None:
        if self.ptr and self.m.h:
            lib().csm_host_free(self.m.h, self.ptr)
        self.ptr = 0


PARSE_GRAPH, PARSE_STREAM = 0, 1


class ParseError(CsmError):
    """A malformed line in a graph or stream text (1-based `line`)."""

    def __init__(self, code: int, line: int):
        super().__init__(code, f"malformed line {line}")
        self.line = line


def parse_text(text: bytes, mode: int = PARSE_STREAM, device: int = 0) -> np.ndarray:
    """Parse graph (PARSE_GRAPH) or stream (PARSE_STREAM) text on the device
    into csm_update records (see csm_parse_text)."""
    data = bytes(text)
    cap = data.count(b"\n") + 1
    out = np.zeros(cap, dtype=UPDATE_DTYPE)
    n, bad = ctypes.c_size_t(0), ctypes.c_uint64(0)
    rc = lib().csm_parse_text(device, mode, data if data else None, len(data), out.ctypes.data,
                              cap, ctypes.byref(n), ctypes.byref(bad))
    if rc == -1 and bad.value:
        raise ParseError(rc, int(bad.value))
    if rc:
        raise CsmError(rc, "csm_parse_text failed")
    return out[: n.value]


def records_to_graph(rec: np.ndarray):
    """Graph-file records to (vlabels, src, dst, elabels) as Graph::AddVertex /
    AddEdge would keep them (graph.cpp:23-93): a vertex keeps its first label,
    vertices never declared carry 0xFFFFFFFF."""
    v = rec[rec["type"] == ord("v")]
    e = rec[rec["type"] == ord("e")]
    nv = 0
    if len(v):
        nv = max(nv, int(v["id1"].max()) + 1)
    if len(e):
        nv = max(nv, int(max(e["id1"].max(), e["id2"].max())) + 1)
    vl = np.full(nv, 0xFFFFFFFF, np.uint32)
    vl[v["id1"][::-1]] = v["label"][::-1]   # the first occurrence is written last
    return vl, e["id1"].copy(), e["id2"].copy(), e["label"].copy()


def run_stream(m: Matching, add, v1, v2, label, batch: Optional[int] = None):
    """Apply a whole stream, batched (default: the engine's max batch) or one
    update at a time (batch=1 through AddEdge/RemoveEdge).  Returns per-update
    (pos, neg) summed over queries when batch == 1, else [n, nq] arrays."""
    n = len(v1)
    if batch == 1:
        pos = np.zeros(n, np.uint64)
        neg = np.zeros(n, np.uint64)
        for i in range(n):
            if add[i]:
                pos[i] = m.AddEdge(int(v1[i]), int(v2[i]), int(label[i]))
            else:
                neg[i] = m.RemoveEdge(int(v1[i]), int(v2[i]))
        return pos, neg
    u = make_updates(add, v1, v2, label)
    if batch is None:
        pos, neg, _, _ = m.apply_batch(u)
        return pos, neg
    outs_p, outs_n = [], []
    for s in range(0, n, batch):
        p, q, _, _ = m.apply_batch(u[s:s + batch])
        outs_p.append(p)
        outs_n.append(q)
    return np.concatenate(outs_p), np.concatenate(outs_n)
