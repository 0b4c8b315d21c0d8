"""ctypes binding of oracle/liboracle.so (built by oracle/Makefile).

Test infrastructure only: it is the checker, never the thing measured or
shipped.  See csm_oracle.cpp for the reference files each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_graph_create.restype = ctypes.c_void_p
        L.orc_graph_create.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, _u32p, _u32p, _u32p]
        L.orc_graph_destroy.argtypes = [ctypes.c_void_p]
        L.orc_graph_num_edges.restype = ctypes.c_uint64
        L.orc_graph_num_edges.argtypes = [ctypes.c_void_p]
        L.orc_classify.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_brute_count.restype = ctypes.c_uint64
        L.orc_brute_count.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint64)]
        L.orc_run.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                              _u8p, _u32p, _u32p, _u32p,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                              ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_size_t), ctypes.c_uint64]
        L.orc_symbi_dump.restype = ctypes.c_long
        L.orc_symbi_dump.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, _u32p,
                                     ctypes.c_size_t]
        L.orc_index_dump.restype = ctypes.c_long
        L.orc_index_dump.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_uint32, _u32p, ctypes.c_size_t]
        L.orc_d2_filtered_count.restype = ctypes.c_uint64
        L.orc_d2_filtered_count.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_dag_parents.restype = ctypes.c_uint32
        L.orc_dag_parents.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _u32p]
        L.orc_graphflow_limited.restype = ctypes.c_int
        L.orc_graphflow_limited.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                            _u8p, _u32p, _u32p, _u32p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double,
                                            ctypes.c_double]
        L.orc_graphflow_orders.restype = ctypes.c_int
        L.orc_graphflow_orders.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u32p]
        L.orc_turboflux_orders.restype = ctypes.c_int
        L.orc_turboflux_orders.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u32p, _u32p]
        L.orc_symbi_work.restype = ctypes.c_int
        L.orc_symbi_work.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _u8p,
                                     _u32p, _u32p, _u32p, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class OGraph:
    """A reference-semantics graph (sorted neighbour vectors) held by the oracle."""

    def __init__(self, vlabels, src, dst, elab):
        vl, s, d, l = _u32(vlabels), _u32(src), _u32(dst), _u32(elab)
        self._keep = (vl, s, d, l)
        self.h = lib().orc_graph_create(len(vl), vl, len(s), s, d, l)

    @classmethod
    def from_query(cls, q):
        return cls(q.vlabels, q.src, q.dst, q.elabels)

    def num_edges(self) -> int:
        return int(lib().orc_graph_num_edges(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_graph_destroy(self.h)
            self.h = None


def classify(query: OGraph, l1: int, l2: int, el: int) -> bool:
    return bool(lib().orc_classify(query.h, l1, l2, el))


def brute_count(data: OGraph, query: OGraph, homo: bool = False, want_hash: bool = False):
    h = ctypes.c_uint64(0)
    n = lib().orc_brute_count(data.h, query.h, int(homo), ctypes.byref(h) if want_hash else None)
    return (int(n), int(h.value)) if want_hash else int(n)


class RunResult:
    def __init__(self, pos, neg, hsh, initial, seconds):
        self.pos, self.neg, self.hash, self.initial, self.seconds = pos, neg, hsh, initial, seconds


def run(algo: str, data: OGraph, query: OGraph, add, v1, v2, lab, *, threads: int = 1,
        homo: bool = False, want_hash: bool = False, want_initial: bool = False,
        check_index: bool = False, limit: int | None = None,
        max_results: int = 0) -> RunResult:
    add = np.ascontiguousarray(np.asarray(add, dtype=np.uint8))
    v1, v2, lab = _u32(v1), _u32(v2), _u32(lab)
    n = len(v1) if limit is None else min(limit, len(v1))
    pos = np.zeros(n, dtype=np.uint64)
    neg = np.zeros(n, dtype=np.uint64)
    hsh = np.zeros(n, dtype=np.uint64) if want_hash else None
    init = ctypes.c_uint64(0)
    secs = ctypes.c_double(0.0)
    fail = ctypes.c_size_t(0)
    rc = lib().orc_run(algo.encode(), data.h, query.h, n, add, v1, v2, lab,
                       pos.ctypes.data, neg.ctypes.data,
                       hsh.ctypes.data if hsh is not None else None,
                       threads, int(homo), ctypes.byref(init) if want_initial else None,
                       int(check_index), ctypes.byref(secs), ctypes.byref(fail), max_results)
    if rc == -4:
        raise ValueError("max_results is supported for algo 'graphflow' only")
    if rc == -1:
        raise ValueError(f"unknown algorithm {algo!r}")
    if rc == -2:
        raise RuntimeError(f"deletion error at update {fail.value}")
    if rc == -3:
        raise AssertionError(f"SymBi index diverged from BuildDCS after update {fail.value}")
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")
    return RunResult(pos, neg, hsh, int(init.value) if want_initial else None, secs.value)


def graphflow_limited(data: OGraph, query: OGraph, add, v1, v2, lab, *, limit: int,
                      max_results: int = 0, update_seconds: float = 0.0,
                      total_seconds: float = 0.0):
    """Sequential GraphFlow over the first `limit` updates with a per-update
    result cap and time limits: (counts, unfinished) where unfinished marks
    the updates whose search ran out of time (see orc_graphflow_limited)."""
    add = np.ascontiguousarray(np.asarray(add, dtype=np.uint8))
    v1, v2, lab = _u32(v1), _u32(v2), _u32(lab)
    n = min(limit, len(v1))
    cnt = np.zeros(n, np.uint64)
    unf = np.zeros(n, np.uint8)
    rc = lib().orc_graphflow_limited(data.h, query.h, n, add, v1, v2, lab, cnt.ctypes.data,
                                     unf.ctypes.data, max_results, update_seconds, total_seconds)
    if rc == -2:
        raise RuntimeError("deletion error")
    return cnt, unf.astype(bool)


def symbi_work(data: OGraph, query: OGraph, add, v1, v2, lab, *, homo: bool = False):
    """Sequential SymBi over the stream: (work, counts) per update, where work
    is the number of DCS entries its searches walked under the
    ExtendableVertex order rule (see orc_symbi_work)."""
    add = np.ascontiguousarray(np.asarray(add, dtype=np.uint8))
    v1, v2, lab = _u32(v1), _u32(v2), _u32(lab)
    n = len(v1)
    work = np.zeros(n, np.uint64)
    cnt = np.zeros(n, np.uint64)
    fail = ctypes.c_size_t(0)
    rc = lib().orc_symbi_work(data.h, query.h, n, add, v1, v2, lab, int(homo), work.ctypes.data,
                              cnt.ctypes.data, ctypes.byref(fail))
    if rc == -2:
        raise RuntimeError(f"deletion error at update {fail.value}")
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")
    return work, cnt


def graphflow_orders(data: OGraph, query: OGraph, k: int, ne: int) -> np.ndarray:
    out = np.zeros(max(ne, 1) * k, dtype=np.uint32)
    n = lib().orc_graphflow_orders(data.h, query.h, out)
    return out[: n * k].reshape(n, k)


def turboflux_orders(data: OGraph, query: OGraph, k: int, ne: int):
    """TurboFlux's matching orders (one per query edge) and explicit-path
    counts per query vertex over the spanning-tree index of `data` (see
    orc_turboflux_orders)."""
    out = np.zeros(max(ne, 1) * k, dtype=np.uint32)
    paths = np.zeros(k, dtype=np.uint32)
    n = lib().orc_turboflux_orders(data.h, query.h, out, paths)
    return out[: n * k].reshape(n, k), paths


def symbi_dump(data: OGraph, query: OGraph, what: int, nv: int, rows: int) -> np.ndarray:
    """From-scratch SymBi index of `data` (see orc_symbi_dump)."""
    out = np.zeros(max(1, rows * nv), dtype=np.uint32)
    n = lib().orc_symbi_dump(data.h, query.h, what, out, out.size)
    if n < 0:
        raise ValueError("orc_symbi_dump failed")
    return out if what == 0 else out.reshape(rows, nv)


def index_dump(data: OGraph, query: OGraph, what: int, nv: int, rows: int, tree: bool = False,
               dag_src: OGraph = None) -> np.ndarray:
    """From-scratch index of `data` in the layout of csm_index_dump: SymBi's
    DAG, or TurboFlux's spanning tree (tree=True); the root and DAG/tree are
    chosen from dag_src's label frequencies (default: data's)."""
    out = np.zeros(max(1, rows * nv), dtype=np.uint32)
    n = lib().orc_index_dump(dag_src.h if dag_src is not None else None, data.h, query.h,
                             1 if tree else 0, what, out, out.size)
    if n < 0:
        raise ValueError("orc_index_dump failed")
    return out if what == 0 else out.reshape(rows, nv)


def d2_filtered_count(data: OGraph, query: OGraph, tree: bool) -> int:
    """Embeddings whose every vertex passes the d2 flag of a from-scratch
    index (SymBi DAG or TurboFlux tree): equals brute_count iff d2 is sound."""
    return int(lib().orc_d2_filtered_count(data.h, query.h, 1 if tree else 0))


def dag_parents(data: OGraph, query: OGraph, k: int, tree: bool):
    """(root, parents) of the index DAG or TurboFlux tree built on `data`."""
    out = np.zeros(k * k, dtype=np.uint32)
    root = lib().orc_dag_parents(data.h, query.h, 1 if tree else 0, out)
    par = [[int(x) for x in out[u * k:(u + 1) * k] if x != 0xFFFFFFFF] for u in range(k)]
    return int(root), par
