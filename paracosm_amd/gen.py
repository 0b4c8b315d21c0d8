"""Seeded synthetic workloads for the incremental match search.

Everything here is deterministic given its seed: R-MAT edge lists, vertex and
edge labels, the initial-graph / update-stream split, and random-walk queries.
The reference ships no generator (its datasets are external downloads, see
ParaCOSM/data/README.md), so these are this project's own definitions:

* R-MAT: ``edge_factor << scale`` draws, each descending ``scale`` levels with
  quadrant probabilities (a, b, c, 1-a-b-c); endpoints canonicalised to
  (min, max), self-loops dropped, duplicates removed.
* Vertex labels: uniform over ``n`` labels, or Zipf with P(k) ~ 1/(k+1)^s.
* Queries: a random walk on the initial graph until k distinct vertices are
  seen; the induced subgraph (with the data labels) relabelled 0..k-1 in
  discovery order.  "sparse" means average degree <= 3, "dense" means > 3.

Update streams use the record layout of the reference's ``InsertUnit``
(ParaCOSM/CSM/utils/types.h:34-42): a type ('e' only here), an add/delete
flag, two vertex ids and a label.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

_RMAT_CHUNK = 1 << 22  # draws per chunk; part of the generator's definition


def rmat_edges(scale: int, edge_factor: int, a: float = 0.57, b: float = 0.19,
               c: float = 0.19, seed: int = 0, progress=None):
    """Unique undirected R-MAT edges as (lo, hi) uint32 arrays, sorted by (lo, hi).
    progress(msg), if given, is called after each chunk of draws."""
    n_draws = edge_factor << scale
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x524D4154]))
    keys = []
    ab = a + b
    abc = a + b + c
    for start in range(0, n_draws, _RMAT_CHUNK):
        n = min(_RMAT_CHUNK, n_draws - start)
        u = np.zeros(n, dtype=np.uint64)
        v = np.zeros(n, dtype=np.uint64)
        for _ in range(scale):
            r = rng.random(n)
            bu = (r >= ab).astype(np.uint64)                       # quadrants c, d
            bv = (((r >= a) & (r < ab)) | (r >= abc)).astype(np.uint64)  # quadrants b, d
            u = (u << np.uint64(1)) | bu
            v = (v << np.uint64(1)) | bv
        lo = np.minimum(u, v)
        hi = np.maximum(u, v)
        keep = lo != hi
        keys.append((lo[keep] << np.uint64(32)) | hi[keep])
        if progress:
            progress(f"R-MAT draws {start + n}/{n_draws}")
    key = np.unique(np.concatenate(keys))
    lo = (key >> np.uint64(32)).astype(np.uint32)
    hi = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return lo, hi


def vertex_labels(nv: int, nlabels: int, dist: str = "uniform", s: float = 1.0,
                  seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x564C4142]))
    if dist == "uniform":
        return rng.integers(0, nlabels, size=nv, dtype=np.uint32)
    if dist == "zipf":
        w = 1.0 / np.power(np.arange(1, nlabels + 1, dtype=np.float64), s)
        cdf = np.cumsum(w / w.sum())
        cdf[-1] = 1.0
        return np.searchsorted(cdf, rng.random(nv), side="right").astype(np.uint32)
    raise ValueError(f"unknown label distribution {dist!r}")


def edge_labels(ne: int, nelabels: int, seed: int = 0) -> np.ndarray:
    if nelabels <= 1:
        return np.zeros(ne, dtype=np.uint32)
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x454C4142]))
    return rng.integers(0, nelabels, size=ne, dtype=np.uint32)


@dataclasses.dataclass
class Query:
    """A labelled query graph: vertex labels and an undirected edge list."""
    vlabels: np.ndarray
    src: np.ndarray
    dst: np.ndarray
    elabels: np.ndarray

    @property
    def k(self) -> int:
        return int(self.vlabels.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.src.shape[0])

    @property
    def avg_degree(self) -> float:
        return 2.0 * self.num_edges / max(self.k, 1)


@dataclasses.dataclass
class Workload:
    nv: int
    vlabels: np.ndarray          # uint32[nv]
    init_src: np.ndarray         # uint32[m]
    init_dst: np.ndarray
    init_elab: np.ndarray
    upd_add: np.ndarray          # uint8[n]; 1 insert, 0 delete
    upd_v1: np.ndarray           # uint32[n]
    upd_v2: np.ndarray
    upd_label: np.ndarray
    queries: List[Query]
    seed: int = 0
    name: str = ""

    @property
    def num_updates(self) -> int:
        return int(self.upd_v1.shape[0])


def build_csr(nv: int, src: np.ndarray, dst: np.ndarray, elab: np.ndarray):
    """Symmetric CSR (offsets, neighbours sorted by id, parallel edge labels)."""
    s = np.concatenate([src, dst]).astype(np.uint64)
    d = np.concatenate([dst, src]).astype(np.uint64)
    e = np.concatenate([elab, elab])
    order = np.argsort((s << np.uint64(32)) | d)
    s, d, e = s[order], d[order], e[order]
    off = np.zeros(nv + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.bincount(s.astype(np.int64), minlength=nv))
    return off, d.astype(np.uint32), e.astype(np.uint32)


def random_walk_query(off, nbr, elab_csr, vlabels, k: int, rng,
                      max_steps: int = 0) -> Optional[Query]:
    """One random-walk query of k vertices, or None if the walk got stuck."""
    nv = off.shape[0] - 1
    deg = off[1:] - off[:-1]
    if max_steps <= 0:
        max_steps = 64 * k
    for _attempt in range(64):
        start = int(rng.integers(0, nv))
        if deg[start] == 0:
            continue
        seen = [start]
        pos = {start: 0}
        cur = start
        for _ in range(max_steps):
            d = int(deg[cur])
            cur = int(nbr[off[cur] + int(rng.integers(0, d))])
            if cur not in pos:
                pos[cur] = len(seen)
                seen.append(cur)
                if len(seen) == k:
                    break
        if len(seen) < k:
            continue
        qs, qd, qe = [], [], []
        for i, v in enumerate(seen):
            lo, hi = off[v], off[v + 1]
            for j in range(lo, hi):
                w = int(nbr[j])
                if w in pos and pos[w] > i:
                    qs.append(i)
                    qd.append(pos[w])
                    qe.append(int(elab_csr[j]))
        return Query(vlabels=np.asarray([vlabels[v] for v in seen], dtype=np.uint32),
                     src=np.asarray(qs, dtype=np.uint32),
                     dst=np.asarray(qd, dtype=np.uint32),
                     elabels=np.asarray(qe, dtype=np.uint32))
    return None


def sample_queries(off, nbr, elab_csr, vlabels, k: int, count: int, seed: int,
                   density: Optional[str] = None, accept=None) -> List[Query]:
    """``count`` random-walk queries of k vertices; density 'sparse'/'dense'/None.

    ``accept`` is an optional predicate applied after the density test; a
    rejected walk is simply resampled, so the result stays seed-determined
    as long as the predicate is deterministic.
    """
    rng = np.random.default_rng(np.random.SeedSequence([seed, k, 0x51555259]))
    out: List[Query] = []
    tries = 0
    while len(out) < count:
        tries += 1
        if tries > 20000:
            raise RuntimeError(f"could not sample {count} {density} queries of size {k}")
        q = random_walk_query(off, nbr, elab_csr, vlabels, k, rng)
        if q is None:
            continue
        if density == "sparse" and q.avg_degree > 3.0:
            continue
        if density == "dense" and q.avg_degree <= 3.0:
            continue
        if accept is not None and not accept(q):
            continue
        out.append(q)
    return out


def insertion_stream(src, dst, elab, frac_stream: float, seed: int):
    """Split unique edges: a seeded permutation, the last ``frac_stream`` become
    insertions in permutation order.  Returns (initial edges, stream edges)."""
    m = src.shape[0]
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x53504C54]))
    perm = rng.permutation(m)
    n_stream = int(round(m * frac_stream))
    init = perm[: m - n_stream]
    strm = perm[m - n_stream:]
    return ((src[init], dst[init], elab[init]), (src[strm], dst[strm], elab[strm]))


def mixed_stream(src, dst, elab, n_updates: int, p_insert: float, seed: int,
                 n_holdout: Optional[int] = None):
    """Hold out edges, then draw ``n_updates`` updates: with probability
    ``p_insert`` insert the next held-out edge, else delete a uniformly chosen
    edge present at that moment.  Returns (initial edges, update arrays)."""
    m = src.shape[0]
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x4D495844]))
    perm = rng.permutation(m)
    if n_holdout is None:
        n_holdout = int(n_updates * p_insert * 1.05) + 16
    n_holdout = min(n_holdout, m // 2)
    hold = perm[m - n_holdout:]
    init = perm[: m - n_holdout]
    coins = rng.random(n_updates) < p_insert
    picks = rng.random(n_updates)
    present = init.tolist()                # edge indices present now
    nxt = 0
    add = np.empty(n_updates, dtype=np.uint8)
    eidx = np.empty(n_updates, dtype=np.int64)
    for t in range(n_updates):
        if (coins[t] and nxt < n_holdout) or not present:
            e = int(hold[nxt]); nxt += 1
            present.append(e)
            add[t] = 1
        else:
            j = int(picks[t] * len(present))
            e = present[j]
            present[j] = present[-1]
            present.pop()
            add[t] = 0
        eidx[t] = e
    return ((src[init], dst[init], elab[init]),
            (add, src[eidx], dst[eidx], elab[eidx]))


def make_workload(scale: int, edge_factor: int, n_vlabels: int, vlabel_dist: str,
                  n_elabels: int, *, stream: str = "insert", frac_stream: float = 0.1,
                  n_updates: int = 0, p_insert: float = 1.0,
                  query_sizes: Sequence[int] = (6,), queries_per_size: int = 5,
                  densities: Sequence[Optional[str]] = (None,), seed: int = 0,
                  zipf_s: float = 1.0, accept=None, name: str = "", progress=None) -> Workload:
    lo, hi = rmat_edges(scale, edge_factor, seed=seed, progress=progress)
    if progress:
        progress(f"{lo.shape[0]} unique edges")
    nv = 1 << scale
    vl = vertex_labels(nv, n_vlabels, vlabel_dist, zipf_s, seed)
    el = edge_labels(lo.shape[0], n_elabels, seed)
    if stream == "insert":
        (isrc, idst, iel), (ssrc, sdst, sel) = insertion_stream(lo, hi, el, frac_stream, seed)
        add = np.ones(ssrc.shape[0], dtype=np.uint8)
    elif stream == "mixed":
        (isrc, idst, iel), (add, ssrc, sdst, sel) = mixed_stream(lo, hi, el, n_updates, p_insert, seed)
    else:
        raise ValueError(stream)
    # orient each update edge by a seeded coin so both orientations occur
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x4F524E54]))
    flip = rng.random(ssrc.shape[0]) < 0.5
    v1 = np.where(flip, sdst, ssrc).astype(np.uint32)
    v2 = np.where(flip, ssrc, sdst).astype(np.uint32)
    if progress:
        progress(f"stream of {ssrc.shape[0]} updates")
    queries: List[Query] = []
    if queries_per_size and query_sizes:
        off, nbr, ecsr = build_csr(nv, isrc, idst, iel)
    for k in query_sizes if queries_per_size else ():
        for di, dens in enumerate(densities):
            queries += sample_queries(off, nbr, ecsr, vl, k, queries_per_size,
                                      seed * 1000003 + di, dens, accept)
    return Workload(nv=nv, vlabels=vl, init_src=isrc.astype(np.uint32),
                    init_dst=idst.astype(np.uint32), init_elab=iel.astype(np.uint32),
                    upd_add=add.astype(np.uint8), upd_v1=v1, upd_v2=v2,
                    upd_label=sel.astype(np.uint32), queries=queries, seed=seed,
                    name=name)


# ---------------------------------------------------------------- text formats
# The reference's formats (ParaCOSM/CSM/graph/graph.cpp:217-285, README.md:42-84):
# graph files hold "t ...", "v <id> <label>" and "e <a> <b> <label>" lines (any
# other leading token is read as an edge); update streams hold "v"/"-v"
# "<id> <label>" and "e"/"-e" "<a> <b> <label>" lines.

def write_graph_file(path, vlabels, src, dst, elab) -> None:
    with open(path, "w") as f:
        f.write(f"t # {len(vlabels)}\n")
        f.writelines(f"v {i} {int(l)}\n" for i, l in enumerate(vlabels))
        f.writelines(f"e {int(a)} {int(b)} {int(l)}\n" for a, b, l in zip(src, dst, elab))


def write_query_file(path, q: Query) -> None:
    write_graph_file(path, q.vlabels, q.src, q.dst, q.elabels)


def write_stream_file(path, add, v1, v2, label) -> None:
    with open(path, "w") as f:
        f.writelines(f"{'e' if a else '-e'} {int(x)} {int(y)} {int(l)}\n"
                     for a, x, y, l in zip(add, v1, v2, label))
