"""Small seeded inputs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

from paracosm_amd.gen import Query


def random_graph(nv, p, n_vl, n_el, rng):
    vl = rng.integers(0, n_vl, size=nv).astype(np.uint32)
    src, dst = [], []
    for a in range(nv):
        for b in range(a + 1, nv):
            if rng.random() < p:
                src.append(a)
                dst.append(b)
    el = rng.integers(0, n_el, size=len(src)).astype(np.uint32)
    return vl, np.asarray(src, np.uint32), np.asarray(dst, np.uint32), el


def random_connected_query(k, extra, n_vl, n_el, rng):
    """A random spanning tree on k vertices plus `extra` random chords."""
    edges = set()
    for v in range(1, k):
        u = int(rng.integers(0, v))
        edges.add((u, v))
    tries = 0
    while extra > 0 and tries < 100:
        tries += 1
        a, b = sorted(int(x) for x in rng.choice(k, 2, replace=False))
        if (a, b) not in edges:
            edges.add((a, b))
            extra -= 1
    edges = sorted(edges)
    return Query(vlabels=rng.integers(0, n_vl, size=k).astype(np.uint32),
                 src=np.asarray([e[0] for e in edges], np.uint32),
                 dst=np.asarray([e[1] for e in edges], np.uint32),
                 elabels=rng.integers(0, n_el, size=len(edges)).astype(np.uint32))


def mixed_stream_small(src, dst, el, nv, n_upd, p_ins, n_el, rng):
    """Split a small edge list into an initial graph and a valid mixed stream
    (inserts of absent edges, deletes of present ones, in both orientations)."""
    m = len(src)
    perm = rng.permutation(m)
    n_init = m // 2
    present = {(int(src[i]), int(dst[i])): int(el[i]) for i in perm[:n_init]}
    absent = [(int(src[i]), int(dst[i]), int(el[i])) for i in perm[n_init:]]
    init = list(present.items())
    ups = []
    for _ in range(n_upd):
        if (rng.random() < p_ins and absent) or not present:
            if not absent:
                break
            a, b, l = absent.pop(int(rng.integers(0, len(absent))))
            present[(a, b)] = l
            ups.append((1, a, b, l))
        else:
            keys = sorted(present)
            a, b = keys[int(rng.integers(0, len(keys)))]
            l = present.pop((a, b))
            absent.append((a, b, l))
            ups.append((0, a, b, l))
    ups = [(t, b, a, l) if rng.random() < 0.5 else (t, a, b, l) for (t, a, b, l) in ups]
    i_src = np.asarray([k[0] for k, _ in init], np.uint32)
    i_dst = np.asarray([k[1] for k, _ in init], np.uint32)
    i_el = np.asarray([v for _, v in init], np.uint32)
    add = np.asarray([u[0] for u in ups], np.uint8)
    v1 = np.asarray([u[1] for u in ups], np.uint32)
    v2 = np.asarray([u[2] for u in ups], np.uint32)
    lab = np.asarray([u[3] for u in ups], np.uint32)
    return (i_src, i_dst, i_el), (add, v1, v2, lab)


def query_triangle(label=0, el=0):
    return Query(vlabels=np.full(3, label, np.uint32), src=np.asarray([0, 0, 1], np.uint32),
                 dst=np.asarray([1, 2, 2], np.uint32), elabels=np.full(3, el, np.uint32))


def query_path2(label=0):
    return Query(vlabels=np.full(3, label, np.uint32), src=np.asarray([0, 1], np.uint32),
                 dst=np.asarray([1, 2], np.uint32), elabels=np.zeros(2, np.uint32))


def query_star(leaves, label=0):
    return Query(vlabels=np.full(leaves + 1, label, np.uint32),
                 src=np.zeros(leaves, np.uint32),
                 dst=np.arange(1, leaves + 1, dtype=np.uint32),
                 elabels=np.zeros(leaves, np.uint32))


def walk_query(vl, src, dst, el, k, rng, return_vertices=False):
    """The subgraph induced by the first k distinct vertices of a random walk
    on the data graph (so the query has at least one embedding), relabelled
    0..k-1."""
    adj = {}
    for a, b in zip(src.tolist(), dst.tolist()):
        adj.setdefault(a, []).append(b)
        adj.setdefault(b, []).append(a)
    starts = sorted(adj)
    cur = starts[int(rng.integers(0, len(starts)))]
    seen = [cur]
    for _ in range(100000):
        if len(seen) == k:
            break
        nb = adj[cur]
        cur = nb[int(rng.integers(0, len(nb)))]
        if cur not in seen:
            seen.append(cur)
    assert len(seen) == k, "the walk's component is smaller than k"
    idx = {v: i for i, v in enumerate(seen)}
    es = [(idx[a], idx[b], l) for a, b, l in zip(src.tolist(), dst.tolist(), el.tolist())
          if a in idx and b in idx]
    q = Query(vlabels=np.asarray(vl[seen], np.uint32),
              src=np.asarray([e[0] for e in es], np.uint32),
              dst=np.asarray([e[1] for e in es], np.uint32),
              elabels=np.asarray([e[2] for e in es], np.uint32))
    return (q, seen) if return_vertices else q


def walk_case(seed, k, nv=150, p=0.05, n_vl=3, n_el=2):
    """A data graph, a k-vertex query induced by a random walk on it, and a
    stream that inserts every edge among the walk's vertices (plus a fifth of
    the others, held out of the initial graph) in random order and then
    deletes half of the inserted edges: the walk's embedding appears during
    the stream and may disappear again."""
    rng = np.random.default_rng(seed)
    vl, s, d, el = random_graph(nv, p, n_vl, n_el, rng)
    q, verts = walk_query(vl, s, d, el, k, rng, return_vertices=True)
    vset = set(verts)
    inq = np.array([a in vset and b in vset for a, b in zip(s.tolist(), d.tolist())])
    other = np.flatnonzero(~inq)
    held = np.concatenate([np.flatnonzero(inq),
                           rng.choice(other, size=len(other) // 5, replace=False)])
    keep = np.setdiff1d(np.arange(len(s)), held)
    ins = rng.permutation(held)
    dels = rng.choice(ins, size=len(ins) // 2, replace=False)
    add = np.concatenate([np.ones(len(ins), np.uint8), np.zeros(len(dels), np.uint8)])
    e = np.concatenate([ins, dels])
    sw = rng.random(len(e)) < 0.5
    v1 = np.where(sw, d[e], s[e]).astype(np.uint32)
    v2 = np.where(sw, s[e], d[e]).astype(np.uint32)
    return vl, (s[keep], d[keep], el[keep]), (add, v1, v2, el[e].astype(np.uint32)), q
