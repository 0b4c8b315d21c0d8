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
