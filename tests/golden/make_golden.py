"""Writes the golden fixtures: tiny cases in the reference's text formats with
their expected outputs, found by complete enumeration.

Each case is a directory holding
  query.graph / data.graph  -- ``t``/``v id label``/``e a b label`` lines, the
                               grammar of Graph::LoadFromFile
                               (ParaCOSM/CSM/graph/graph.cpp:217-249);
  stream.txt                -- ``e a b l`` / ``-e a b l`` lines, the grammar of
                               Graph::LoadUpdateStream (graph.cpp:251-285);
  expected.json             -- per update, the embeddings the update creates
                               (insertions) or destroys (deletions), their
                               numbers, the totals, the initial count, and
                               the SHA-256 of the three input files.

An embedding is an injective map f from query vertices to data vertices that
keeps vertex labels and maps every query edge onto a data edge with the same
edge label; embeddings are maps, so automorphic images count separately (the
reference's count, e.g. a triangle of one label has 6).  The expected values
come from enumerating every embedding of the whole graph before and after
each update and taking the set difference: nothing here shares code with the
oracle or the engine, and no search order or pruning is involved.

Run ``python tests/golden/make_golden.py`` to rewrite the fixtures; the
output is deterministic (seeded ``random.Random``).
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))

# name, seed, data vertices, vertex labels, edge labels, edge probability,
# query edges (u, w, edge label) with query vertex labels, updates, insert share
CASES = [
    ("triangle_two_labels", 17, 12, 2, 1, 0.6,
     ([0, 0, 1], [(0, 1, 0), (1, 2, 0), (0, 2, 0)]), 40, 0.6),
    ("square_with_chord_two_edge_labels", 16, 11, 1, 2, 0.9,
     ([0, 0, 0, 0], [(0, 1, 0), (1, 2, 1), (2, 3, 0), (3, 0, 1), (0, 2, 0)]), 40, 0.6),
    ("path_of_five_one_label", 13, 11, 1, 1, 0.25,
     ([0, 0, 0, 0, 0], [(0, 1, 0), (1, 2, 0), (2, 3, 0), (3, 4, 0)]), 30, 0.55),
    ("star_three_leaves", 14, 13, 2, 1, 0.20,
     ([0, 1, 1, 1], [(0, 1, 0), (0, 2, 0), (0, 3, 0)]), 35, 0.6),
    ("five_vertices_tree_and_chord", 15, 10, 1, 2, 0.7,
     ([0, 0, 0, 0, 0], [(0, 1, 0), (1, 2, 1), (1, 3, 0), (3, 4, 1), (2, 4, 0)]), 40, 0.6),
]


def embeddings(qvl, qedges, vl, edges):
    """Every embedding, as a tuple (f(0), ..., f(k-1)); edges: {(a, b): l}, a < b."""
    k = len(qvl)
    by_label = {}
    for v, l in enumerate(vl):
        by_label.setdefault(l, []).append(v)
    out = []
    for f in itertools.product(*(by_label.get(l, []) for l in qvl)):
        if len(set(f)) != k:
            continue
        if all(edges.get((min(f[u], f[w]), max(f[u], f[w]))) == l for u, w, l in qedges):
            out.append(tuple(f))
    return set(out)


def make_case(name, seed, nv, n_vl, n_el, p, query, n_upd, p_ins, out_dir=HERE):
    rng = random.Random(seed)
    vl = [rng.randrange(n_vl) for _ in range(nv)]
    all_edges = [(a, b) for a in range(nv) for b in range(a + 1, nv) if rng.random() < p]
    rng.shuffle(all_edges)
    labels = {e: rng.randrange(n_el) for e in all_edges}
    present = dict((e, labels[e]) for e in all_edges[: len(all_edges) // 2])
    absent = [e for e in all_edges[len(all_edges) // 2:]]
    initial = dict(present)
    ups = []
    for _ in range(n_upd):
        if (rng.random() < p_ins and absent) or not present:
            e = absent.pop(rng.randrange(len(absent)))
            ups.append((1, e, labels[e]))
            present[e] = labels[e]
        else:
            e = sorted(present)[rng.randrange(len(present))]
            ups.append((0, e, present.pop(e)))
            absent.append(e)
    # orientation of each update line, as a stream would list it
    ups = [(t, (e[1], e[0]) if rng.random() < 0.5 else e, l) for t, e, l in ups]

    qvl, qedges = query
    g = dict(initial)
    before = embeddings(qvl, qedges, vl, g)
    n_initial = len(before)
    pos, neg, changed = [], [], []
    for i, (t, (a, b), l) in enumerate(ups):
        key = (min(a, b), max(a, b))
        if t:
            g[key] = l
        else:
            del g[key]
        after = embeddings(qvl, qedges, vl, g)
        diff = sorted(after - before) if t else sorted(before - after)
        pos.append(len(diff) if t else 0)
        neg.append(0 if t else len(diff))
        changed.append([list(m) for m in diff])
        before = after

    d = os.path.join(out_dir, name)
    os.makedirs(d, exist_ok=True)
    files = {
        "query.graph": "t # 0\n" + "".join(f"v {i} {l}\n" for i, l in enumerate(qvl))
        + "".join(f"e {u} {w} {l}\n" for u, w, l in qedges),
        "data.graph": "t # 0\n" + "".join(f"v {i} {l}\n" for i, l in enumerate(vl))
        + "".join(f"e {a} {b} {l}\n" for (a, b), l in sorted(initial.items())),
        "stream.txt": "".join(f"{'e' if t else '-e'} {a} {b} {l}\n" for t, (a, b), l in ups),
    }
    sha = {}
    for fn, text in files.items():
        with open(os.path.join(d, fn), "w") as f:
            f.write(text)
        sha[fn] = hashlib.sha256(text.encode()).hexdigest()
    expected = {
        "case": name, "seed": seed, "sha256": sha,
        "query_vertices": len(qvl), "data_vertices": nv, "initial_edges": len(initial),
        "updates": len(ups), "initial_embeddings": n_initial,
        "positive": pos, "negative": neg,
        "total_positive": sum(pos), "total_negative": sum(neg),
        "embeddings": changed,
    }
    with open(os.path.join(d, "expected.json"), "w") as f:
        json.dump(expected, f, indent=0, separators=(",", ":"))
        f.write("\n")
    return expected


def main():
    for c in CASES:
        e = make_case(*c)
        print(f"{e['case']}: {e['updates']} updates, initial {e['initial_embeddings']}, "
              f"+{e['total_positive']} -{e['total_negative']}")


if __name__ == "__main__":
    main()
