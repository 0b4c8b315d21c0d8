"""Loading the golden fixtures under tests/golden/ (one directory per case)."""
from __future__ import annotations

import json
import os

import numpy as np

from paracosm_amd.gen import Query

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M64 = (1 << 64) - 1


def case_names():
    return sorted(d for d in os.listdir(GOLDEN)
                  if os.path.isfile(os.path.join(GOLDEN, d, "expected.json")))


def read_graph(path):
    """``t``/``v``/``e`` lines -> vertex labels and edge arrays (the fixtures
    list every vertex once, in id order)."""
    vl, src, dst, el = [], [], [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0] == "t":
                continue
            if tok[0] == "v":
                assert int(tok[1]) == len(vl)
                vl.append(int(tok[2]))
            else:
                src.append(int(tok[1]))
                dst.append(int(tok[2]))
                el.append(int(tok[3]))
    u32 = lambda a: np.asarray(a, np.uint32)
    return u32(vl), u32(src), u32(dst), u32(el)


def read_stream(path):
    add, v1, v2, lab = [], [], [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            add.append(1 if tok[0] == "e" else 0)
            v1.append(int(tok[1]))
            v2.append(int(tok[2]))
            lab.append(int(tok[3]))
    return (np.asarray(add, np.uint8), np.asarray(v1, np.uint32), np.asarray(v2, np.uint32),
            np.asarray(lab, np.uint32))


def load(name):
    d = os.path.join(GOLDEN, name)
    with open(os.path.join(d, "expected.json")) as f:
        exp = json.load(f)
    qvl, qs, qd, qel = read_graph(os.path.join(d, "query.graph"))
    q = Query(vlabels=qvl, src=qs, dst=qd, elabels=qel)
    vl, s, t, el = read_graph(os.path.join(d, "data.graph"))
    ups = read_stream(os.path.join(d, "stream.txt"))
    return d, exp, q, vl, (s, t, el), ups


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def set_hash(embs):
    """The engine's embedding-set hash: the sum over embeddings of a mix64
    chain over the data vertex of each query vertex."""
    tot = 0
    for m in embs:
        h = 0
        for v in m:
            h = _mix64((h + int(v) + 1) & M64)
        tot = (tot + h) & M64
    return tot
