"""Hand-worked cases of SymBi's search-time order rule (the ExtendableVertex
rule, SymBi/symbi.cpp:474-506 and Parallel_SymBi/parrallel_symbi.cpp:4231-4262).

Each case is a query, a data graph and a stream that inserts one edge and
deletes it again.  `work` is the number of DCS entries the search walks for
each update (the candidates of the vertex the rule picks, valid or not, at
every node of the search tree) and `count` its embeddings, both worked out by
hand in the comments below.  Every case is built so that a plausible mistake
in the rule changes the work: taking the largest E or the lowest id instead of
the least E, ignoring the preference for non-isolated vertices, ranking by
the DCS list's size instead of its d2-valid entries (n2), breaking ties the
other way, or walking the candidates of the wrong matched neighbour.

All edge labels are 0; every vertex label is distinct within a query.
"""
from __future__ import annotations

import numpy as np

from paracosm_amd.gen import Query


def _q(vlabels, edges):
    return Query(vlabels=np.asarray(vlabels, np.uint32),
                 src=np.asarray([a for a, _ in edges], np.uint32),
                 dst=np.asarray([b for _, b in edges], np.uint32),
                 elabels=np.zeros(len(edges), np.uint32))


def _data(vlabels, edges):
    return (np.asarray(vlabels, np.uint32), np.asarray([a for a, _ in edges], np.uint32),
            np.asarray([b for _, b in edges], np.uint32), np.zeros(len(edges), np.uint32))


def _insert_delete(a, b):
    """Insert edge (a, b), then delete it: the same search runs twice."""
    return (np.array([1, 0], np.uint8), np.array([a, a], np.uint32), np.array([b, b], np.uint32),
            np.zeros(2, np.uint32))


def star():
    """Star query: centre 0 (label 1), leaves 1..4 (labels 2..5).  Data: centre
    x = 0 with 5 neighbours of label 3, 2 of label 4 and 3 of label 5; vertex 1
    (label 2) is joined to x by the update.  After the seed (0, 1) the three
    other leaves are isolated with E = 5, 2, 3 (every leaf candidate is
    d2-valid), so the rule takes them in the order E = 2, 3, 5:
      work = 2 + 2*3 + 2*3*5 = 38, count = 2*3*5 = 30.
    Query-id order (E = 5, 2, 3) would walk 5 + 10 + 30 = 45 and the largest E
    first 5 + 15 + 30 = 50."""
    q = _q([1, 2, 3, 4, 5], [(0, 1), (0, 2), (0, 3), (0, 4)])
    labels = [1, 2] + [3] * 5 + [4] * 2 + [5] * 3
    edges = [(0, v) for v in range(2, 12)]
    return q, _data(labels, edges), _insert_delete(0, 1), [38, 38], [30, 30]


def two_arms():
    """Query: b=0 (1) - c=1 (2) the seed edge; arm p=2 (3) - p'=4 (5) at b and
    arm r=3 (4) - r'=5 (6) at c.  Data: xb = 0, xc = 1 (joined by the update);
    xb's label-3 neighbours are 2, 3 (each with its own label-5 neighbour 7, 8)
    and 4, 5, 6 (without one: not d2-valid); xc's label-4 neighbours are 9, 10
    (with label-6 neighbours 11, 12).
    After the seed p and r are both non-isolated with E = 2 (p's DCS has 5
    entries, 2 valid); the tie goes to the lower id, p: walk 5, 2 candidates.
    Then r (non-isolated, E = 2) comes before p' (isolated, E = 1): walk 2, 2
    candidates.  Then p' and r' (both isolated, E = 1; p' by id): walk 1 each.
      work = 5 + 2 * (2 + 2 * (1 + 1)) = 17, count = 2 * 2 = 4.
    Ties to the higher id, or ranking by DCS size, take r first: 2 + 2 * (5 +
    2 * 2) = 20; ignoring isolation takes p' before r: 5 + 2 * (1 + 2 + 2) = 15."""
    q = _q([1, 2, 3, 4, 5, 6], [(0, 1), (0, 2), (1, 3), (2, 4), (3, 5)])
    labels = [1, 2, 3, 3, 3, 3, 3, 5, 5, 4, 4, 6, 6]
    edges = [(0, 2), (0, 3), (0, 4), (0, 5), (0, 6), (2, 7), (3, 8), (1, 9), (1, 10),
             (9, 11), (10, 12)]
    return q, _data(labels, edges), _insert_delete(0, 1), [17, 17], [4, 4]


def triangle(extra_at_c: bool):
    """Triangle query b=0 (1), c=1 (2), s=2 (3).  Data: xb = 0, xc = 1 (joined by
    the update), vertex 2 (label 3) adjacent to both, and three more label-3
    vertices 3, 4, 5 adjacent to xb only (extra_at_c: to xc only).  Only vertex
    2 is d2-valid for s, so s has E = 1 from either neighbour and u_min is the
    neighbour whose counter came first: the seed's DAG parent.
    The query edge (b, c) is the rarest in the data graph before the update (it
    has no copy), its endpoint labels are equally frequent and their query
    degrees equal, so the DAG root is c (the later endpoint) and c is b's
    parent: E(s) is set from c first.  s's candidates are then xc's label-3
    neighbours: work = 1 (extra_at_c: 4), count = 1.  Taking u_min from b (the
    lower id) would give 4 (extra_at_c: 1)."""
    q = _q([1, 2, 3], [(0, 1), (0, 2), (1, 2)])
    hub = 1 if extra_at_c else 0
    labels = [1, 2, 3, 3, 3, 3]
    edges = [(0, 2), (1, 2)] + [(hub, v) for v in (3, 4, 5)]
    w = 4 if extra_at_c else 1
    return q, _data(labels, edges), _insert_delete(0, 1), [w, w], [1, 1]


CASES = {
    "star": star,
    "two_arms": two_arms,
    "triangle_extra_at_b": lambda: triangle(False),
    "triangle_extra_at_c": lambda: triangle(True),
}
