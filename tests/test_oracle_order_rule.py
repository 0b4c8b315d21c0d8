"""Pins the oracle's restatement of SymBi's search-time order rule
(ExtendableVertex, SymBi/symbi.cpp:474-506; the same rule in
Parallel_SymBi/parrallel_symbi.cpp:4231-4262) and its work count.

The per-update work (DCS entries walked at every node of the search tree) is
checked against hand-worked cases whose values follow from the rule alone;
each case tells apart the rule from a plausible mistake (largest E, query-id
order, no preference for non-isolated vertices, ties to the higher id, ranking
by DCS size, u_min from the other matched neighbour).  The counts are checked
against brute force, so the rule is known to order an exact search.
"""
import numpy as np
import pytest

from tests.helpers import mixed_stream_small, random_connected_query, random_graph
from tests.order_rule_cases import CASES


@pytest.mark.parametrize("name", sorted(CASES))
def test_hand_worked_work_and_counts(oracle_lib, name):
    O = oracle_lib
    q, data, ups, work_want, count_want = CASES[name]()
    work, cnt = O.symbi_work(O.OGraph(*data), O.OGraph.from_query(q), *ups)
    assert list(map(int, work)) == work_want
    assert list(map(int, cnt)) == count_want


@pytest.mark.parametrize("seed", range(4))
def test_counts_equal_the_symbi_run(oracle_lib, seed):
    """symbi_work runs the same search as the sequential SymBi run: the same
    counts, and no work for updates that no seed can grow into a match."""
    O = oracle_lib
    rng = np.random.default_rng(500 + seed)
    vl, s, d, el = random_graph(26, 0.3, 2, 2, rng)
    init, ups = mixed_stream_small(s, d, el, 26, 60, 0.6, 2, rng)
    q = random_connected_query(3 + seed, seed % 2, 2, 2, rng)
    work, cnt = O.symbi_work(O.OGraph(vl, *init), O.OGraph.from_query(q), *ups)
    ref = O.run("symbi", O.OGraph(vl, *init), O.OGraph.from_query(q), *ups)
    np.testing.assert_array_equal(cnt, ref.pos + ref.neg)
    assert (work >= cnt).all() or q.src.size == 1
    # brute force on the final state of a prefix: pos - neg telescopes
    add, v1, v2, lab = ups
    n = 20
    g0 = O.OGraph(vl, *init)
    src, dst, el2 = list(init[0]), list(init[1]), list(init[2])
    edges = {(int(a), int(b)): int(l) for a, b, l in zip(src, dst, el2)}
    for i in range(n):
        a, b = int(v1[i]), int(v2[i])
        key = (a, b) if (a, b) in edges or add[i] else (b, a)
        if add[i]:
            edges[(a, b)] = int(lab[i])
        else:
            edges.pop(key, None)
    g1 = O.OGraph(vl, np.array([e[0] for e in edges], np.uint32),
                  np.array([e[1] for e in edges], np.uint32),
                  np.array(list(edges.values()), np.uint32))
    qg = O.OGraph.from_query(q)
    delta = int(O.brute_count(g1, qg)) - int(O.brute_count(g0, qg))
    signed = [int(cnt[i]) if add[i] else -int(cnt[i]) for i in range(n)]
    assert sum(signed) == delta
