"""TurboFlux's matching orders on the device: with reference_orders,
parallel_turboflux ranks query vertices by the explicit-path counts of its
spanning-tree index (Parallel_TurboFlux::CountDownwards and
GenerateMatchingOrder, parallel_turboflux.cpp:469-700), computed by k_tf_paths
over the device index.  The orders and path counts read back through
csm_query_orders equal the oracle's, and the counts of a mixed stream searched
in those orders equal the oracle's and the golden fixtures'.
"""
import numpy as np
import pytest

from tests.golden_cases import case_names, load
from tests.helpers import walk_query

pytestmark = pytest.mark.gpu


def _case(seed, nv=800, m=12000, n_vl=4, n_el=2, n_upd=2048, k=6):
    """A random graph (average degree 30, so walk queries pick up chords:
    non-tree query edges), a k-vertex query from a random walk on it, and a
    stream that inserts the last n_upd edges and then deletes every other
    one of them."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, nv, size=3 * m)
    b = rng.integers(0, nv, size=3 * m)
    keep = a != b
    pairs = np.unique(np.stack([np.minimum(a, b), np.maximum(a, b)], 1)[keep], axis=0)
    pairs = pairs[rng.permutation(len(pairs))[:m]]
    src, dst = pairs[:, 0].astype(np.uint32), pairs[:, 1].astype(np.uint32)
    el = rng.integers(0, n_el, size=m).astype(np.uint32)
    vl = rng.integers(0, n_vl, size=nv).astype(np.uint32)
    q = walk_query(vl, src, dst, el, k, rng)
    n0 = m - n_upd
    init = (src[:n0], dst[:n0], el[:n0])
    idx = np.concatenate([np.arange(n0, m), np.arange(n0, m, 2)])
    add = np.concatenate([np.ones(n_upd, np.uint8), np.zeros(len(idx) - n_upd, np.uint8)])
    return vl, init, (add, src[idx], dst[idx], el[idx]), q


def _n_edges(q):
    return len({(min(a, b), max(a, b)) for a, b in zip(q.src.tolist(), q.dst.tolist())})


@pytest.mark.parametrize("seed,k", [(1, 4), (2, 5), (3, 6), (4, 7), (5, 8)])
def test_orders_and_paths_equal_oracle(oracle_lib, seed, k):
    from paracosm_amd.engine import Matching, make_updates
    O = oracle_lib
    vl, init, ups, q = _case(700 + seed, k=k)
    dg, qg = O.OGraph(vl, *init), O.OGraph.from_query(q)
    want_o, want_p = O.turboflux_orders(dg, qg, q.k, _n_edges(q))
    ref = O.run("graphflow", dg, qg, *ups)
    print("seed", seed, "k", k, "paths", want_p.tolist(), "total",
          int(ref.pos.sum() + ref.neg.sum()), flush=True)
    with Matching("parallel_turboflux", q, vl, *init, reference_orders=True) as m:
        m.Preprocessing()
        got_o, got_p = m.query_orders(0)
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
    np.testing.assert_array_equal(got_p, want_p)
    np.testing.assert_array_equal(got_o, want_o)
    np.testing.assert_array_equal(pos[:, 0], ref.pos)
    np.testing.assert_array_equal(neg[:, 0], ref.neg)


@pytest.mark.parametrize("name", case_names())
def test_golden_fixtures_in_turboflux_orders(oracle_lib, name):
    from paracosm_amd.engine import Matching, make_updates
    O = oracle_lib
    _, exp, q, vl, init, ups = load(name)
    want_o, want_p = O.turboflux_orders(O.OGraph(vl, *init), O.OGraph.from_query(q), q.k,
                                        _n_edges(q))
    with Matching("parallel_turboflux", q, vl, *init, reference_orders=True) as m:
        m.Preprocessing()
        got_o, got_p = m.query_orders(0)
        assert m.InitialMatching() == exp["initial_embeddings"]
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
    np.testing.assert_array_equal(got_p, want_p)
    np.testing.assert_array_equal(got_o, want_o)
    np.testing.assert_array_equal(pos[:, 0], np.asarray(exp["positive"], np.uint64))
    np.testing.assert_array_equal(neg[:, 0], np.asarray(exp["negative"], np.uint64))


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_turboflux"])
def test_default_orders_report_no_paths(algo):
    """Without reference_orders the engine's own cost-ranked orders are
    reported: each a permutation of the query vertices, and every query edge
    the first two positions of exactly one; the path counts are zero."""
    from paracosm_amd.engine import Matching
    vl, init, _, q = _case(801, k=6)
    with Matching(algo, q, vl, *init) as m:
        m.Preprocessing()
        orders, paths = m.query_orders(0)
    edges = sorted({(min(a, b), max(a, b)) for a, b in zip(q.src.tolist(), q.dst.tolist())})
    assert orders.shape == (len(edges), q.k)
    assert not paths.any()
    for o in orders.tolist():
        assert sorted(o) == list(range(q.k))
    assert sorted(tuple(sorted(o[:2])) for o in orders.tolist()) == edges
