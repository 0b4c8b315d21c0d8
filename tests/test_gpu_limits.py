"""Result cap (--max-results) and time limit (--time-limit) on the device.

The reference stops an update's search once it has max_num_results_ results
(ParaCOSM/CSM/matching/Parallel_GraphFlow/parallel_graphflow.cpp:386, 2101)
and every search once the watchdog has set reach_time_limit
(utils/utils.h:19-36; parallel_graphflow.cpp:288, 386-387).  Here the cap
gives min(exact count, cap) per update and query, and the time limit stops
the search, keeps every update applied, and flags the updates whose counts
are partial.
"""
import time

import numpy as np
import pytest

from tests.helpers import mixed_stream_small, random_connected_query, random_graph

pytestmark = pytest.mark.gpu


def _case(seed, nv=45, p=0.3, n_vl=2, n_el=1, n_upd=150, k=5, extra=2):
    rng = np.random.default_rng(seed)
    vl, s, d, el = random_graph(nv, p, n_vl, n_el, rng)
    init, ups = mixed_stream_small(s, d, el, nv, n_upd, 0.7, n_el, rng)
    q = random_connected_query(k, extra, n_vl, n_el, rng)
    return vl, init, ups, q


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
@pytest.mark.parametrize("cap", [1, 7, 40])
def test_result_cap_is_min_of_exact_and_cap(oracle_lib, algo, cap):
    from paracosm_amd.engine import USTAT_CAPPED, Matching, make_updates
    vl, init, ups, q = _case(31)
    ref = oracle_lib.run("graphflow", oracle_lib.OGraph(vl, *init),
                         oracle_lib.OGraph.from_query(q), *ups)
    exact = (ref.pos + ref.neg).astype(np.uint64)
    print("cap", algo, cap, "max exact", int(exact.max()), flush=True)
    with Matching(algo, q, vl, *init, max_results=cap) as m:
        m.Preprocessing()
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
        st = m.update_status(len(exact))
    got = (pos + neg)[:, 0]
    np.testing.assert_array_equal(got, np.minimum(exact, cap))
    np.testing.assert_array_equal((st & USTAT_CAPPED) != 0, exact >= cap)
    assert (exact > cap).any() and (exact < cap).any() or cap == 1


def test_result_cap_applies_to_initial_matching(oracle_lib):
    from paracosm_amd.engine import Matching
    vl, init, ups, q = _case(32, k=4, extra=1)
    full = oracle_lib.brute_count(oracle_lib.OGraph(vl, *init), oracle_lib.OGraph.from_query(q))
    assert full > 10
    with Matching("parallel_graphflow", q, vl, *init, max_results=10) as m:
        m.Preprocessing()
        assert m.InitialMatching() == 10
    with Matching("parallel_graphflow", q, vl, *init, max_results=full + 5) as m:
        m.Preprocessing()
        assert m.InitialMatching() == full


def _heavy():
    """A dense graph (400 vertices, average degree about 80) with one label
    and a 7-vertex cycle query: each inserted edge closes billions of cycles,
    seconds of search for 32 insertions."""
    rng = np.random.default_rng(5)
    nv = 400
    vl, s, d, el = random_graph(nv, 0.4, 1, 1, rng)
    init, ups = mixed_stream_small(s, d, el, nv, 64, 1.0, 1, rng)
    from paracosm_amd.gen import Query
    k = 7
    q = Query(vlabels=np.zeros(k, np.uint32), src=np.arange(k, dtype=np.uint32),
              dst=np.asarray([(i + 1) % k for i in range(k)], np.uint32),
              elabels=np.zeros(k, np.uint32))
    return vl, init, ups, q


def test_time_limit_returns_promptly_with_flags():
    """A tiny limit on a heavy query: the call returns CSM_E_TIME_LIMIT well
    within a second or two, flags the stopped updates, and leaves the engine
    usable."""
    from paracosm_amd.engine import CSM_E_TIME_LIMIT, USTAT_TIME_LIMIT, CsmError, Matching, make_updates
    vl, init, ups, q = _heavy()
    u = make_updates(*ups)
    with Matching("parallel_graphflow", q, vl, *init, time_limit_s=0.15) as m:
        m.Preprocessing()
        t0 = time.perf_counter()
        pos, _, _, _ = m.apply_batch(u[:32], partial_ok=True)
        dt = time.perf_counter() - t0
        st = m.update_status(32)
        print("time limit: returned after", round(dt, 3), "s; flagged", int((st & 1).sum()),
              "counted", int(pos.sum()), flush=True)
        assert dt < 3.0
        assert (st & USTAT_TIME_LIMIT).any()
        assert m.query_status().tolist() == [2]   # the time limit dropped its work
        with pytest.raises(CsmError) as ei:
            m.apply_batch(u[32:40])
        assert ei.value.code == CSM_E_TIME_LIMIT
        # the engine goes on: without a limit the same kind of call completes
        m.set_time_limit(0)
        _, _, _, _ = m.apply_batch(u[40:41])
        assert not m.update_status(1).any()


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_time_limit_keeps_graph_exact(oracle_lib, algo):
    """Updates not flagged keep their exact counts, flagged ones are lower
    bounds, and the graph and index are fully updated: a later call without a
    limit is exact again."""
    from paracosm_amd.engine import USTAT_TIME_LIMIT, Matching, make_updates
    vl, init, ups, q = _case(33, nv=60, p=0.4, k=6, extra=3, n_upd=200)
    ref = oracle_lib.run("graphflow", oracle_lib.OGraph(vl, *init),
                         oracle_lib.OGraph.from_query(q), *ups)
    exact = (ref.pos + ref.neg).astype(np.uint64)
    u = make_updates(*ups)
    with Matching(algo, q, vl, *init, time_limit_s=1e-6) as m:
        m.Preprocessing()
        pos, neg, _, _ = m.apply_batch(u[:120], partial_ok=True)
        st = m.update_status(120)
        got = (pos + neg)[:, 0]
        flagged = (st & USTAT_TIME_LIMIT) != 0
        print(algo, "flagged", int(flagged.sum()), "of 120", flush=True)
        assert flagged.any()
        np.testing.assert_array_equal(got[~flagged], exact[:120][~flagged])
        assert (got[flagged] <= exact[:120][flagged]).all()
        m.set_time_limit(0)
        pos, neg, _, _ = m.apply_batch(u[120:])
        np.testing.assert_array_equal((pos + neg)[:, 0], exact[120:])
