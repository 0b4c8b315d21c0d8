"""Serving loop: every request's results are the same alone and in the mix,
whatever batches continuous batching forms; HBM admission queues work."""
import numpy as np
import pytest

from tests.helpers import mixed_stream_small, random_connected_query, random_graph

pytestmark = pytest.mark.gpu


def _setup(seed=12, n_upd=600):
    rng = np.random.default_rng(seed)
    nv = 70
    vl, s, d, el = random_graph(nv, 0.3, 3, 2, rng)
    init, ups = mixed_stream_small(s, d, el, nv, n_upd, 0.7, 2, rng)
    qs = [random_connected_query(k, 1, 3, 2, rng) for k in (4, 5, 6)]
    return vl, init, ups, qs


def _engine(vl, init, qs, algo="parallel_graphflow"):
    from paracosm_amd.engine import Matching
    m = Matching(algo, qs[0], vl, *init, max_batch=4096)
    for q in qs[1:]:
        m.register_query(q)
    m.Preprocessing()
    return m


def _expected(O, vl, init, ups, qs):
    cols = []
    for q in qs:
        r = O.run("graphflow", O.OGraph(vl, *init), O.OGraph.from_query(q), *ups)
        cols.append(r.pos + r.neg)
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("offered", [2e3, 5e4, 1e9])
@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_results_same_alone_and_in_the_mix(oracle_lib, offered, algo):
    """Low load: mostly one request per batch.  High load: many requests per
    batch.  Every update's counts equal the one-at-a-time oracle."""
    from paracosm_amd.engine import make_updates
    from paracosm_amd.serve import Server, poisson_requests
    vl, init, ups, qs = _setup()
    want = _expected(oracle_lib, vl, init, ups, qs)
    m = _engine(vl, init, qs, algo)
    reqs = poisson_requests(make_updates(*ups), offered, seed=1, max_request=64)
    st = Server(m, max_batch=4096).serve(reqs)
    got = np.concatenate([r.counts for r in reqs])
    m.close()
    print("serve", algo, offered, st, flush=True)
    np.testing.assert_array_equal(got, want)
    if offered >= 1e9:
        assert st.mean_batch > 64      # requests were mixed into shared batches


def test_hbm_budget_queues_requests(oracle_lib):
    from paracosm_amd.engine import make_updates
    from paracosm_amd.serve import Server, poisson_requests
    vl, init, ups, qs = _setup(seed=13, n_upd=400)
    want = _expected(oracle_lib, vl, init, ups, qs)
    m = _engine(vl, init, qs)
    reqs = poisson_requests(make_updates(*ups), 1e9, seed=2, max_request=32)
    srv = Server(m, max_batch=4096)
    srv.hbm_budget = m.device_bytes() + srv.working_set(100)
    st = srv.serve(reqs)
    got = np.concatenate([r.counts for r in reqs])
    np.testing.assert_array_equal(got, want)
    assert st.refused_for_memory > 0 and st.mean_batch <= 100
    from paracosm_amd.serve import Request
    srv.hbm_budget = m.device_bytes() + srv.working_set(1)
    with pytest.raises(MemoryError):
        srv.serve([Request(arrival=0.0, updates=make_updates(*ups)[:50])])
    m.close()
