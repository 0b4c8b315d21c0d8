"""parallel_symbi with reference orders on the device: the next query vertex
is chosen during the search by SymBi's ExtendableVertex rule
(Parallel_SymBi/parrallel_symbi.cpp:4231-4262), in k_search_dyn.

On the plain per-update path the device index is exactly the reference's at
each update's search (after the insertion, before the deletion), so the rule
picks the same vertices as the oracle's sequential SymBi and the per-update
work (DCS entries walked) must be equal, not only the counts.  In larger
batches the index describes the batch's union graph, so the order may differ;
the counts and hashes must not.
"""
import numpy as np
import pytest

from tests.helpers import mixed_stream_small, random_connected_query, random_graph, walk_case
from tests.order_rule_cases import CASES

pytestmark = pytest.mark.gpu


def _dyn(q, vl, init, **kw):
    from paracosm_amd.engine import Matching
    m = Matching("parallel_symbi", q, vl, init[0], init[1], init[2], reference_orders=True, **kw)
    m.Preprocessing()
    return m


def _plain(m, ups):
    """Each update through AddEdge / RemoveEdge: counts and work per update."""
    add, v1, v2, lab = ups
    cnt, work = [], []
    for i in range(len(v1)):
        c = m.AddEdge(int(v1[i]), int(v2[i]), int(lab[i])) if add[i] else \
            m.RemoveEdge(int(v1[i]), int(v2[i]))
        cnt.append(c)
        work.append(int(m.update_work(1)[0]))
    return np.asarray(cnt, np.uint64), np.asarray(work, np.uint64)


@pytest.mark.parametrize("name", sorted(CASES))
def test_hand_worked_cases(name):
    q, data, ups, work_want, count_want = CASES[name]()
    vl, s, d, el = data
    with _dyn(q, vl, (s, d, el), want_work=True) as m:
        cnt, work = _plain(m, ups)
    assert list(map(int, work)) == work_want
    assert list(map(int, cnt)) == count_want


def _case(seed, nv, p, n_vl, n_el, n_upd, k, extra):
    rng = np.random.default_rng(seed)
    vl, s, d, el = random_graph(nv, p, n_vl, n_el, rng)
    init, ups = mixed_stream_small(s, d, el, nv, n_upd, 0.65, n_el, rng)
    q = random_connected_query(k, extra, n_vl, n_el, rng)
    return vl, init, ups, q


@pytest.mark.parametrize("seed", range(6))
def test_plain_path_work_equals_oracle(oracle_lib, seed):
    O = oracle_lib
    vl, init, ups, q = _case(900 + seed, nv=40, p=0.35, n_vl=2, n_el=1 + seed % 2,
                             n_upd=80, k=3 + seed, extra=seed % 3)
    work_ref, cnt_ref = O.symbi_work(O.OGraph(vl, *init), O.OGraph.from_query(q), *ups)
    with _dyn(q, vl, init, want_work=True) as m:
        cnt, work = _plain(m, ups)
    np.testing.assert_array_equal(cnt, cnt_ref)
    np.testing.assert_array_equal(work, work_ref)


@pytest.mark.parametrize("k", [5, 8, 12, 20])
def test_batches_match_oracle(oracle_lib, k):
    """Batches of several sizes (union-graph index, chunked seeds) against the
    oracle's counts and embedding-set hashes; k = 12 and 20 use the larger
    LDS layouts of the kernel."""
    from paracosm_amd.engine import make_updates
    O = oracle_lib
    vl, init, ups, q = walk_case(950 + k, k)
    ref = O.run("symbi", O.OGraph(vl, *init), O.OGraph.from_query(q), *ups, want_hash=True)
    assert ref.pos.sum() > 0
    u = make_updates(*ups)
    for batch, chunk in ((len(u), 0), (37, 4)):
        with _dyn(q, vl, init, want_hash=True, max_batch=batch, split_chunk=chunk) as m:
            pos, neg, _, hsh = m.apply_batch(u)
        np.testing.assert_array_equal(pos[:, 0], ref.pos)
        np.testing.assert_array_equal(neg[:, 0], ref.neg)
        np.testing.assert_array_equal(hsh[:, 0], ref.hash)


def test_initial_matching_and_result_cap(oracle_lib):
    """InitialMatching under the rule, and the per-update result cap
    (min(exact, cap))."""
    from paracosm_amd.engine import make_updates
    O = oracle_lib
    vl, init, ups, q = _case(977, nv=50, p=0.35, n_vl=2, n_el=1, n_upd=150, k=5, extra=1)
    ref = O.run("symbi", O.OGraph(vl, *init), O.OGraph.from_query(q), *ups, want_initial=True)
    with _dyn(q, vl, init) as m:
        assert m.InitialMatching() == ref.initial
    cap = 3
    with _dyn(q, vl, init, max_results=cap) as m:
        pos, neg, _, _ = m.apply_batch(make_updates(*ups), partial_ok=True)
    np.testing.assert_array_equal(pos[:, 0], np.minimum(ref.pos, cap))
    np.testing.assert_array_equal(neg[:, 0], np.minimum(ref.neg, cap))
