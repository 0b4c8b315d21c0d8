"""GPU parity: the HIP engine through the C-ABI against the CPU oracle.

Bit-exact per update and per query (counts and embedding-set hashes) on
seeded inputs small enough for the oracle to finish in seconds.
"""
import numpy as np
import pytest

from tests.helpers import (mixed_stream_small, query_path2, query_star, query_triangle,
                           random_connected_query, random_graph, walk_case)

pytestmark = pytest.mark.gpu


def _engine(algo, q, vl, init, **kw):
    from paracosm_amd.engine import Matching
    m = Matching(algo, q, vl, init[0], init[1], init[2], **kw)
    m.Preprocessing()
    return m


def _oracle(O, q, vl, init, ups, algo="graphflow", **kw):
    return O.run(algo, O.OGraph(vl, *init), O.OGraph.from_query(q), *ups, **kw)


def _case(seed, nv=40, p=0.25, n_vl=3, n_el=2, n_upd=120, k=None, extra=None):
    rng = np.random.default_rng(seed)
    vl, s, d, el = random_graph(nv, p, n_vl, n_el, rng)
    init, ups = mixed_stream_small(s, d, el, nv, n_upd, 0.65, n_el, rng)
    k = k or int(3 + seed % 4)
    extra = seed % 3 if extra is None else extra
    q = random_connected_query(k, extra, n_vl, n_el, rng)
    return vl, init, ups, q


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"])
@pytest.mark.parametrize("seed", range(4))
def test_batch_matches_oracle(oracle_lib, algo, seed):
    vl, init, ups, q = _case(seed)
    ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
    from paracosm_amd.engine import USTAT_CERTIFIED, make_updates
    with _engine(algo, q, vl, init, want_hash=True) as m:
        pos, neg, safe, hsh = m.apply_batch(make_updates(*ups))
        st = m.update_status(len(ups[0]))
    print("case", algo, seed, "k", q.k, "total", int(ref.pos.sum() + ref.neg.sum()),
          "certified", int(((st & USTAT_CERTIFIED) != 0).sum()), flush=True)
    np.testing.assert_array_equal(pos[:, 0], ref.pos)
    np.testing.assert_array_equal(neg[:, 0], ref.neg)
    np.testing.assert_array_equal(hsh[:, 0], ref.hash)
    # Classify (the label filter) equals the oracle's per update; the
    # neighbourhood certificate is sound (certified updates change nothing)
    # and applies only past the label filter
    add, v1, v2, lab = ups
    qg = oracle_lib.OGraph.from_query(q)
    want_safe = [oracle_lib.classify(qg, int(vl[a]), int(vl[b]), int(c))
                 for a, b, c in zip(v1, v2, lab)]
    np.testing.assert_array_equal(safe.astype(bool), np.asarray(want_safe))
    cert = (st & USTAT_CERTIFIED) != 0
    assert (ref.pos[cert] + ref.neg[cert] == 0).all()
    assert not (cert & safe.astype(bool)).any()


@pytest.mark.parametrize("batch", [1, 5, 33])
def test_batch_sizes_agree(oracle_lib, batch):
    vl, init, ups, q = _case(11, nv=50, p=0.3, n_upd=150, k=5, extra=2)
    ref = _oracle(oracle_lib, q, vl, init, ups)
    from paracosm_amd.engine import run_stream
    with _engine("parallel_graphflow", q, vl, init) as m:
        pos, neg = run_stream(m, *ups, batch=batch)
    pos = pos if pos.ndim == 1 else pos[:, 0]
    neg = neg if neg.ndim == 1 else neg[:, 0]
    np.testing.assert_array_equal(pos, ref.pos)
    np.testing.assert_array_equal(neg, ref.neg)


def test_plain_per_update_path(oracle_lib):
    vl, init, ups, q = _case(5, nv=30, p=0.3, n_upd=60, k=4, extra=1)
    ref = _oracle(oracle_lib, q, vl, init, ups)
    with _engine("parallel_graphflow", q, vl, init) as m:
        add, v1, v2, lab = ups
        for i in range(len(v1)):
            if add[i]:
                assert m.AddEdge(int(v1[i]), int(v2[i]), int(lab[i])) == ref.pos[i], i
            else:
                assert m.RemoveEdge(int(v1[i]), int(v2[i])) == ref.neg[i], i
        init_n, p, n = m.totals()
        assert p == int(ref.pos.sum()) and n == int(ref.neg.sum())


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_initial_matching(oracle_lib, algo):
    O = oracle_lib
    for seed in range(3):
        vl, init, ups, q = _case(20 + seed, nv=35, p=0.3, k=4 + seed % 2)
        want = O.brute_count(O.OGraph(vl, *init), O.OGraph.from_query(q))
        with _engine(algo, q, vl, init) as m:
            assert m.InitialMatching() == want


def test_closed_forms_on_complete_graph():
    """Triangle query on a clique built up by insertions: pos = 6*common."""
    from paracosm_amd.engine import make_updates
    nv = 24
    vl = np.zeros(nv, np.uint32)
    pairs = [(a, b) for a in range(nv) for b in range(a + 1, nv)]
    rng = np.random.default_rng(0)
    rng.shuffle(pairs)
    ups = make_updates(np.ones(len(pairs), np.uint8), [p[0] for p in pairs],
                       [p[1] for p in pairs], np.zeros(len(pairs), np.uint32))
    empty = (np.zeros(0, np.uint32),) * 3
    with _engine("parallel_graphflow", query_triangle(), vl, empty) as m:
        pos, _, _, _ = m.apply_batch(ups)
    adj = [set() for _ in range(nv)]
    for i, (a, b) in enumerate(pairs):
        adj[a].add(b); adj[b].add(a)
        assert pos[i, 0] == 6 * len(adj[a] & adj[b]), i


def test_churn_within_one_batch(oracle_lib):
    """The same edges deleted and re-inserted many times inside one batch:
    each update must see exactly the graph before it."""
    rng = np.random.default_rng(3)
    nv = 20
    vl, s, d, el = random_graph(nv, 0.5, 2, 1, rng)
    init = (s, d, el)
    hot = [(int(s[i]), int(d[i]), int(el[i])) for i in rng.choice(len(s), 6, replace=False)]
    add, v1, v2, lab = [], [], [], []
    present = {(a, b): True for a, b, _ in hot}
    for _ in range(200):
        a, b, l = hot[int(rng.integers(0, len(hot)))]
        if present[(a, b)]:
            add.append(0); present[(a, b)] = False
        else:
            add.append(1); present[(a, b)] = True
        if rng.random() < 0.5:
            a, b = b, a
        v1.append(a); v2.append(b); lab.append(l)
    ups = (np.asarray(add, np.uint8), np.asarray(v1, np.uint32), np.asarray(v2, np.uint32),
           np.asarray(lab, np.uint32))
    q = query_star(3)
    q.vlabels[:] = 0
    vl[:] = 0
    ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
    from paracosm_amd.engine import make_updates
    with _engine("parallel_graphflow", q, vl, init, want_hash=True) as m:
        pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
    np.testing.assert_array_equal(pos[:, 0], ref.pos)
    np.testing.assert_array_equal(neg[:, 0], ref.neg)
    np.testing.assert_array_equal(hsh[:, 0], ref.hash)


def test_duplicate_insert_follows_reference(oracle_lib):
    """Inserting an edge that exists leaves the graph unchanged but is still
    searched (graph.cpp:63-64, graphflow.cpp:296-352)."""
    vl = np.zeros(6, np.uint32)
    init = (np.asarray([0, 1, 2, 0], np.uint32), np.asarray([1, 2, 3, 2], np.uint32),
            np.zeros(4, np.uint32))
    ups = (np.asarray([1, 1, 0, 1, 1], np.uint8), np.asarray([0, 1, 0, 2, 0], np.uint32),
           np.asarray([1, 0, 1, 3, 1], np.uint32), np.zeros(5, np.uint32))
    q = query_path2()
    ref = _oracle(oracle_lib, q, vl, init, ups)
    from paracosm_amd.engine import make_updates
    with _engine("parallel_graphflow", q, vl, init) as m:
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
    np.testing.assert_array_equal(pos[:, 0], ref.pos)
    np.testing.assert_array_equal(neg[:, 0], ref.neg)


def test_delete_missing_edge_is_an_error():
    from paracosm_amd.engine import CSM_E_DELETE_MISSING, CsmError, make_updates
    vl = np.zeros(5, np.uint32)
    init = (np.asarray([0, 1], np.uint32), np.asarray([1, 2], np.uint32), np.zeros(2, np.uint32))
    with _engine("parallel_graphflow", query_path2(), vl, init) as m:
        with pytest.raises(CsmError) as ei:
            m.RemoveEdge(3, 4)
        assert ei.value.code == CSM_E_DELETE_MISSING
    with _engine("parallel_graphflow", query_path2(), vl, init) as m:
        with pytest.raises(CsmError) as ei:
            m.apply_batch(make_updates([1, 0, 0], [2, 0, 0], [3, 1, 1], [0, 0, 0]))
        assert ei.value.code == CSM_E_DELETE_MISSING


def test_multi_query_columns(oracle_lib):
    vl, init, ups, q0 = _case(7, nv=40, p=0.3, n_upd=100, k=4)
    rng = np.random.default_rng(77)
    qs = [q0] + [random_connected_query(k, 1, 3, 2, rng) for k in (3, 5, 6)]
    from paracosm_amd.engine import make_updates
    with _engine("parallel_graphflow", q0, vl, init) as m:
        for q in qs[1:]:
            m.register_query(q)
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
    for j, q in enumerate(qs):
        ref = _oracle(oracle_lib, q, vl, init, ups)
        np.testing.assert_array_equal(pos[:, j], ref.pos, err_msg=f"query {j}")
        np.testing.assert_array_equal(neg[:, j], ref.neg, err_msg=f"query {j}")


def test_classify_matches_oracle(oracle_lib):
    O = oracle_lib
    vl, init, ups, q = _case(9)
    qg = O.OGraph.from_query(q)
    with _engine("parallel_graphflow", q, vl, init) as m:
        add, v1, v2, lab = ups
        for i in range(len(v1)):
            want = O.classify(qg, int(vl[v1[i]]), int(vl[v2[i]]), int(lab[i]))
            assert m.Classify(int(v1[i]), int(v2[i]), int(lab[i])) == want


def test_rmat_config1_slice(oracle_lib):
    """Config-1 shape (R-MAT, 16 uniform labels, 1 edge label, 6-vertex
    random-walk queries, insertion stream) at scale 12."""
    from paracosm_amd import gen
    from paracosm_amd.engine import make_updates
    w = gen.make_workload(12, 8, 16, "uniform", 1, stream="insert", frac_stream=0.1,
                          query_sizes=(6,), queries_per_size=3, seed=4)
    ups = (w.upd_add, w.upd_v1, w.upd_v2, w.upd_label)
    init = (w.init_src, w.init_dst, w.init_elab)
    for qi, q in enumerate(w.queries):
        ref = _oracle(oracle_lib, q, w.vlabels, init, ups)
        for algo in ("parallel_graphflow", "parallel_symbi"):
            with _engine(algo, q, w.vlabels, init) as m:
                pos, neg, _, _ = m.apply_batch(make_updates(*ups))
            print("rmat12 query", qi, algo, int(ref.pos.sum()), flush=True)
            np.testing.assert_array_equal(pos[:, 0], ref.pos)


def _edges_after(init, ups, upto):
    edges = {(min(int(a), int(b)), max(int(a), int(b))): int(l) for a, b, l in zip(*init)}
    add, v1, v2, lab = ups
    for i in range(upto):
        key = (min(int(v1[i]), int(v2[i])), max(int(v1[i]), int(v2[i])))
        if add[i]:
            edges.setdefault(key, int(lab[i]))
        else:
            edges.pop(key, None)
    ks = sorted(edges)
    return (np.asarray([k[0] for k in ks], np.uint32), np.asarray([k[1] for k in ks], np.uint32),
            np.asarray([edges[k] for k in ks], np.uint32))


@pytest.mark.parametrize("algo", ["parallel_symbi", "parallel_turboflux"])
@pytest.mark.parametrize("seed", range(5))
def test_index_is_the_fixpoint(oracle_lib, algo, seed):
    """After every batch the device index (d1, d2, np1, nc2, n1, n2) equals a
    from-scratch build of the graph at that point: over SymBi's DAG, or over
    TurboFlux's spanning tree (its DCG).  The DAG/tree is chosen from the
    initial graph's label frequencies, as the engine does at Preprocessing."""
    O = oracle_lib
    tree = algo == "parallel_turboflux"
    vl, init, ups, q = _case(40 + seed, nv=45, p=0.25, n_vl=2, n_el=2, n_upd=150,
                             k=4 + seed % 3, extra=seed % 3)
    from paracosm_amd.engine import make_updates
    nv = len(vl)
    k = q.k
    g0 = O.OGraph(vl, *init)
    qg = O.OGraph.from_query(q)
    u = make_updates(*ups)
    n_dag_edges = k - 1 if tree else q.num_edges
    with _engine(algo, q, vl, init) as m:
        bounds = [0, 40, 41, 100, len(u)]
        for bi in range(len(bounds)):
            if bi > 0:
                m.apply_batch(u[bounds[bi - 1]:bounds[bi]])
            dg = O.OGraph(vl, *_edges_after(init, ups, bounds[bi]))
            assert m.index_dump(0, 0, nv, 1)[0] == O.index_dump(dg, qg, 0, nv, 1, tree=tree,
                                                                dag_src=g0)[0]
            # rows per dump kind: k, k, k, k, #DAG or tree edges, 2|E_q|
            for what, rows in ((1, k), (2, k), (3, k), (4, k), (5, n_dag_edges),
                               (6, 2 * q.num_edges)):
                dev = m.index_dump(0, what, nv, rows)
                ref = O.index_dump(dg, qg, what, nv, rows, tree=tree, dag_src=g0)
                np.testing.assert_array_equal(dev, ref, err_msg=f"batch {bi} what {what}")


def test_match_budget_stops_only_queries_over_it(oracle_lib):
    """A query whose matches in a batch exceed the budget is stopped and
    flagged; a query under it is exact and unflagged."""
    from paracosm_amd.engine import make_updates
    nv = 30
    vl = np.zeros(nv, np.uint32)
    rng = np.random.default_rng(8)
    _, s, d, el = random_graph(nv, 0.6, 1, 1, rng)
    init, ups = mixed_stream_small(s, d, el, nv, 80, 1.0, 1, rng)
    heavy, light = query_star(4), query_path2()
    rh = _oracle(oracle_lib, heavy, vl, init, ups)
    rl = _oracle(oracle_lib, light, vl, init, ups)
    budget = int(rl.pos.sum()) + 10
    assert int(rh.pos.sum()) > budget
    with _engine("parallel_graphflow", light, vl, init, match_budget=budget) as m:
        m.register_query(heavy)
        pos, _, _, _ = m.apply_batch(make_updates(*ups))
        st = m.query_status()
    assert st.tolist() == [0, 1]
    np.testing.assert_array_equal(pos[:, 0], rl.pos)


@pytest.mark.parametrize("budget", [1, 2, 3, 17])
@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_spilled_subtrees_are_exact(oracle_lib, budget, algo):
    """Tasks that stop after `budget` DFS iterations hand their remaining
    ranges and buffered candidates to the next round; counts and hashes stay
    bit-exact however finely the search is cut."""
    from paracosm_amd.engine import make_updates
    for seed in (0, 3, 6):
        vl, init, ups, q = _case(300 + seed, nv=40, p=0.35, n_vl=2, n_upd=60, k=5 + seed % 3,
                                 extra=1)
        ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
        with _engine(algo, q, vl, init, want_hash=True, spill_budget=budget, split_chunk=4) as m:
            pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
            rounds = m.stats()["max_rounds"]
        print("spill", algo, budget, seed, "rounds", rounds, "total",
              int(ref.pos.sum() + ref.neg.sum()), flush=True)
        np.testing.assert_array_equal(pos[:, 0], ref.pos)
        np.testing.assert_array_equal(neg[:, 0], ref.neg)
        np.testing.assert_array_equal(hsh[:, 0], ref.hash)
        assert budget > 5 or rounds > 1


def _hub_case(seed, nv=300, p=0.03, n_upd=40):
    """One hub adjacent to every vertex plus a sparse random graph, one label:
    the hub's segment is far longer than a 64-entry window."""
    from tests.helpers import mixed_stream_small
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for b in range(1, nv):
        src.append(0)
        dst.append(b)
    for a in range(1, nv):
        for b in range(a + 1, nv):
            if rng.random() < p:
                src.append(a)
                dst.append(b)
    src, dst = np.array(src, np.uint32), np.array(dst, np.uint32)
    el = np.zeros(len(src), np.uint32)
    vl = np.zeros(nv, np.uint32)
    init, ups = mixed_stream_small(src, dst, el, nv, n_upd, 0.65, 1, rng)
    return vl, init, ups


@pytest.mark.parametrize("budget", [1, 2])
@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"])
def test_tiny_budget_with_long_fixed_segments(oracle_lib, budget, algo):
    """Entering position k-2 charges the walk of the last position's fixed
    segment to the task's budget, which a hub's segment exceeds at once; a
    task still takes one step before it spills, so the rounds end and the
    counts stay exact (k = 4: seeds start at position k-2)."""
    from paracosm_amd.engine import make_updates
    from paracosm_amd.gen import Query
    vl, init, ups = _hub_case(11)
    for edges in ([(0, 1), (1, 2), (2, 3), (0, 3)], [(0, 1), (1, 2), (0, 2), (2, 3)]):
        q = Query(vlabels=np.zeros(4, np.uint32), src=np.asarray([e[0] for e in edges], np.uint32),
                  dst=np.asarray([e[1] for e in edges], np.uint32),
                  elabels=np.zeros(len(edges), np.uint32))
        ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
        with _engine(algo, q, vl, init, want_hash=True, spill_budget=budget, split_chunk=4) as m:
            pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
            rounds = m.stats()["max_rounds"]
        print("hub", algo, budget, edges, "rounds", rounds, "total",
              int(ref.pos.sum() + ref.neg.sum()), flush=True)
        assert ref.pos.sum() > 0
        np.testing.assert_array_equal(pos[:, 0], ref.pos)
        np.testing.assert_array_equal(neg[:, 0], ref.neg)
        np.testing.assert_array_equal(hsh[:, 0], ref.hash)


@pytest.mark.parametrize("budget", [0, 2])
@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"])
def test_fixed_segment_beyond_the_leaf_filter(oracle_lib, budget, algo):
    """A hub of ~1,300 neighbours under one label: the last position's fixed
    segment exceeds the 512-entry leaf filter, so the shared leaf count is
    walked in global memory and reused across the entries of position k-2
    (no hashes: the count-only path)."""
    from paracosm_amd.engine import make_updates
    from paracosm_amd.gen import Query
    vl, init, ups = _hub_case(21, nv=2600, p=0.001, n_upd=40)
    edges = [(0, 1), (1, 2), (0, 2), (0, 3)]   # triangle with a pendant on vertex 0
    q = Query(vlabels=np.zeros(4, np.uint32), src=np.asarray([e[0] for e in edges], np.uint32),
              dst=np.asarray([e[1] for e in edges], np.uint32),
              elabels=np.zeros(len(edges), np.uint32))
    ref = _oracle(oracle_lib, q, vl, init, ups)
    kw = {"spill_budget": budget, "split_chunk": 4} if budget else {}
    with _engine(algo, q, vl, init, **kw) as m:
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
    print("wide hub", algo, budget, "total", int(ref.pos.sum() + ref.neg.sum()), flush=True)
    assert ref.pos.sum() > 0 and ref.neg.sum() > 0
    np.testing.assert_array_equal(pos[:, 0], ref.pos)
    np.testing.assert_array_equal(neg[:, 0], ref.neg)


@pytest.mark.parametrize("k", [10, 12, 14, 17, 24, 28])
@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_large_queries(oracle_lib, k, algo):
    """Queries beyond 8 vertices run the larger LDS layouts of the search
    (10, 12, 16, 24 and 32 vertices); a small spill budget also cuts their
    deep prefixes into tasks."""
    from paracosm_amd.engine import make_updates
    for seed in (0, 1):
        vl, init, ups, q = walk_case(700 + 10 * k + seed, k)
        ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
        assert ref.pos.sum() > 0
        for budget in (0, 2):
            kw = {"spill_budget": budget, "split_chunk": 4} if budget else {}
            with _engine(algo, q, vl, init, want_hash=True, **kw) as m:
                pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
            print("large", algo, k, seed, budget, "edges", len(q.src), "total",
                  int(ref.pos.sum() + ref.neg.sum()), flush=True)
            np.testing.assert_array_equal(pos[:, 0], ref.pos)
            np.testing.assert_array_equal(neg[:, 0], ref.neg)
            np.testing.assert_array_equal(hsh[:, 0], ref.hash)


def test_query_beyond_32_vertices_is_rejected():
    from paracosm_amd.engine import CsmError, Matching
    rng = np.random.default_rng(5)
    vl, s, d, el = random_graph(60, 0.2, 2, 1, rng)
    q = random_connected_query(33, 0, 2, 1, rng)
    with pytest.raises(CsmError):
        Matching("parallel_graphflow", q, vl, s, d, el).close()


def _present_after(init, ups, n):
    """Edge set {(min, max): label} after the first n updates."""
    present = {(min(a, b), max(a, b)): l for a, b, l in zip(*[x.tolist() for x in init])}
    add, v1, v2, lab = [x.tolist() for x in ups]
    for i in range(n):
        key = (min(v1[i], v2[i]), max(v1[i], v2[i]))
        if add[i]:
            present[key] = lab[i]
        else:
            present.pop(key, None)
    return present


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_remove_vertex_matches_oracle(oracle_lib, algo):
    """A vertex deletion inside a stream removes the vertex and its edges
    without counting matches (as the reference's RemoveVertex); the updates
    after it are exact against the oracle run on the graph without those
    edges, and the SymBi index is the fix-point of that graph."""
    O = oracle_lib
    from paracosm_amd.engine import make_updates
    for seed in (0, 1, 2):
        vl, init, ups, q = _case(900 + seed, nv=45, p=0.3, n_vl=2, n_el=2, n_upd=140, k=4,
                                 extra=seed % 2)
        cut = 60
        present = _present_after(init, ups, cut)
        deg = {}
        for a, b in present:
            deg[a] = deg.get(a, 0) + 1
            deg[b] = deg.get(b, 0) + 1
        h = max(deg, key=lambda v: (deg[v], -v))
        inc = sorted(k for k in present if h in k)
        # the rest of the stream without updates that touch h
        add, v1, v2, lab = ups
        rest = [i for i in range(cut, len(v1)) if v1[i] != h and v2[i] != h]
        B = tuple(x[rest] for x in ups)
        A = tuple(x[:cut] for x in ups)
        D = (np.zeros(len(inc), np.uint8), np.asarray([a for a, _ in inc], np.uint32),
             np.asarray([b for _, b in inc], np.uint32),
             np.asarray([present[k] for k in inc], np.uint32))
        ref = _oracle(O, q, vl, init, tuple(np.concatenate(z) for z in zip(A, D, B)))
        stream = make_updates(*(np.concatenate(z) for z in zip(A, ([0], [h], [0], [0]), B)),
                              vertex=np.r_[np.zeros(cut, bool), [True], np.zeros(len(rest), bool)])
        with _engine(algo, q, vl, init) as m:
            pos, neg, safe, _ = m.apply_batch(stream)
            print("remove_vertex", algo, seed, "vertex", h, "edges", len(inc), "total",
                  int(ref.pos.sum() + ref.neg.sum()), flush=True)
            np.testing.assert_array_equal(pos[:cut, 0], ref.pos[:cut])
            np.testing.assert_array_equal(neg[:cut, 0], ref.neg[:cut])
            assert pos[cut, 0] == 0 and neg[cut, 0] == 0
            np.testing.assert_array_equal(pos[cut + 1:, 0], ref.pos[cut + len(inc):])
            np.testing.assert_array_equal(neg[cut + 1:, 0], ref.neg[cut + len(inc):])
            if algo == "parallel_symbi":
                nv = len(vl)
                vl2 = vl.copy()
                vl2[h] = 1000  # a label no query vertex carries: h is in no index row
                fin = _present_after(init, tuple(np.concatenate(z) for z in zip(A, D, B)),
                                     cut + len(inc) + len(rest))
                es = sorted(fin)
                dg = O.OGraph(vl2, np.asarray([a for a, _ in es], np.uint32),
                              np.asarray([b for _, b in es], np.uint32),
                              np.asarray([fin[k] for k in es], np.uint32))
                qg = O.OGraph.from_query(q)
                for what, rows in ((1, q.k), (2, q.k), (3, q.k), (4, q.k), (5, q.num_edges),
                                   (6, 2 * q.num_edges)):
                    np.testing.assert_array_equal(m.index_dump(0, what, nv, rows),
                                                  O.symbi_dump(dg, qg, what, nv, rows),
                                                  err_msg=f"what {what}")


def test_vertex_record_in_a_device_batch_is_an_error():
    from paracosm_amd.engine import CsmError, make_updates
    vl, init, ups, q = _case(950)
    u = make_updates(*(x[:8] for x in ups))
    u["type"][3] = ord("v")
    with _engine("parallel_graphflow", q, vl, init) as m:
        d_u = m.device_buffer(u.nbytes)
        d_c = m.device_buffer(8 * len(u))
        d_u.upload(u.view(np.uint8))
        with pytest.raises(CsmError):
            m.apply_batch_device(d_u.ptr, len(u), d_c.ptr)


@pytest.mark.parametrize("chunk", [1, 3])
@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_fine_chunks_exact(oracle_lib, chunk, algo):
    """Tasks cut into chunks of 1 or 3 candidates: each wave interleaves many
    tasks and plans, so state a wave keeps between chunks (cached segment
    bounds and dirty bits, the leaf filter, plan fields) is exercised in
    many orders; counts and hashes stay bit-exact."""
    from paracosm_amd.engine import make_updates
    for seed in range(6):
        vl, init, ups, q = _case(1200 + seed, nv=60, p=0.3, n_vl=2, n_upd=150,
                                 k=5 + seed % 4, extra=seed % 3)
        ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
        with _engine(algo, q, vl, init, want_hash=True, split_chunk=chunk) as m:
            pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
        print("fine", algo, chunk, seed, "k", q.k, "total", int(ref.pos.sum() + ref.neg.sum()),
              flush=True)
        np.testing.assert_array_equal(pos[:, 0], ref.pos)
        np.testing.assert_array_equal(neg[:, 0], ref.neg)
        np.testing.assert_array_equal(hsh[:, 0], ref.hash)


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"])
def test_reference_orders_agree(oracle_lib, algo):
    """The GraphFlow plugin's own matching orders and the default
    estimated-cost orders enumerate the same embeddings (counts and hashes)."""
    from paracosm_amd.engine import make_updates
    for seed in (2, 5, 9):
        vl, init, ups, q = _case(1400 + seed, nv=50, p=0.3, n_vl=3, n_upd=120, k=5 + seed % 3,
                                 extra=seed % 2)
        ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
        for reference_orders in (False, True):
            with _engine(algo, q, vl, init, want_hash=True,
                         reference_orders=reference_orders) as m:
                pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
            np.testing.assert_array_equal(pos[:, 0], ref.pos)
            np.testing.assert_array_equal(neg[:, 0], ref.neg)
            np.testing.assert_array_equal(hsh[:, 0], ref.hash)


@pytest.mark.parametrize("algo,budget", [("parallel_graphflow", 0), ("parallel_symbi", 0),
                                         ("parallel_graphflow", 2), ("parallel_turboflux", 3)])
def test_graph_replays_match_eager(oracle_lib, algo, budget):
    """Steady-state batches replayed as HIP graphs (from the second batch of a
    size and buffer parity on) count exactly what the oracle and an engine
    without graphs count, on every replay, with eager batches of another size
    and host reads between the replays.  A tiny spill budget makes batches
    need more search rounds than the graph holds (eager continuation)."""
    from paracosm_amd.engine import make_updates
    vl, init, ups, q = _case(77, nv=70, p=0.3, n_upd=640, k=5, extra=2)
    ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
    u = make_updates(*ups)
    kw = dict(want_hash=True, split_chunk=8, spill_budget=budget) if budget else dict(want_hash=True)
    with _engine(algo, q, vl, init, **kw) as m, \
            _engine(algo, q, vl, init, no_graphs=True, **kw) as e:
        s = 0
        sizes = [40] * 6 + [17] + [40] * 6 + [33] + [40] * 4
        for b in sizes:
            p1, n1, _, h1 = m.apply_batch(u[s:s + b])
            p2, n2, _, h2 = e.apply_batch(u[s:s + b])
            np.testing.assert_array_equal(p1[:, 0], ref.pos[s:s + b])
            np.testing.assert_array_equal(n1[:, 0], ref.neg[s:s + b])
            np.testing.assert_array_equal(h1[:, 0], ref.hash[s:s + b])
            np.testing.assert_array_equal(p1, p2)
            np.testing.assert_array_equal(h1, h2)
            s += b
        g = m.stats()["graph_batches"]
        print("graph replays", algo, budget, g, "rounds", m.stats()["max_rounds"], flush=True)
        assert g >= 10 and e.stats()["graph_batches"] == 0


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"])
def test_add_vertex_then_edges_matches_oracle(oracle_lib, algo):
    """Vertex insertions inside a stream ('v' records, Graph::AddVertex,
    graph.cpp:24-44, through the plugins' AddVertex, parrallel_symbi.cpp:
    3812-3872) followed by edges to the new vertices: counts equal the oracle
    run on a graph that held the new vertices, isolated, from the start; the
    SymBi/TurboFlux index gets rows for them (its fix-point is checked
    against a from-scratch build afterwards)."""
    O = oracle_lib
    from paracosm_amd.engine import make_updates
    rng = np.random.default_rng(4242)
    nv0, extra_v = 40, 12
    vl_all, s, d, el = random_graph(nv0 + extra_v, 0.3, 2, 2, rng)
    keep = (s < nv0) & (d < nv0)
    init = (s[keep], d[keep], el[keep])
    new_edges = [(int(a), int(b), int(c)) for a, b, c in zip(s[~keep], d[~keep], el[~keep])]
    rng.shuffle(new_edges)
    q = random_connected_query(4, 1, 2, 2, rng)
    # stream: each new vertex announced before its first edge
    recs, announced = [], set()
    for a, b, c in new_edges:
        for v in (a, b):
            if v >= nv0 and v not in announced:
                announced.add(v)
                recs.append((True, 1, v, 0, int(vl_all[v])))
        recs.append((False, 1, a, b, c))
    for v in range(nv0, nv0 + extra_v):   # vertices without edges, announced last
        if v not in announced:
            recs.append((True, 1, v, 0, int(vl_all[v])))
    vert = np.asarray([r[0] for r in recs])
    add = np.asarray([r[1] for r in recs], np.uint8)
    v1 = np.asarray([r[2] for r in recs], np.uint32)
    v2 = np.asarray([r[3] for r in recs], np.uint32)
    lab = np.asarray([r[4] for r in recs], np.uint32)
    e_idx = np.flatnonzero(~vert)
    ref = _oracle(O, q, vl_all, init, (add[e_idx], v1[e_idx], v2[e_idx], lab[e_idx]))
    with _engine(algo, q, vl_all[:nv0], init) as m:
        pos, neg, _, _ = m.apply_batch(make_updates(add, v1, v2, lab, vertex=vert))
        print("add_vertex", algo, "new vertices", len(announced), "edges", len(e_idx), "total",
              int(ref.pos.sum()), flush=True)
        np.testing.assert_array_equal(pos[e_idx, 0], ref.pos)
        assert (pos[vert, 0] == 0).all() and (neg[:, 0] == 0).all()
        assert ref.pos.sum() > 0
        if algo != "parallel_graphflow":
            nv = nv0 + extra_v
            final = (np.concatenate([init[0], v1[e_idx]]), np.concatenate([init[1], v2[e_idx]]),
                     np.concatenate([init[2], lab[e_idx]]))
            fg = O.OGraph(vl_all, *final)
            # the index's DAG/tree comes from the initial graph, whose label
            # frequencies do not count the vertices added later
            vl0 = vl_all.copy()
            vl0[nv0:] = 0xFFFFFFFF
            ig = O.OGraph(vl0, *init)
            for what in (1, 2):
                got = m.index_dump(0, what, nv, q.k)
                want = O.index_dump(fg, O.OGraph.from_query(q), what, nv, q.k,
                                    tree=algo == "parallel_turboflux", dag_src=ig)
                np.testing.assert_array_equal(got[:, :nv], want)


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_hub_lists(oracle_lib, algo):
    """Lists long enough for their key directories to be built by one binary
    search per key (at least 16 entries per directory slot): a hub joined to
    every other vertex, a mixed stream around it, several batch sizes."""
    from paracosm_amd.engine import make_updates
    rng = np.random.default_rng(4242)
    nv = 600
    vl, s, d, el = random_graph(nv, 0.01, 1, 2, rng)
    keep = s != 0                      # the hub (vertex 0) gets every other vertex
    s = np.concatenate([s[keep], np.zeros(nv - 1, np.uint32)])
    d = np.concatenate([d[keep], np.arange(1, nv, dtype=np.uint32)])
    el = rng.integers(0, 2, size=len(s)).astype(np.uint32)
    init, ups = mixed_stream_small(s, d, el, nv, 400, 0.6, 2, rng)
    q = random_connected_query(4, 1, 1, 2, rng)
    ref = _oracle(oracle_lib, q, vl, init, ups, want_hash=True)
    assert ref.pos.sum() > 1000
    for batch in (0, 37):
        with _engine(algo, q, vl, init, want_hash=True, max_batch=batch) as m:
            pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
        np.testing.assert_array_equal(pos[:, 0], ref.pos)
        np.testing.assert_array_equal(neg[:, 0], ref.neg)
        np.testing.assert_array_equal(hsh[:, 0], ref.hash)


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_hub_list_merges(oracle_lib, algo):
    """A hub list of 6,000-12,000 entries merged with thousands of new
    entries per batch (and its key directory rebuilt by binary searches):
    6,000 initial hub edges and 6,000 more inserted, among insertions and
    deletions of 40,000 random edges, in batches of 4,000; triangles."""
    from paracosm_amd.engine import make_updates
    from paracosm_amd.gen import Query
    rng = np.random.default_rng(99)
    nv = 12000
    a = rng.integers(1, nv, size=60000)
    b = rng.integers(1, nv, size=60000)
    rand = sorted({(int(min(x, y)), int(max(x, y))) for x, y in zip(a, b) if x != y})[:40000]
    order = rng.permutation(len(rand))
    held = [rand[i] for i in order[:5000]]                 # inserted by the stream
    dels = [rand[i] for i in order[5000:6000]]             # deleted by the stream
    init_pairs = [rand[i] for i in order[5000:]] + [(0, v) for v in range(1, 6001)]
    ups = ([(1, 0, v) for v in range(6001, nv)] + [(1, x, y) for x, y in held] +
           [(0, x, y) for x, y in dels])
    ups = [ups[i] for i in rng.permutation(len(ups))]
    init = (np.array([p[0] for p in init_pairs], np.uint32),
            np.array([p[1] for p in init_pairs], np.uint32),
            np.zeros(len(init_pairs), np.uint32))
    stream = (np.array([u[0] for u in ups], np.uint8), np.array([u[1] for u in ups], np.uint32),
              np.array([u[2] for u in ups], np.uint32), np.zeros(len(ups), np.uint32))
    vl = np.zeros(nv, np.uint32)
    q = Query(vlabels=np.zeros(3, np.uint32), src=np.array([0, 0, 1], np.uint32),
              dst=np.array([1, 2, 2], np.uint32), elabels=np.zeros(3, np.uint32))
    ref = _oracle(oracle_lib, q, vl, init, stream, want_hash=True)
    assert ref.pos.sum() > 0 and ref.neg.sum() > 0
    with _engine(algo, q, vl, init, want_hash=True, max_batch=4000) as m:
        pos, neg, _, hsh = m.apply_batch(make_updates(*stream))
    np.testing.assert_array_equal(pos[:, 0], ref.pos)
    np.testing.assert_array_equal(neg[:, 0], ref.neg)
    np.testing.assert_array_equal(hsh[:, 0], ref.hash)
