"""Pins the oracle's index and planning functions from their definitions.

The SymBi/TurboFlux index (d1, d2, np1, nc2, n1, n2; SymBi::BuildDCS,
ParaCOSM/CSM/matching/SymBi/symbi.cpp:246-377) and the SymBi DAG
(SymBi::BuildDAG, symbi.cpp:113-244) are recomputed here in plain Python from
the set definitions, with no incremental rule and no code shared with the
oracle, and compared element by element with orc_index_dump / orc_dag_parents
on small random graphs.  The GraphFlow matching orders
(Graphflow::GenerateMatchingOrder, graphflow.cpp:36-160) are checked on a
hand-worked query whose orders depend on both tie rules.  TurboFlux's
explicit-path counts and matching orders (Parallel_TurboFlux::CountDownwards
and GenerateMatchingOrder, parallel_turboflux.cpp:469-700) are checked against
a closed form of the memoised count and the order definition on random cases,
and on two hand-worked ones.
"""
import numpy as np
import pytest

from paracosm_amd.gen import Query
from tests.helpers import random_connected_query, random_graph


def _adj(nv, src, dst, el):
    """Per vertex: sorted list of (neighbour, edge label)."""
    a = [[] for _ in range(nv)]
    for s, d, l in zip(src.tolist(), dst.tolist(), el.tolist()):
        a[s].append((d, l))
        a[d].append((s, l))
    return [sorted(x) for x in a]


def _qadj(q):
    return _adj(q.k, q.src, q.dst, q.elabels)


def dag_from_definition(q, vl, adj):
    """Root and parents (in neighbour-id order) of SymBi's BFS DAG.

    Root: the endpoint of the query edge with the fewest label-matching
    directed data edges (first such edge over u ascending, neighbours
    ascending, counted from the smaller endpoint's side), the endpoint whose
    label is rarer in the data, on equal frequency the first endpoint if its
    degree is larger, else the second.  DAG: BFS from the root over
    neighbours in id order; an edge points from the vertex dequeued first."""
    qa = _qadj(q)
    best, mv = None, None
    for u in range(q.k):
        for uo, ul in qa[u]:
            if u > uo:
                continue
            f = sum(1 for v in range(len(vl)) if vl[v] == q.vlabels[u]
                    for w, l in adj[v] if vl[w] == q.vlabels[uo] and l == ul)
            if best is None or f < best:
                best, mv = f, (u, uo)
    f0 = int(np.sum(vl == q.vlabels[mv[0]]))
    f1 = int(np.sum(vl == q.vlabels[mv[1]]))
    if f0 != f1:
        root = mv[1] if f0 > f1 else mv[0]
    else:
        root = mv[0] if len(qa[mv[0]]) > len(qa[mv[1]]) else mv[1]
    popped, queued, fifo = set(), {root}, [root]
    parents = [[] for _ in range(q.k)]
    while fifo:
        u = fifo.pop(0)
        popped.add(u)
        for uo, _ in qa[u]:
            if uo not in queued:
                queued.add(uo)
                fifo.append(uo)
            if uo in popped:
                parents[u].append(uo)
    return root, parents


def index_from_definition(q, vl, adj, parents):
    """The index over a DAG given by `parents`, from its definitions:

    S(v, u->x)     = neighbours w of v with L(w) = L(x) and l(v,w) = l(u,x)
    n1[u<-p][v]    = |{w in S(v, u->p) : d1[p][w]}|
    np1[u][v]      = #parents p with n1[u<-p][v] > 0
    d1[u][v]       = L(v) = L(u) and np1[u][v] = #parents(u)
    n2[u->x][v]    = |{w in S(v, u->x) : d2[x][w]}| for x a parent or child
    nc2[u][v]      = #children c with n2[u->c][v] > 0
    d2[u][v]       = d1[u][v] and nc2[u][v] = #children(u)
    with every counter 0 where L(v) != L(u)."""
    k, nv = q.k, len(vl)
    qa = _qadj(q)
    lab = {(u, x): l for u in range(k) for x, l in qa[u]}
    children = [[c for c in range(k) if u in parents[c]] for u in range(k)]
    # topological order: repeatedly take vertices whose parents are placed
    order, placed = [], set()
    while len(order) < k:
        for u in range(k):
            if u not in placed and all(p in placed for p in parents[u]):
                order.append(u)
                placed.add(u)

    def S(v, u, x):
        return [w for w, l in adj[v] if vl[w] == q.vlabels[x] and l == lab[(u, x)]]

    d1 = np.zeros((k, nv), np.uint32)
    d2 = np.zeros((k, nv), np.uint32)
    np1 = np.zeros((k, nv), np.uint32)
    nc2 = np.zeros((k, nv), np.uint32)
    n1 = {}
    n2 = {}
    for u in order:
        for p in parents[u]:
            n1[(u, p)] = np.zeros(nv, np.uint32)
        for v in range(nv):
            if vl[v] != q.vlabels[u]:
                continue
            for p in parents[u]:
                n1[(u, p)][v] = sum(1 for w in S(v, u, p) if d1[p][w])
                np1[u][v] += n1[(u, p)][v] > 0
            d1[u][v] = np1[u][v] == len(parents[u])
    for u in reversed(order):
        for c in children[u]:
            n2[(u, c)] = np.zeros(nv, np.uint32)
        for v in range(nv):
            if vl[v] != q.vlabels[u]:
                continue
            for c in children[u]:
                n2[(u, c)][v] = sum(1 for w in S(v, u, c) if d2[c][w])
                nc2[u][v] += n2[(u, c)][v] > 0
            d2[u][v] = d1[u][v] and nc2[u][v] == len(children[u])
    for u in range(k):
        for p in parents[u]:
            n2[(u, p)] = np.zeros(nv, np.uint32)
            for v in range(nv):
                if vl[v] == q.vlabels[u]:
                    n2[(u, p)][v] = sum(1 for w in S(v, u, p) if d2[p][w])
    n1_rows = np.asarray([n1[(u, p)] for u in range(k) for p in parents[u]]).reshape(-1, nv)
    n2_rows = np.asarray([n2.get((u, x), np.zeros(nv, np.uint32))
                          for u in range(k) for x, _ in qa[u]]).reshape(-1, nv)
    return {1: d1, 2: d2, 3: np1, 4: nc2, 5: n1_rows, 6: n2_rows}


def _case(seed):
    rng = np.random.default_rng(9000 + seed)
    nv = 12 + seed % 9
    vl, s, d, el = random_graph(nv, 0.3 + 0.05 * (seed % 4), 2 + seed % 2, 1 + seed % 2, rng)
    q = random_connected_query(3 + seed % 4, seed % 3, 2 + seed % 2, 1 + seed % 2, rng)
    return vl, s, d, el, q


@pytest.mark.parametrize("seed", range(12))
def test_symbi_dag_from_definition(oracle_lib, seed):
    O = oracle_lib
    vl, s, d, el, q = _case(seed)
    root, par = dag_from_definition(q, vl, _adj(len(vl), s, d, el))
    got_root, got_par = O.dag_parents(O.OGraph(vl, s, d, el), O.OGraph.from_query(q), q.k,
                                      tree=False)
    assert (got_root, got_par) == (root, par)


@pytest.mark.parametrize("tree", [False, True])
@pytest.mark.parametrize("seed", range(12))
def test_index_from_definition(oracle_lib, tree, seed):
    """Every row of the oracle's from-scratch index equals the definition, on
    SymBi's DAG and on TurboFlux's spanning tree (parents taken from the
    oracle; the tree rule is pinned in test_oracle.py)."""
    O = oracle_lib
    vl, s, d, el, q = _case(seed)
    dg, qg = O.OGraph(vl, s, d, el), O.OGraph.from_query(q)
    root, par = O.dag_parents(dg, qg, q.k, tree=tree)
    want = index_from_definition(q, vl, _adj(len(vl), s, d, el), par)
    nv = len(vl)
    assert int(O.index_dump(dg, qg, 0, nv, 1, tree=tree)[0]) == root
    for what in range(1, 7):
        rows = want[what].shape[0]
        got = O.index_dump(dg, qg, what, nv, rows, tree=tree)
        np.testing.assert_array_equal(got, want[what], err_msg=f"what={what}")


def test_index_definition_is_not_trivial():
    """The cases above include d1 without d2 and partial counters, so a
    dropped term or a wrong parent/child role changes some row."""
    d1_not_d2 = partial = 0
    for seed in range(12):
        vl, s, d, el, q = _case(seed)
        root, par = dag_from_definition(q, vl, _adj(len(vl), s, d, el))
        idx = index_from_definition(q, vl, _adj(len(vl), s, d, el), par)
        d1_not_d2 += int(np.sum((idx[1] == 1) & (idx[2] == 0)))
        partial += int(np.sum((idx[3] > 0) & (idx[1] == 0)))
    assert d1_not_d2 > 10 and partial > 0


def test_graphflow_orders_tie_rules(oracle_lib):
    """Hand-worked orders (graphflow.cpp:36-160) for the query
    0-1, 0-2, 1-3, 2-3, 2-4 (degrees 2, 2, 3, 2, 1).

    Order 0 starts at the first vertex of largest degree (2); each next vertex
    has the most already-placed neighbours, ties to the strictly larger degree
    and otherwise to the smaller id: 2, 0, 1, 3, 4.  Its backward lists are
    [2], [0], [1, 2], [2], so orders 1..4 start on the edges (1,0), (3,1),
    (3,2), (4,2) (the lower_bound walk over the offsets [0, 1, 2, 4, 5]).
    Order 2 takes 2 before 0 (equal adjacency, larger degree); orders 0, 3
    and 4 take the smaller id among equal degrees."""
    O = oracle_lib
    q = Query(vlabels=np.zeros(5, np.uint32), src=np.asarray([0, 0, 1, 2, 2], np.uint32),
              dst=np.asarray([1, 2, 3, 3, 4], np.uint32), elabels=np.zeros(5, np.uint32))
    g = O.OGraph(np.zeros(2, np.uint32), [0], [1], [0])
    got = O.graphflow_orders(g, O.OGraph.from_query(q), 5, 5).tolist()
    assert got == [[2, 0, 1, 3, 4], [1, 0, 2, 3, 4], [3, 1, 2, 0, 4], [3, 2, 0, 1, 4],
                   [4, 2, 0, 1, 3]]


# ------------------------------------------------ TurboFlux's matching orders
U32 = 1 << 32


def turboflux_paths_from_definition(q, vl, adj, root, parents, d2):
    """TurboFlux's explicit-path counts (Parallel_TurboFlux::CountDownwards,
    parallel_turboflux.cpp:469-491, started once per data vertex of the root's
    label, :536-541) in closed form, tree edge by tree edge:

    C(c, v)      = {w in S(v, u->c) : d2[c][w]} for the tree edge u -> c
    calls(r, v)  = 1 for every v with L(v) = L(r), r the root
    calls(c, w)  = #{v : calls(u, v) > 0 and w in C(c, v)}
    paths[c]     = sum_v calls(u, v) * |C(c, v)|  (mod 2^32)
    paths[r]     = 2^32 - 2.

    calls(c, w) counts the distinct parent vertices that reach w, not the
    paths into w: the memo of a (c, parent vertex) pair replays its count on
    later runs without walking again."""
    k, nv = q.k, len(vl)
    qa = _qadj(q)
    lab = {(u, x): l for u in range(k) for x, l in qa[u]}
    children = [[c for c in range(k) if parents[c] == [u]] for u in range(k)]
    calls = {root: [1 if vl[v] == q.vlabels[root] else 0 for v in range(nv)]}
    paths = [0] * k
    frontier = [root]
    while frontier:
        u = frontier.pop(0)
        for c in children[u]:
            calls[c] = [0] * nv
            for v in range(nv):
                if not calls[u][v]:
                    continue
                cand = [w for w, l in adj[v] if vl[w] == q.vlabels[c] and l == lab[(u, c)]
                        and d2[c][w]]
                paths[c] += calls[u][v] * len(cand)
                for w in cand:
                    calls[c][w] += 1
            frontier.append(c)
    paths = [p % U32 for p in paths]
    paths[root] = U32 - 2
    return paths


def turboflux_orders_from_definition(q, root, parents, paths):
    """GenerateMatchingOrder (parallel_turboflux.cpp:533-700): the initial
    order lists the tree so that the leaf with the fewest paths (ties: the
    smallest id) goes last, then the next leaf of what remains, a parent
    becoming a leaf when its children are gone; the order of a query edge
    {a, b} (a < b, a tree edge written child first) starts with its two
    endpoints and the second one's ancestors up to the root, followed by the
    other vertices in initial order.  Edges go by a, then b, ascending."""
    k = q.k
    par = [p[0] if p else None for p in parents]
    remaining = set(range(k))
    initial = []
    while remaining:
        leaves = [u for u in remaining if not any(par[c] == u for c in remaining)]
        u = min(leaves, key=lambda x: (paths[x], x))
        initial.insert(0, u)
        remaining.remove(u)
    edges = sorted({(min(a, b), max(a, b)) for a, b in zip(q.src.tolist(), q.dst.tolist())})
    orders = []
    for a, b in edges:
        first, second = (b, a) if par[b] == a else (a, b)
        o = [first, second]
        x = second
        while x != root:
            x = par[x]
            o.append(x)
        o += [x for x in initial if x not in o]
        orders.append(o)
    return orders


def test_turboflux_paths_hand_worked(oracle_lib):
    """Query path 0-1-2-3 (labels 0, 1, 2, 3).  Data: two label-0 vertices
    0 and 1 on one label-1 vertex 2, which has the label-2 vertex 5, which has
    the label-3 vertices 7 and 8; vertices 3, 4, 6, 9, 10 (labels 1, 1, 2, 2,
    3) make label 0 the rarest, and 3-5 keeps the three query edges at two
    matching data edges each, so the root is 0 and the tree is the path.
    calls(1, 2) = 2 (both roots reach 2), calls(2, 5) = 1 (one parent vertex),
    so paths = [2^32 - 2, 2, 2, 2], where the true numbers of root-to-vertex
    paths are 2, 2 and 4: the memo's count, not the path count."""
    O = oracle_lib
    q = Query(vlabels=np.asarray([0, 1, 2, 3], np.uint32), src=np.asarray([0, 1, 2], np.uint32),
              dst=np.asarray([1, 2, 3], np.uint32), elabels=np.zeros(3, np.uint32))
    vl = np.asarray([0, 0, 1, 1, 1, 2, 2, 3, 3, 2, 3], np.uint32)
    e = [(0, 2), (1, 2), (2, 5), (3, 5), (5, 7), (5, 8)]
    dg = O.OGraph(vl, [a for a, _ in e], [b for _, b in e], [0] * len(e))
    qg = O.OGraph.from_query(q)
    root, par = O.dag_parents(dg, qg, 4, tree=True)
    assert (root, par) == (0, [[], [0], [1], [2]])
    orders, paths = O.turboflux_orders(dg, qg, 4, 3)
    assert paths.tolist() == [U32 - 2, 2, 2, 2]
    assert orders.tolist() == [[1, 0, 2, 3], [2, 1, 0, 3], [3, 2, 1, 0]]


def test_turboflux_initial_order_hand_worked(oracle_lib):
    """Query star 0-{1, 2, 3} (labels 1, 2, 3 on the leaves) with 2-4 below
    leaf 2 (label 4).  Data: the root vertex 0 has three label-1, one label-2
    and two label-3 neighbours, and the label-2 one has one label-4
    neighbour: paths [2^32 - 2, 3, 1, 2, 1].  Filling the initial order from
    the back: 4 (1 path; 2 becomes a leaf), 2 (1), 3 (2), 1 (3), the root:
    0, 1, 3, 2, 4."""
    O = oracle_lib
    q = Query(vlabels=np.asarray([0, 1, 2, 3, 4], np.uint32),
              src=np.asarray([0, 0, 0, 2], np.uint32), dst=np.asarray([1, 2, 3, 4], np.uint32),
              elabels=np.zeros(4, np.uint32))
    vl = np.asarray([0, 1, 1, 1, 2, 3, 3, 4], np.uint32)
    e = [(0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (0, 6), (4, 7)]
    dg = O.OGraph(vl, [a for a, _ in e], [b for _, b in e], [0] * len(e))
    qg = O.OGraph.from_query(q)
    root, par = O.dag_parents(dg, qg, 5, tree=True)
    assert (root, par) == (0, [[], [0], [0], [0], [2]])
    orders, paths = O.turboflux_orders(dg, qg, 5, 4)
    assert paths.tolist() == [U32 - 2, 3, 1, 2, 1]
    # edge orders: endpoints (child first), the ancestors, then 0, 1, 3, 2, 4
    assert orders.tolist() == [[1, 0, 3, 2, 4], [2, 0, 1, 3, 4], [3, 0, 1, 2, 4],
                               [4, 2, 0, 1, 3]]


@pytest.mark.parametrize("seed", range(12))
def test_turboflux_orders_from_definition(oracle_lib, seed):
    """The oracle's recursive, memoised CountDownwards and its order rule
    equal the closed form and the order definition above on random cases
    (tree and d2 from the oracle; both are pinned by the tests above)."""
    O = oracle_lib
    vl, s, d, el, q = _case(seed)
    dg, qg = O.OGraph(vl, s, d, el), O.OGraph.from_query(q)
    root, par = O.dag_parents(dg, qg, q.k, tree=True)
    d2 = index_from_definition(q, vl, _adj(len(vl), s, d, el), par)[2]
    paths = turboflux_paths_from_definition(q, vl, _adj(len(vl), s, d, el), root, par, d2)
    ne = len({(min(a, b), max(a, b)) for a, b in zip(q.src.tolist(), q.dst.tolist())})
    got_orders, got_paths = O.turboflux_orders(dg, qg, q.k, ne)
    assert got_paths.tolist() == paths
    assert got_orders.tolist() == turboflux_orders_from_definition(q, root, par, paths)


def test_turboflux_order_cases_are_not_trivial(oracle_lib):
    """Across the random cases the memo count differs from the plain
    root-to-vertex path count somewhere, and some order's tail is not the
    tree's BFS order, so a dropped memo or a wrong ranking changes a result."""
    memo_differs = order_differs = 0
    for seed in range(12):
        vl, s, d, el, q = _case(seed)
        adj = _adj(len(vl), s, d, el)
        root, par = oracle_lib.dag_parents(oracle_lib.OGraph(vl, s, d, el),
                                           oracle_lib.OGraph.from_query(q), q.k, tree=True)
        d2 = index_from_definition(q, vl, adj, par)[2]
        paths = turboflux_paths_from_definition(q, vl, adj, root, par, d2)
        # plain path counts: number of root-to-vertex walks through d2 vertices
        qa = _qadj(q)
        lab = {(u, x): l for u in range(q.k) for x, l in qa[u]}
        cnt = {root: [1 if vl[v] == q.vlabels[root] else 0 for v in range(len(vl))]}
        order = [root]
        for u in order:
            for c in range(q.k):
                if par[c] == [u]:
                    cnt[c] = [0] * len(vl)
                    for v in range(len(vl)):
                        for w, l in adj[v]:
                            if cnt[u][v] and vl[w] == q.vlabels[c] and l == lab[(u, c)] and d2[c][w]:
                                cnt[c][w] += cnt[u][v]
                    order.append(c)
        true_paths = [sum(cnt[u]) for u in range(q.k)]
        memo_differs += any(true_paths[u] != paths[u] for u in range(q.k) if u != root)
        o = turboflux_orders_from_definition(q, root, par, paths)
        bfs = [root] + [x for x in order if x != root]
        order_differs += any(oo[2:] != [x for x in bfs if x not in oo[:2]] for oo in o)
    assert memo_differs > 0 and order_differs > 0
