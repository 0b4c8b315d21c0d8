"""--print-enum on the device: the embedding rows of every call.

The reference prints each embedding as its search reaches it
(SymBi/symbi.cpp:564-569, Parallel_GraphFlow/parallel_graphflow.cpp:367).
Here the rows of a call are checked three ways: every row is an embedding of
its query in the graph its update sees and contains the updated edge; no row
repeats; and per update and query the rows number exactly the oracle's count
and hash to the oracle's embedding-set hash.  Distinct valid embeddings as
many as the exact count are the whole set.
"""
import numpy as np
import pytest

from tests.helpers import mixed_stream_small, random_connected_query, random_graph

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
NONE = 0xFFFFFFFF


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _row_hash(m):
    h = 0
    for v in m:
        h = _mix64((h + int(v) + 1) & M64)
    return h


def _edges(q):
    return list(zip(q.src.tolist(), q.dst.tolist(), q.elabels.tolist()))


def _check_row(m, q, vl, E, upd=None):
    """m: data vertex per query vertex; E: {(a, b): label} with a < b."""
    k = len(q.vlabels)
    assert len(m) == k and len(set(m)) == k, ("not injective", m)
    for u in range(k):
        assert vl[m[u]] == q.vlabels[u], ("vertex label", m, u)
    hit = False
    for a, b, l in _edges(q):
        x, y = sorted((m[a], m[b]))
        assert E.get((x, y)) == l, ("query edge not in the graph", m, a, b)
        hit = hit or (upd is not None and (x, y) == upd)
    assert upd is None or hit, ("update edge not in the embedding", m, upd)


def _replay(rows, vl, init, ups, qs, shape):
    """Check every row against the graph its update sees (insertions
    applied before their own search, deletions after) and return the rows'
    count and embedding-set hash per update and query."""
    add, v1, v2, lab = (np.asarray(a) for a in ups)
    E = {}
    for a, b, l in zip(*init):
        E[tuple(sorted((int(a), int(b))))] = int(l)
    by_update = {}
    for r in rows:
        by_update.setdefault(int(r[0]), []).append(r)
    got_cnt = np.zeros(shape, np.uint64)
    got_hash = np.zeros(shape, np.uint64)
    for t in range(len(v1)):
        e = tuple(sorted((int(v1[t]), int(v2[t]))))
        if add[t] and e not in E:
            E[e] = int(lab[t])
        seen = set()
        for r in by_update.get(t, []):
            qi = int(r[1])
            k = len(qs[qi].vlabels)
            mm = [int(x) for x in r[2:2 + k]]
            assert all(int(x) == NONE for x in r[2 + k:])
            _check_row(mm, qs[qi], vl, E, upd=e)
            key = (qi, tuple(mm))
            assert key not in seen, ("repeated row", t, key)
            seen.add(key)
            got_cnt[t, qi] += 1
            got_hash[t, qi] = (int(got_hash[t, qi]) + _row_hash(mm)) & M64
        if not add[t]:
            E.pop(e, None)
    assert not set(by_update) - set(range(len(v1)))
    return got_cnt, got_hash


def _case(seed, nv=40, p=0.3, n_vl=2, n_el=2, n_upd=120, ks=(4, 5)):
    rng = np.random.default_rng(seed)
    vl, s, d, el = random_graph(nv, p, n_vl, n_el, rng)
    init, ups = mixed_stream_small(s, d, el, nv, n_upd, 0.7, n_el, rng)
    qs = [random_connected_query(k, 1, n_vl, n_el, rng) for k in ks]
    return vl, init, ups, qs


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"])
def test_rows_are_each_updates_embeddings(oracle_lib, algo):
    from paracosm_amd.engine import Matching, make_updates
    O = oracle_lib
    vl, init, ups, qs = _case(41, p=0.4, n_el=1)
    with Matching(algo, qs[0], vl, *init, want_hash=True, print_enum=True) as m:
        for q in qs[1:]:
            m.register_query(q)
        m.Preprocessing()
        pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
        rows, found = m.embeddings()
    cnt = pos + neg
    print(algo, "rows", len(rows), "found", found, "matches", int(cnt.sum()), flush=True)
    assert found == int(cnt.sum()) == len(rows) > 0
    # the oracle's counts and hashes per update and query
    for j, q in enumerate(qs):
        ref = O.run("graphflow", O.OGraph(vl, *init), O.OGraph.from_query(q), *ups,
                    want_hash=True)
        np.testing.assert_array_equal(cnt[:, j], ref.pos + ref.neg)
        np.testing.assert_array_equal(hsh[:, j], ref.hash)
    # every row valid in the graph its update sees, none repeated
    got_cnt, got_hash = _replay(rows, vl, init, ups, qs, cnt.shape)
    np.testing.assert_array_equal(got_cnt, cnt)
    np.testing.assert_array_equal(got_hash, hsh)


@pytest.mark.parametrize("ks,dense", [((2,), False), ((3,), False), ((4,), False),
                                      ((6,), True)])
def test_initial_rows_equal_brute_force(oracle_lib, ks, dense):
    """InitialMatching rows over every code path: k = 2 (seed kernel), 3 (the
    position-k-1 loop), 4 and 6 (the fused last position)."""
    from paracosm_amd.engine import Matching
    O = oracle_lib
    kw = dict(p=0.5, n_vl=1) if dense else dict(p=0.4)
    vl, init, ups, qs = _case(42, nv=30, n_el=1, ks=ks, **kw)
    q = qs[0]
    want, want_h = O.brute_count(O.OGraph(vl, *init), O.OGraph.from_query(q), want_hash=True)
    with Matching("parallel_graphflow", q, vl, *init, print_enum=True) as m:
        m.Preprocessing()
        n0 = m.InitialMatching()
        rows, found = m.embeddings()
    print("initial k", ks, "matches", n0, "rows", len(rows), flush=True)
    assert n0 == want == found == len(rows) > 0
    E = {tuple(sorted((int(a), int(b)))): int(l) for a, b, l in zip(*init)}
    k = len(q.vlabels)
    seen = set()
    h = 0
    for r in rows:
        assert int(r[0]) == 0 and int(r[1]) == 0
        mm = [int(x) for x in r[2:2 + k]]
        _check_row(mm, q, vl, E)
        assert tuple(mm) not in seen
        seen.add(tuple(mm))
        h = (h + _row_hash(mm)) & M64
    assert h == want_h


def test_row_cap_keeps_the_count():
    """More embeddings than enum_cap: every one is counted, enum_cap rows
    are kept, and each kept row is valid."""
    from paracosm_amd.engine import Matching, make_updates
    vl, init, ups, qs = _case(43, p=0.4, n_el=1, ks=(4,))
    with Matching("parallel_graphflow", qs[0], vl, *init, print_enum=True, enum_cap=7) as m:
        m.Preprocessing()
        pos, neg, _, _ = m.apply_batch(make_updates(*ups))
        rows, found = m.embeddings()
    total = int((pos + neg).sum())
    assert total > 7
    assert found == total and len(rows) == 7
    got_cnt, _ = _replay(rows, vl, init, ups, qs, pos.shape)
    assert (got_cnt <= pos + neg).all() and int(got_cnt.sum()) == 7


def test_plain_path_rows_follow_each_call():
    """csm_add_edge / csm_remove_edge: the rows are those of the last call."""
    from paracosm_amd.engine import Matching
    vl, init, ups, qs = _case(44, p=0.4, n_el=1, ks=(4,))
    add, v1, v2, lab = (np.asarray(a) for a in ups)
    with Matching("parallel_graphflow", qs[0], vl, *init, print_enum=True) as m:
        m.Preprocessing()
        checked = 0
        for t in range(40):
            if add[t]:
                c = m.AddEdge(int(v1[t]), int(v2[t]), int(lab[t]))
            else:
                c = m.RemoveEdge(int(v1[t]), int(v2[t]))
            rows, found = m.embeddings()
            assert found == c == len(rows)
            assert (rows[:, 0] == 0).all() if len(rows) else True
            checked += c > 0
        assert checked > 0


def test_rows_need_print_enum():
    from paracosm_amd.engine import CsmError, Matching
    vl, init, ups, qs = _case(45, ks=(3,))
    with Matching("parallel_graphflow", qs[0], vl, *init) as m:
        m.Preprocessing()
        with pytest.raises(CsmError):
            m.embeddings()


def test_cli_print_enum_lines(tmp_path, oracle_lib):
    """The drop-in CLI with --print-enum 1: one line per embedding of the
    initial matching and of every update, k vertex ids each."""
    import os
    import re
    import subprocess
    from paracosm_amd import gen
    O = oracle_lib
    vl, init, ups, qs = _case(46, p=0.4, n_el=1, ks=(4,))
    q = qs[0]
    gen.write_query_file(tmp_path / "q.graph", q)
    gen.write_graph_file(tmp_path / "d.graph", vl, *init)
    gen.write_stream_file(tmp_path / "s.stream", *ups)
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "paracosm_amd", "csm")
    p = subprocess.run([cli, "-q", str(tmp_path / "q.graph"), "-d", str(tmp_path / "d.graph"),
                        "-u", str(tmp_path / "s.stream"), "-a", "parallel_graphflow",
                        "--print-enum", "1"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [ln for ln in p.stdout.splitlines() if re.fullmatch(r"(\d+ ){4}", ln)]
    ref = O.run("graphflow", O.OGraph(vl, *init), O.OGraph.from_query(q), *ups,
                want_initial=True)
    print("cli rows", len(lines), "initial", ref.initial, "incremental",
          int((ref.pos + ref.neg).sum()), flush=True)
    assert len(lines) == ref.initial + int((ref.pos + ref.neg).sum()) > 0
    E = {tuple(sorted((int(a), int(b)))): int(l) for a, b, l in zip(*init)}
    for ln in lines[:ref.initial]:
        _check_row([int(x) for x in ln.split()], q, vl, E)
