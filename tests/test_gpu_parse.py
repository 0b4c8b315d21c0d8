"""Device text ingestion against a line-by-line Python restatement of the
reference grammar (Graph::LoadFromFile, graph.cpp:217-249, and
Graph::LoadUpdateStream, graph.cpp:251-285) on the same bytes: records equal
field by field, malformed lines reported with their 1-based number."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U32MAX = 4294967295


def py_parse(text: bytes, graph: bool):
    """(type, is_add, id1, id2, label) per record, or the first bad line.
    Graph files take the type from the line's first char ('t' header, 'v'
    vertex, anything else an edge; graph.cpp:226-246) and the first field may
    follow it without a space; streams take a whole token ("v"/"-v" vertex,
    anything else an edge with is_add = token == "e"; graph.cpp:263-281)."""
    out = []
    for no, raw in enumerate(text.split(b"\n"), 1):
        line = raw.replace(b"\r", b" ").replace(b"\t", b" ")
        body = line.lstrip(b" ")
        if not body:
            continue
        if graph:
            t, rest = body[:1], body[1:]
            if t == b"t":
                continue
            typ, add = (ord("v"), 1) if t == b"v" else (ord("e"), 1)
            fields = rest.split()
        else:
            toks = body.split()
            t, fields = toks[0], toks[1:]
            if t in (b"v", b"-v"):
                typ, add = ord("v"), int(t == b"v")
            else:
                typ, add = ord("e"), int(t == b"e")
        need = 2 if typ == ord("v") else 3
        if len(fields) != need or not all(x.isdigit() for x in fields):
            return no
        vals = [int(x) for x in fields]
        if any(v > U32MAX for v in vals):
            return no
        if need == 2:
            out.append((typ, add, vals[0], 0, vals[1]))
        else:
            out.append((typ, add, vals[0], vals[1], vals[2]))
    return out


def as_tuples(rec):
    return [(int(r["type"]), int(r["is_add"]), int(r["id1"]), int(r["id2"]), int(r["label"]))
            for r in rec]


def random_text(rng, n, graph):
    lines = []
    if graph:
        lines.append("t # %d" % n)
    for i in range(n):
        sp = " " * int(rng.integers(1, 3)) if rng.random() < 0.2 else " "
        kind = rng.random()
        if kind < 0.3:
            tok = "v" if graph or rng.random() < 0.5 else "-v"
            s = f"{tok}{sp}{rng.integers(0, 1 << 20)}{sp}{rng.integers(0, 64)}"
        else:
            if graph:
                tok = "e" if rng.random() < 0.9 else "x"
            else:
                tok = "e" if rng.random() < 0.6 else ("-e" if rng.random() < 0.8 else "ev")
            s = f"{tok}{sp}{rng.integers(0, 1 << 32)}{sp}{rng.integers(0, 1 << 20)}{sp}{rng.integers(0, 4)}"
        if rng.random() < 0.05:
            s = "\t" + s + " \r"
        lines.append(s)
        if rng.random() < 0.03:
            lines.append("")
    return ("\n".join(lines)).encode()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("seed", range(3))
def test_parse_matches_restatement(graph, seed):
    from paracosm_amd.engine import PARSE_GRAPH, PARSE_STREAM, parse_text
    rng = np.random.default_rng(seed)
    n = [10, 5000, 70000][seed]
    text = random_text(rng, n, graph)
    if seed == 1:
        text += b"\n"          # with and without a final newline
    ref = py_parse(text, graph)
    rec = parse_text(text, PARSE_GRAPH if graph else PARSE_STREAM)
    print("parse", graph, seed, "bytes", len(text), "records", len(rec), flush=True)
    assert as_tuples(rec) == ref


@pytest.mark.parametrize("case,line", [
    (b"e 1 2\n", 1),                        # missing field
    (b"v 1 2 3\n", 1),                      # extra field
    (b"e 1 2 4294967296\n", 1),             # overflow
    (b"e 1 2 3\n\ne 1 2a 3\n", 3),          # non-digit
    (b"e 1 2 3\n-t 1\n", 2),                # an edge record split over lines
])
def test_parse_reports_malformed_line(case, line):
    from paracosm_amd.engine import PARSE_STREAM, ParseError, parse_text
    with pytest.raises(ParseError) as ei:
        parse_text(case, PARSE_STREAM)
    assert ei.value.line == line


def test_unknown_tokens_are_edges():
    """Any token other than v/-v is an edge record, an insertion only for
    "e" (LoadUpdateStream, graph.cpp:276-281); in graph files any leading
    char other than t/v is an edge (LoadFromFile, graph.cpp:241-246)."""
    from paracosm_amd.engine import PARSE_GRAPH, PARSE_STREAM, ParseError, parse_text
    E = ord("e")
    assert as_tuples(parse_text(b"x 1 2 3\nev 4 5 6\n-t 7 8 9\ne 1 2 0\n+e 3 4 1",
                                PARSE_STREAM)) == [
        (E, 0, 1, 2, 3), (E, 0, 4, 5, 6), (E, 0, 7, 8, 9), (E, 1, 1, 2, 0), (E, 0, 3, 4, 1)]
    assert as_tuples(parse_text(b"t # 2\nx 1 2 3\nv0 5\ne1 2 3\n", PARSE_GRAPH)) == [
        (E, 1, 1, 2, 3), (ord("v"), 1, 0, 0, 5), (E, 1, 1, 2, 3)]
    # the reference stops reading where a field is not a number; here that
    # line is an error
    with pytest.raises(ParseError) as ei:
        parse_text(b"e 1 2 3\n-e 1 2 3\n", PARSE_GRAPH)
    assert ei.value.line == 2


def test_parse_edge_cases():
    from paracosm_amd.engine import PARSE_GRAPH, PARSE_STREAM, parse_text
    text = b"t # 3\r\nv 0 4294967295\r\n\r\n  e\t0  1 7  \nv 1 2"
    assert as_tuples(parse_text(text, PARSE_GRAPH)) == [
        (ord("v"), 1, 0, 0, U32MAX), (ord("e"), 1, 0, 1, 7), (ord("v"), 1, 1, 0, 2)]
    assert as_tuples(parse_text(b"", PARSE_STREAM)) == []
    assert as_tuples(parse_text(b"\n\n  \n", PARSE_STREAM)) == []
    assert as_tuples(parse_text(b"-v 5 1\n-e 1 2 0", PARSE_STREAM)) == [
        (ord("v"), 0, 5, 0, 1), (ord("e"), 0, 1, 2, 0)]


def test_device_text_feeds_the_engine(oracle_lib):
    """A stream text resident in HBM is parsed into device records and applied
    with csm_apply_batch_device; counts equal the host path's."""
    from paracosm_amd.engine import PARSE_STREAM, make_updates
    from tests.helpers import mixed_stream_small, random_connected_query, random_graph
    rng = np.random.default_rng(77)
    vl, s, d, el = random_graph(50, 0.3, 2, 2, rng)
    init, ups = mixed_stream_small(s, d, el, 50, 150, 0.65, 2, rng)
    q = random_connected_query(5, 1, 2, 2, rng)
    add, v1, v2, lab = ups
    text = "\n".join(("e" if a else "-e") + f" {x} {y} {l}"
                     for a, x, y, l in zip(add, v1, v2, lab)).encode()
    ref = oracle_lib.run("graphflow", oracle_lib.OGraph(vl, *init),
                         oracle_lib.OGraph.from_query(q), *ups)
    from paracosm_amd.engine import Matching
    with Matching("parallel_graphflow", q, vl, *init) as m:
        m.Preprocessing()
        n = len(v1)
        d_text = m.device_buffer(len(text))
        d_text.upload(np.frombuffer(text, np.uint8))
        d_rec = m.device_buffer(16 * n)
        got = m.parse_text_device(d_text.ptr, len(text), PARSE_STREAM, d_rec.ptr, n)
        assert got == n
        host = np.zeros(n, make_updates(*ups).dtype)
        d_rec.download(host.view(np.uint8))
        np.testing.assert_array_equal(host, make_updates(*ups))
        d_cnt = m.device_buffer(8 * n)
        m.apply_batch_device(d_rec.ptr, n, d_cnt.ptr)
        cnt = np.zeros(n, np.uint64)
        d_cnt.download(cnt.view(np.uint8))
    np.testing.assert_array_equal(cnt, np.where(add == 1, ref.pos, ref.neg))


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
def test_engine_from_graph_text_equals_host_build(oracle_lib, algo):
    """A data graph file parsed on the device and built into the CSR there
    (csm_parse_text_to_device + csm_create_from_records) gives the engine
    that host arrays give: same counts per update, same SymBi index.  The
    file has what LoadFromFile tolerates: a header, a vertex declared twice
    (the first label stays), a vertex never declared (no label), a repeated
    edge (the first copy stays) and an edge line with another leading
    char."""
    from paracosm_amd.engine import Matching, make_updates, records_to_graph, parse_text, PARSE_GRAPH
    from tests.helpers import mixed_stream_small, random_connected_query, random_graph
    rng = np.random.default_rng(606)
    nv = 60
    vl, s, d, el = random_graph(nv, 0.35, 2, 1, rng)
    init, ups = mixed_stream_small(s, d, el, nv, 200, 0.7, 1, rng)
    q = random_connected_query(4, 1, 2, 1, rng)
    lines = ["t # 0"] + [f"v {v} {vl[v]}" for v in range(nv - 1)]   # vertex nv-1 undeclared
    lines.append(f"v 3 {(int(vl[3]) + 1) % 2}")                       # ignored re-declaration
    src, dst, lab = init
    for i, (a, b, c) in enumerate(zip(src, dst, lab)):
        lines.append(f"{'x' if i % 7 == 0 else 'e'} {a} {b} {c}")
    lines.append(f"e {src[0]} {dst[0]} 7")                            # repeated edge, other label
    text = ("\n".join(lines) + "\n").encode()
    hv, hs, hd, he = records_to_graph(parse_text(text, PARSE_GRAPH))
    u = make_updates(*ups)
    keep = (ups[1] != nv - 1) & (ups[2] != nv - 1)   # updates of the unlabelled vertex are refused
    u = u[keep]
    with Matching(algo, q, hv, hs, hd, he) as ref, Matching.from_text(algo, q, text) as m:
        ref.Preprocessing()
        m.Preprocessing()
        assert m.GetMemoryCost() == ref.GetMemoryCost()
        if algo == "parallel_symbi":
            for what in (0, 1, 2, 3, 4):
                rows = 1 if what == 0 else q.k
                np.testing.assert_array_equal(m.index_dump(0, what, nv, rows),
                                              ref.index_dump(0, what, nv, rows))
        p1, n1, _, _ = m.apply_batch(u)
        p2, n2, _, _ = ref.apply_batch(u)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(n1, n2)
    assert (p1.sum() + n1.sum()) > 0
