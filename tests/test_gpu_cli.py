"""The drop-in CLI (paracosm_amd/csm) against the oracle: report lines,
default per-update mode, the reference's window mode (--compat 1), and the
reference's error exits."""
import os
import re
import subprocess

import numpy as np
import pytest

from paracosm_amd import gen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "paracosm_amd", "csm")


def _files(tmp_path, seed=3, mixed=False):
    if mixed:
        w = gen.make_workload(11, 8, 8, "uniform", 1, stream="mixed", n_updates=2000,
                              p_insert=0.7, query_sizes=(5,), queries_per_size=1, seed=seed)
    else:
        w = gen.make_workload(11, 8, 8, "uniform", 1, stream="insert", frac_stream=0.1,
                              query_sizes=(5,), queries_per_size=1, seed=seed)
    q = w.queries[0]
    gen.write_query_file(tmp_path / "q.graph", q)
    gen.write_graph_file(tmp_path / "d.graph", w.vlabels, w.init_src, w.init_dst, w.init_elab)
    gen.write_stream_file(tmp_path / "s.stream", w.upd_add, w.upd_v1, w.upd_v2, w.upd_label)
    return w, q


def _run(tmp_path, *extra):
    cmd = [CLI, "-q", str(tmp_path / "q.graph"), "-d", str(tmp_path / "d.graph"),
           "-u", str(tmp_path / "s.stream"), *extra]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout


def _num(out, what):
    m = re.search(r"^(\d+) " + re.escape(what), out, re.M)
    assert m, (what, out)
    return int(m.group(1))


@pytest.mark.parametrize("algo,ref_orders", [("parallel_graphflow", 0), ("parallel_symbi", 0),
                                             ("parallel_symbi", 1)])
def test_cli_default_mode_totals(tmp_path, oracle_lib, algo, ref_orders):
    """Totals and report lines; parallel_symbi with --reference-orders 1 runs
    SymBi's search-time order rule."""
    O = oracle_lib
    w, q = _files(tmp_path, mixed=True)
    rc, out = _run(tmp_path, "-a", algo, "--reference-orders", str(ref_orders))
    assert rc == 0, out
    ref = O.run("graphflow", O.OGraph(w.vlabels, w.init_src, w.init_dst, w.init_elab),
                O.OGraph.from_query(q), w.upd_add, w.upd_v1, w.upd_v2, w.upd_label,
                want_initial=True)
    assert _num(out, "matches.") == ref.initial
    assert _num(out, "positive matches.") == int(ref.pos.sum())
    assert _num(out, "negative matches.") == int(ref.neg.sum())
    assert _num(out, "unsafe updates.") == int(((ref.pos + ref.neg) > 0).sum())
    for line in ("Load Graphs:", "Preprocessing:", "Initial Matching:", "Incremental Matching:",
                 "Peak Virtual Memory:", "----------------- End -----------------"):
        assert line in out


def test_cli_compat_mode_replays_window_loop(tmp_path, oracle_lib):
    """--compat 1: only unsafe insertions are applied and searched
    (main.cpp:253-397); deletions and safe insertions are skipped."""
    O = oracle_lib
    w, q = _files(tmp_path, seed=5, mixed=True)
    rc, out = _run(tmp_path, "-a", "parallel_graphflow", "--compat", "1")
    assert rc == 0, out
    qg = O.OGraph.from_query(q)
    keep = np.asarray([bool(a) and not O.classify(qg, int(w.vlabels[x]), int(w.vlabels[y]), int(l))
                       for a, x, y, l in zip(w.upd_add, w.upd_v1, w.upd_v2, w.upd_label)])
    ref = O.run("graphflow", O.OGraph(w.vlabels, w.init_src, w.init_dst, w.init_elab), qg,
                w.upd_add[keep], w.upd_v1[keep], w.upd_v2[keep], w.upd_label[keep])
    assert _num(out, "positive matches.") == int(ref.pos.sum())
    assert _num(out, "negative matches.") == 0
    assert _num(out, "unsafe updates.") == int((ref.pos > 0).sum())
    # the window loop's own update count (it adds 16 for a safe final window)
    unsafe_flags = [not O.classify(qg, int(w.vlabels[x]), int(w.vlabels[y]), int(l))
                    for x, y, l in zip(w.upd_v1, w.upd_v2, w.upd_label)]
    base = n_e = 0
    while base < len(unsafe_flags):
        hit = next((i for i in range(16) if base + i < len(unsafe_flags)
                    and unsafe_flags[base + i]), None)
        step = 16 if hit is None else hit + 1
        base += step
        n_e += step
    assert _num(out, "edge updates.") == n_e


def test_cli_unknown_algorithm(tmp_path):
    _files(tmp_path)
    rc, out = _run(tmp_path, "-a", "no_such_algorithm")
    assert rc == 255 and "Unknown algorithm" in out


def test_cli_deletion_error(tmp_path):
    _files(tmp_path)
    with open(tmp_path / "s.stream", "a") as f:
        f.write("-e 0 0 0\n")  # a self-pair can never be an edge
    rc, out = _run(tmp_path, "-a", "parallel_graphflow")
    assert rc == 255 and "deletion error" in out
