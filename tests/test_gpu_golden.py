"""The golden fixtures (tests/golden/) through the HIP engine: per-update
counts, embedding-set hashes, the embeddings themselves (--print-enum rows),
the initial count, the plain per-update path, and the drop-in CLI reading the
fixture files.  Expected values come from complete enumeration
(tests/golden/make_golden.py), not from the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.golden_cases import case_names, load, set_hash

pytestmark = pytest.mark.gpu

CASES = case_names()
ALGOS = ["parallel_graphflow", "parallel_symbi", "parallel_turboflux"]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "paracosm_amd", "csm")


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("name", CASES)
def test_batch_equals_fixture(name, algo):
    from paracosm_amd.engine import Matching, make_updates
    _, exp, q, vl, init, ups = load(name)
    k = exp["query_vertices"]
    with Matching(algo, q, vl, *init, want_hash=True, print_enum=True) as m:
        m.Preprocessing()
        assert m.InitialMatching() == exp["initial_embeddings"]
        pos, neg, _, hsh = m.apply_batch(make_updates(*ups))
        rows, found = m.embeddings()
    np.testing.assert_array_equal(pos[:, 0], np.asarray(exp["positive"], np.uint64))
    np.testing.assert_array_equal(neg[:, 0], np.asarray(exp["negative"], np.uint64))
    np.testing.assert_array_equal(
        hsh[:, 0], np.asarray([set_hash(e) for e in exp["embeddings"]], np.uint64))
    # the rows of each update are exactly its embeddings
    assert found == len(rows) == exp["total_positive"] + exp["total_negative"]
    got = [set() for _ in range(exp["updates"])]
    for r in rows:
        got[int(r[0])].add(tuple(int(x) for x in r[2:2 + k]))
    for i, e in enumerate(exp["embeddings"]):
        assert got[i] == {tuple(x) for x in e}, (name, algo, i)


@pytest.mark.parametrize("name", CASES)
def test_plain_path_equals_fixture(name):
    from paracosm_amd.engine import Matching
    _, exp, q, vl, init, ups = load(name)
    add, v1, v2, lab = ups
    with Matching("parallel_graphflow", q, vl, *init) as m:
        m.Preprocessing()
        got = [m.AddEdge(int(a), int(b), int(l)) if t else m.RemoveEdge(int(a), int(b))
               for t, a, b, l in zip(add, v1, v2, lab)]
    want = [p + n for p, n in zip(exp["positive"], exp["negative"])]
    assert got == want


def _num(out, what):
    m = re.search(r"^(\d+) " + re.escape(what), out, re.M)
    assert m, (what, out)
    return int(m.group(1))


@pytest.mark.parametrize("algo", ["parallel_graphflow", "parallel_symbi"])
@pytest.mark.parametrize("name", CASES)
def test_cli_reads_fixture_files(name, algo):
    d, exp, *_ = load(name)
    cmd = [CLI, "-a", algo, "-q", os.path.join(d, "query.graph"), "-d",
           os.path.join(d, "data.graph"), "-u", os.path.join(d, "stream.txt")]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert _num(p.stdout, "matches.") == exp["initial_embeddings"]
    assert _num(p.stdout, "positive matches.") == exp["total_positive"]
    assert _num(p.stdout, "negative matches.") == exp["total_negative"]
    changed = sum(1 for pp, nn in zip(exp["positive"], exp["negative"]) if pp + nn > 0)
    assert _num(p.stdout, "unsafe updates.") == changed
