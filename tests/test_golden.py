"""The golden fixtures (tests/golden/): their files are intact and
reproducible, and the oracle's per-update counts, embedding-set hashes and
initial counts equal the expected values found by complete enumeration."""
import hashlib
import os

import numpy as np
import pytest

from tests.golden import make_golden
from tests.golden_cases import case_names, load, set_hash

CASES = case_names()


def test_fixtures_present():
    assert len(CASES) == len(make_golden.CASES)


@pytest.mark.parametrize("name", CASES)
def test_fixture_files_are_intact(name):
    d, exp, *_ = load(name)
    for fn, digest in exp["sha256"].items():
        with open(os.path.join(d, fn), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == digest, fn
    assert len(exp["positive"]) == len(exp["negative"]) == exp["updates"]
    assert [len(e) for e in exp["embeddings"]] == [p + n for p, n in
                                                   zip(exp["positive"], exp["negative"])]
    assert exp["total_positive"] + exp["total_negative"] > 0


def test_generator_reproduces_the_fixtures(tmp_path):
    spec = make_golden.CASES[0]
    make_golden.make_case(*spec, out_dir=str(tmp_path))
    for fn in ("query.graph", "data.graph", "stream.txt", "expected.json"):
        a = open(os.path.join(make_golden.HERE, spec[0], fn)).read()
        b = open(os.path.join(tmp_path, spec[0], fn)).read()
        assert a == b, fn


@pytest.mark.parametrize("algo", ["graphflow", "symbi", "parallel_graphflow"])
@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_fixture(oracle_lib, name, algo):
    """Counts per update from every oracle formulation; the embedding-set
    hash and the initial count from the sequential ones (the oracle's
    parallel GraphFlow shape counts only); SymBi's index checked against a
    from-scratch build after every update."""
    O = oracle_lib
    _, exp, q, vl, init, ups = load(name)
    seq = algo != "parallel_graphflow"
    r = O.run(algo, O.OGraph(vl, *init), O.OGraph.from_query(q), *ups, want_hash=seq,
              want_initial=seq, check_index=algo == "symbi", threads=2)
    np.testing.assert_array_equal(r.pos, np.asarray(exp["positive"], np.uint64))
    np.testing.assert_array_equal(r.neg, np.asarray(exp["negative"], np.uint64))
    if seq:
        np.testing.assert_array_equal(
            r.hash, np.asarray([set_hash(e) for e in exp["embeddings"]], np.uint64))
        assert r.initial == exp["initial_embeddings"]
