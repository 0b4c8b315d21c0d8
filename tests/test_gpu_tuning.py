"""The autotuner picks the search settings (split chunk, spill budget) per
query size class: queries of up to 8, 12, 16 and 32 vertices.  On its tuning
batch it times each class's candidates with the other classes' queries
skipped, keeps the fastest per class, and then counts the batch in full; the
counts match the oracle on every batch, during and after tuning.
"""
import numpy as np
import pytest

from tests.helpers import walk_query

pytestmark = pytest.mark.gpu

CANDIDATES = {(64, 1024), (64, 2048), (64, 512), (64, 384), (64, 320), (64, 256), (32, 320)}
BATCH = 4096   # the smallest batch the autotuner scores


def _case(seed=5, nv=3000, m=24000, n_vl=8, n_el=2, n_ins=3 * BATCH):
    """A random graph whose held-out edges form an insertion-only stream of
    n_ins updates, and random-walk queries of 5, 10 and 14 vertices (size
    classes 0, 1 and 2) on the final graph."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, nv, size=3 * m)
    b = rng.integers(0, nv, size=3 * m)
    keep = a != b
    pairs = np.unique(np.stack([np.minimum(a, b), np.maximum(a, b)], 1)[keep], axis=0)
    pairs = pairs[rng.permutation(len(pairs))[:m]]
    src, dst = pairs[:, 0].astype(np.uint32), pairs[:, 1].astype(np.uint32)
    el = rng.integers(0, n_el, size=m).astype(np.uint32)
    vl = rng.integers(0, n_vl, size=nv).astype(np.uint32)
    queries = [walk_query(vl, src, dst, el, k, rng) for k in (5, 10, 14)]
    n0 = m - n_ins
    init = (src[:n0], dst[:n0], el[:n0])
    ups = (np.ones(n_ins, np.uint8), src[n0:], dst[n0:], el[n0:])
    return vl, init, ups, queries


def test_settings_are_tuned_per_query_size(oracle_lib):
    from paracosm_amd.engine import Matching, make_updates
    vl, init, ups, queries = _case()
    O = oracle_lib
    dg = O.OGraph(vl, *init)
    ref = np.stack([O.run("graphflow", dg, O.OGraph.from_query(q), *ups).pos for q in queries], 1)
    print("oracle totals per query", ref.sum(axis=0), flush=True)
    assert (ref.sum(axis=0) > 0).all()
    u = make_updates(*ups)
    m = Matching("parallel_graphflow", queries[0], vl, *init, max_batch=BATCH)
    for q in queries[1:]:
        m.register_query(q)
    m.Preprocessing()
    try:
        for i in range(3):
            pos, _, _, _ = m.apply_batch(u[i * BATCH:(i + 1) * BATCH])
            np.testing.assert_array_equal(pos, ref[i * BATCH:(i + 1) * BATCH])
            st = m.stats()
            print("batch", i, "tuning", st["tuning"], st["size_class_settings"], flush=True)
            # batch 0 settles, batch 1 tunes
            assert st["tuning"] == (1 if i == 0 else 0)
        cls = st["size_class_settings"]
        for c in (0, 1, 2):
            assert tuple(cls[c]) in CANDIDATES
        assert tuple(cls[3]) == (64, 1024)   # no query of 17-32 vertices: untouched
        assert (st["split_chunk"], st["spill_budget"]) == tuple(cls[0])
    finally:
        m.close()


def test_fixed_settings_skip_tuning():
    """A caller that fixes either knob gets it for every size class."""
    from paracosm_amd.engine import Matching
    vl, init, _, queries = _case()
    m = Matching("parallel_graphflow", queries[0], vl, *init, split_chunk=16)
    for q in queries[1:]:
        m.register_query(q)
    m.Preprocessing()
    try:
        st = m.stats()
        assert st["tuning"] == 0
        assert all(tuple(s) == (16, 1024) for s in st["size_class_settings"])
    finally:
        m.close()
