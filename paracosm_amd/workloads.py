"""The five benchmark workloads (seeded R-MAT graphs, streams and queries)
and the query admission rule.

Query admission.  Random walks on R-MAT gravitate to hubs, and a tree-shaped
8-vertex query through hubs can have 10^8-10^9 embeddings per update (seen on
the scale-20 workload), which no CPU or GPU enumerates over a million updates.
A candidate query is therefore admitted only if its exact number of
embeddings created or destroyed over the first ``sample`` updates of the
stream is at most ``budget_per_update * sample``.  The rule depends only on
exact counts, so any correct engine admits the same queries; candidates are
drawn in seeded order and the first ones admitted are kept.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import gen


@dataclasses.dataclass(frozen=True)
class Config:
    name: str
    scale: int
    edge_factor: int
    n_vlabels: int
    vlabel_dist: str
    n_elabels: int
    stream: str              # "insert" (held-out fraction) or "mixed"
    frac_stream: float = 0.1
    n_updates: int = 0
    p_insert: float = 1.0
    query_sizes: Tuple[int, ...] = (6,)
    per_size: int = 5
    densities: Tuple[Optional[str], ...] = (None,)
    description: str = ""


CONFIGS: Dict[str, Config] = {
    "c1": Config("c1", 14, 8, 16, "uniform", 1, "insert", frac_stream=0.1, query_sizes=(6,),
                 per_size=5, densities=(None,),
                 description="rmat14-ef8-uniform16-el1-10%insert-5q-k6"),
    "c2": Config("c2", 20, 16, 32, "zipf", 2, "mixed", n_updates=1_000_000, p_insert=0.8,
                 query_sizes=(8,), per_size=5, densities=("sparse", "dense"),
                 description="rmat20-ef16-zipf32-el2-1Mupd-80/20-10q-k8"),
    "c3": Config("c3", 20, 16, 32, "zipf", 2, "mixed", n_updates=100_000, p_insert=1.0,
                 query_sizes=(12, 16, 24), per_size=20, densities=(None,),
                 description="rmat20-ef16-zipf32-el2-100kins-60q-k12/16/24"),
    "c4": Config("c4", 21, 16, 32, "zipf", 2, "mixed", n_updates=1_000_000, p_insert=0.7,
                 query_sizes=tuple(range(6, 17)), per_size=6, densities=(None,),
                 description="rmat21-ef16-zipf32-el2-serving-64q-k6..16"),
    "c5": Config("c5", 23, 16, 64, "zipf", 4, "mixed", n_updates=10_000_000, p_insert=0.5,
                 query_sizes=(10,), per_size=8, densities=("dense",),
                 description="rmat23-ef16-zipf64-el4-10Mupd-50/50-8q-k10"),
}


@dataclasses.dataclass
class Data:
    """A workload without its queries, plus the initial graph's CSR for sampling."""
    cfg: Config
    seed: int
    w: gen.Workload
    off: np.ndarray
    nbr: np.ndarray
    ecsr: np.ndarray


def build(name: str, seed: int = 1, scale: Optional[int] = None,
          n_updates: Optional[int] = None, progress=None) -> Data:
    cfg = CONFIGS[name]
    if scale is not None:
        cfg = dataclasses.replace(cfg, scale=scale)
    if n_updates is not None:
        cfg = dataclasses.replace(cfg, n_updates=n_updates)
    w = gen.make_workload(cfg.scale, cfg.edge_factor, cfg.n_vlabels, cfg.vlabel_dist,
                          cfg.n_elabels, stream=cfg.stream, frac_stream=cfg.frac_stream,
                          n_updates=cfg.n_updates, p_insert=cfg.p_insert, query_sizes=(),
                          queries_per_size=0, seed=seed, name=name, progress=progress)
    off, nbr, ecsr = gen.build_csr(w.nv, w.init_src, w.init_dst, w.init_elab)
    if progress:
        progress("initial CSR built")
    return Data(cfg, seed, w, off, nbr, ecsr)


def candidates(d: Data, k: int, density: Optional[str], count: int, round_: int = 0):
    return gen.sample_queries(d.off, d.nbr, d.ecsr, d.w.vlabels, k, count,
                              d.seed * 1000003 + 7919 * round_ + (0 if density != "dense" else 1),
                              density)


ADMISSION_SAMPLE = 4096
ADMISSION_BUDGET_PER_UPDATE = 3000
ADMISSION_RULE = (f"a candidate query is kept only if its exact count over the stream's first "
                  f"{ADMISSION_SAMPLE} updates is at most {ADMISSION_BUDGET_PER_UPDATE} per "
                  f"update; candidates in seeded order, first admitted kept")


def admit(d: Data, *, sample: int = ADMISSION_SAMPLE,
          budget_per_update: float = ADMISSION_BUDGET_PER_UPDATE,
          algo: str = "parallel_graphflow", max_rounds: int = 16, log=print) -> List[gen.Query]:
    """Admitted queries for every (size, density) slot of the config."""
    from .engine import Matching, make_updates
    w = d.w
    n = min(sample, w.num_updates)
    ups = make_updates(w.upd_add[:n], w.upd_v1[:n], w.upd_v2[:n], w.upd_label[:n])
    out: List[gen.Query] = []
    for k in d.cfg.query_sizes:
        for dens in d.cfg.densities:
            got: List[gen.Query] = []
            for rnd in range(max_rounds):
                cands = candidates(d, k, dens, 4 * d.cfg.per_size, rnd)
                budget = int(budget_per_update * n)
                m = Matching(algo, cands[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                             match_budget=budget)
                for q in cands[1:]:
                    m.register_query(q)
                m.Preprocessing()
                pos, neg, _, _ = m.apply_batch(ups)
                over = m.query_status()
                m.close()
                tot = (pos + neg).sum(axis=0)
                for q, t, o in zip(cands, tot, over):
                    if not o and t <= budget and len(got) < d.cfg.per_size:
                        got.append(q)
                log(f"admission k={k} {dens}: round {rnd}: {len(got)}/{d.cfg.per_size} "
                    f"(sample totals {sorted(int(x) for x in tot)[:6]}...)")
                if len(got) >= d.cfg.per_size:
                    break
            if len(got) < d.cfg.per_size:
                raise RuntimeError(f"admission found only {len(got)} queries for k={k} {dens}")
            out += got
    return out


def query_fingerprint(qs: Sequence[gen.Query]) -> str:
    import hashlib
    h = hashlib.sha256()
    for q in qs:
        for a in (q.vlabels, q.src, q.dst, q.elabels):
            h.update(np.ascontiguousarray(a, np.uint32).tobytes())
    return h.hexdigest()[:16]
