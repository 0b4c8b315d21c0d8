"""Figures of one benchmark workload beside the headline bench: the fraction
of updates the label filter certifies safe (Classify, as
parallel_graphflow.cpp:1990-2016), the device memory the engine holds for the
graph, indexes and standing queries, and the batched update rate of a
plugin over a sample of the stream.

    python -m paracosm_amd.config_report --config c5 --algo parallel_symbi

Prints one JSON line; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

HBM_BYTES = 288 * 10**9   # one MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None) -> int:
    from . import workloads
    from .engine import Matching, make_updates

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c5")
    ap.add_argument("--algo", default="parallel_symbi")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sample", type=int, default=1_000_000, help="stream updates run")
    ap.add_argument("--batch", type=int, default=62500)
    ap.add_argument("--scale", type=int, default=None)
    args = ap.parse_args(argv)

    t0 = time.time()
    d = workloads.build(args.config, seed=args.seed, scale=args.scale, progress=log)
    w = d.w
    log(f"workload {args.config}: {w.nv} vertices, {len(w.init_src)} initial edges, "
        f"{w.num_updates} updates in {time.time() - t0:.1f}s")
    t0 = time.time()
    queries = workloads.admit(d, log=log)
    log(f"admitted {len(queries)} queries in {time.time() - t0:.1f}s")

    t0 = time.time()
    m = Matching(args.algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                 max_batch=args.batch)
    for q in queries[1:]:
        m.register_query(q)
    m.Preprocessing()
    build_s = time.time() - t0
    engine_bytes = m.device_bytes()
    log(f"engine built and preprocessed in {build_s:.1f}s, {engine_bytes} device bytes")

    n = min(args.sample, w.num_updates)
    nq = len(queries)
    d_ups = m.device_buffer(min(n, args.batch) * 16)
    d_cnt = m.device_buffer(min(n, args.batch) * nq * 8)
    d_safe = m.device_buffer(min(n, args.batch))
    safe_total, matches, peak_bytes = 0, 0, engine_bytes
    host_cnt = np.zeros(args.batch * nq, np.uint64)
    host_safe = np.zeros(args.batch, np.uint8)
    t_search = 0.0
    for s in range(0, n, args.batch):
        k = min(args.batch, n - s)
        ups = make_updates(w.upd_add[s:s + k], w.upd_v1[s:s + k], w.upd_v2[s:s + k],
                           w.upd_label[s:s + k])
        d_ups.upload(ups.view(np.uint8))
        m.sync()
        t = time.perf_counter()
        m.apply_batch_device(d_ups.ptr, k, d_cnt.ptr, d_safe.ptr)
        m.sync()
        t_search += time.perf_counter() - t
        d_cnt.download(host_cnt.view(np.uint8)[: k * nq * 8])
        d_safe.download(host_safe[:k])
        safe_total += int(host_safe[:k].sum())
        matches += int(host_cnt[: k * nq].sum())
        peak_bytes = max(peak_bytes, m.device_bytes())
        log(f"batch at {s}: {k} updates, {int(host_safe[:k].sum())} safe, "
            f"{time.perf_counter() - t:.3f}s")
    out = {
        "config": d.cfg.description, "algo": args.algo, "seed": args.seed,
        "vertices": int(w.nv), "initial_edges": int(len(w.init_src)),
        "queries": nq, "query_sizes": [q.k for q in queries],
        "updates_run": n, "batch": args.batch,
        "safe_fraction": round(safe_total / max(n, 1), 6),
        "updates_per_s": round(n / t_search, 1) if t_search else None,
        "matches": matches,
        "engine_device_bytes": int(peak_bytes),
        "fraction_of_288GB": round(peak_bytes / HBM_BYTES, 5),
        "engine_build_s": round(build_s, 2),
        "stats": m.stats(),
    }
    m.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
