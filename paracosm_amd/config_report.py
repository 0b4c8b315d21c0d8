"""Figures of one benchmark workload beside the headline bench.

--mode rate (default): the fraction of updates the label filter certifies
safe (Classify, as parallel_graphflow.cpp:1990-2016), the device memory the
engine holds for the graph, indexes and standing queries, and the batched
update rate of a plugin over a sample of the stream.

    python -m paracosm_amd.config_report --config c5 --algo parallel_symbi

--mode large: large queries (config c3: 20 random-walk queries each of 12, 16
and 24 vertices on the scale-20 graph, 100k insertions).  Every query runs
alone on its own engine over the stream under a time limit for the whole
stream (the reference's --time-limit); a query succeeds if its search never
hit the limit.  Reported per size: success rate, time and matches, and the
histogram of per-update search-tree sizes (candidate entries tested, log2
buckets).  No query admission rule applies.

    python -m paracosm_amd.config_report --config c3 --mode large --time-limit 60

--mode serve: the serving loop (config c4: scale-21 graph, 70/30 stream, 64
standing random-walk queries of 6-16 vertices, 6 per size in seeded order,
no admission rule).  Requests of log-uniform size in 1..4096 arrive as a
Poisson process at each offered load; batches are admitted by free HBM and
each batch's search runs under a time limit (--time-limit, per batch), so
every request completes.  Reported per load: updates/s, p50/p99 latency,
the fraction of updates whose counts are complete, and per query the
fraction of batches it completed within the limit.

    python -m paracosm_amd.config_report --config c4 --mode serve --time-limit 0.2

--mode orders: the same admitted queries and stream sample under three
matching-order rules: the engine's estimated-cost orders (parallel_symbi),
the GraphFlow plugin's static orders (parallel_graphflow with reference
orders, graphflow.cpp:36-160) and SymBi's search-time rule (parallel_symbi
with reference orders, parrallel_symbi.cpp:4231-4262).  Batches of --batch
updates, each under --time-limit seconds.  Reported per rule: updates/s,
updates complete, and candidate entries tested per update over the updates
every rule completed (whose counts must agree).

    python -m paracosm_amd.config_report --config c2 --mode orders --sample 4000 --batch 1000 --time-limit 20

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
    ap.add_argument("--mode", default="rate", choices=("rate", "large", "serve", "orders"))
    ap.add_argument("--loads", default="1e4,1e5,1e6,1e7",
                    help="serve: offered loads in updates/s")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="serve: seconds of arrivals per offered load")
    ap.add_argument("--queries", type=int, default=64, help="serve: standing queries")
    ap.add_argument("--time-limit", type=float, default=60.0,
                    help="large: seconds per query for the whole stream sample")
    ap.add_argument("--per-size", type=int, default=0,
                    help="large: queries per size (0 = the config's)")
    ap.add_argument("--sizes", default="",
                    help="large: query sizes to run (comma list; default: the config's)")
    args = ap.parse_args(argv)
    if args.mode == "large":
        return large_queries(args)
    if args.mode == "serve":
        return serve_loads(args)
    if args.mode == "orders":
        return order_rules(args)

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


def histogram_log2(x: np.ndarray) -> dict:
    """Counts of zeros ("0") and of values in [2^j, 2^(j+1)) ("2^j")."""
    x = np.asarray(x, np.uint64)
    out = {"0": int((x == 0).sum())}
    nz = x[x > 0].astype(np.float64)
    if nz.size:
        b = np.floor(np.log2(nz)).astype(np.int64)
        for j, c in zip(*np.unique(b, return_counts=True)):
            out[f"2^{int(j)}"] = int(c)
    return out


def large_queries(args) -> int:
    """Config c3: per-query success under a time limit, and the per-update
    search-tree size histogram per query size."""
    from . import workloads
    from .engine import USTAT_TIME_LIMIT, Matching, make_updates
    t0 = time.time()
    d = workloads.build(args.config, seed=args.seed, scale=args.scale, progress=log)
    w = d.w
    n = min(args.sample, w.num_updates)
    log(f"workload {args.config}: {w.nv} vertices, {len(w.init_src)} initial edges, "
        f"{w.num_updates} updates in {time.time() - t0:.1f}s; running {n}")
    ups = make_updates(w.upd_add[:n], w.upd_v1[:n], w.upd_v2[:n], w.upd_label[:n])
    per = args.per_size or d.cfg.per_size
    sizes = {}
    run_sizes = [int(x) for x in args.sizes.split(",") if x] or list(d.cfg.query_sizes)
    for k in run_sizes:
        qs = workloads.candidates(d, k, None, per, 0)
        rows = []
        work_all = []
        for i, q in enumerate(qs):
            tb = time.time()
            m = Matching(args.algo, q, w.vlabels, w.init_src, w.init_dst, w.init_elab,
                         max_batch=args.batch, want_work=True)
            m.Preprocessing()
            build = time.time() - tb
            d_ups = m.device_buffer(min(n, args.batch) * 16)
            d_cnt = m.device_buffer(min(n, args.batch) * 8)
            cnt = np.zeros(min(n, args.batch), np.uint64)
            matches, partial, done = 0, 0, 0
            t_run = time.perf_counter()
            work = []
            for s0 in range(0, n, args.batch):
                kk = min(args.batch, n - s0)
                left = args.time_limit - (time.perf_counter() - t_run)
                m.set_time_limit(max(left, 1e-9))
                d_ups.upload(ups[s0:s0 + kk].view(np.uint8))
                hit = m.apply_batch_device(d_ups.ptr, kk, d_cnt.ptr, partial_ok=True)
                d_cnt.download(cnt.view(np.uint8)[: kk * 8])
                matches += int(cnt[:kk].sum())
                st = m.update_status(kk)
                partial += int(((st & USTAT_TIME_LIMIT) != 0).sum())
                work.append(m.update_work(kk))
                done += kk
                if hit:
                    break
            elapsed = time.perf_counter() - t_run
            m.close()
            ok = partial == 0 and done == n
            wk = np.concatenate(work) if work else np.zeros(0, np.uint64)
            work_all.append(wk)
            rows.append({"query": i, "edges": q.num_edges, "success": ok,
                         "seconds": round(elapsed, 3), "updates_searched": done,
                         "matches": matches, "partial_updates": partial,
                         "max_update_work": int(wk.max()) if wk.size else 0,
                         "engine_build_s": round(build, 2)})
            log(f"k={k} query {i}: {rows[-1]}")
        wk = np.concatenate(work_all) if work_all else np.zeros(0, np.uint64)
        succ = sum(r["success"] for r in rows)
        sizes[str(k)] = {
            "queries": len(rows), "success": succ, "success_rate": round(succ / max(len(rows), 1), 3),
            "median_seconds": float(np.median([r["seconds"] for r in rows])) if rows else None,
            "matches_successful": int(sum(r["matches"] for r in rows if r["success"])),
            "update_work_histogram_log2": histogram_log2(wk),
            "per_query": rows,
        }
    out = {"config": d.cfg.description, "mode": "large", "algo": args.algo, "seed": args.seed,
           "updates": n, "batch": args.batch, "time_limit_s_per_query": args.time_limit,
           "work_unit": "candidate entries tested by the search of one update",
           "sizes": sizes}
    print(json.dumps(out), flush=True)
    return 0


ORDER_RULES = (
    ("estimated_cost", "parallel_symbi", False),
    ("graphflow_static", "parallel_graphflow", True),
    ("symbi_search_time", "parallel_symbi", True),
)


def order_rules(args) -> int:
    """Search effort and rate of the three order rules on one workload."""
    from . import workloads
    from .engine import USTAT_TIME_LIMIT, Matching, make_updates
    t0 = time.time()
    d = workloads.build(args.config, seed=args.seed, scale=args.scale, progress=log)
    w = d.w
    n = min(args.sample, w.num_updates)
    log(f"workload {args.config}: {w.nv} vertices, {len(w.init_src)} initial edges; running {n}")
    queries = workloads.admit(d, log=log)
    log(f"admitted {len(queries)} queries in {time.time() - t0:.1f}s")
    ups = make_updates(w.upd_add[:n], w.upd_v1[:n], w.upd_v2[:n], w.upd_label[:n])
    res = {}
    for name, algo, ref in ORDER_RULES:
        m = Matching(algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                     max_batch=args.batch, want_work=True, reference_orders=ref)
        for q in queries[1:]:
            m.register_query(q)
        m.Preprocessing()
        work, done, cnts = [], [], []
        t_all = 0.0
        for s0 in range(0, n, args.batch):
            kk = min(args.batch, n - s0)
            m.set_time_limit(args.time_limit)
            t = time.perf_counter()
            pos, neg, _, _ = m.apply_batch(ups[s0:s0 + kk], partial_ok=True)
            t_all += time.perf_counter() - t
            st = m.update_status(kk)
            work.append(m.update_work(kk))
            done.append((st & USTAT_TIME_LIMIT) == 0)
            cnts.append((pos + neg).sum(axis=1))
            log(f"{name}: batch at {s0}: {time.perf_counter() - t:.3f}s, "
                f"{int(done[-1].sum())}/{kk} complete")
        m.close()
        res[name] = {"algo": algo, "reference_orders": ref, "seconds": t_all,
                     "work": np.concatenate(work), "done": np.concatenate(done),
                     "counts": np.concatenate(cnts)}
    all_done = np.logical_and.reduce([r["done"] for r in res.values()])
    base = res[ORDER_RULES[0][0]]["counts"][all_done]
    rules = {}
    for name, r in res.items():
        wk = r["work"][all_done]
        rules[name] = {
            "algo": r["algo"], "reference_orders": r["reference_orders"],
            "updates_per_s": round(n / r["seconds"], 1) if r["seconds"] else None,
            "updates_complete": int(r["done"].sum()),
            "candidates_per_update": round(float(wk.mean()), 1) if wk.size else None,
            "max_update_candidates": int(wk.max()) if wk.size else 0,
            "counts_equal_estimated_cost": bool(np.array_equal(r["counts"][all_done], base)),
        }
    out = {"config": d.cfg.description, "mode": "orders", "seed": args.seed, "updates": n,
           "batch": args.batch, "time_limit_s_per_batch": args.time_limit,
           "queries": len(queries), "query_fingerprint": workloads.query_fingerprint(queries),
           "updates_complete_in_all": int(all_done.sum()),
           "work_unit": "candidate entries tested by the search of one update (all queries)",
           "rules": rules}
    print(json.dumps(out), flush=True)
    return 0


def serve_loads(args) -> int:
    """Config c4: the serving loop over all standing queries at several
    offered loads, each on a fresh engine from the initial graph."""
    from . import workloads
    from .engine import Matching, make_updates
    from .serve import Server, poisson_requests
    t0 = time.time()
    d = workloads.build(args.config, seed=args.seed, scale=args.scale, progress=log)
    w = d.w
    log(f"workload {args.config}: {w.nv} vertices, {len(w.init_src)} initial edges, "
        f"{w.num_updates} updates in {time.time() - t0:.1f}s")
    per = args.per_size or d.cfg.per_size
    qs = []
    for k in d.cfg.query_sizes:
        qs += workloads.candidates(d, k, None, per, 0)
    qs = qs[: args.queries]
    log(f"{len(qs)} standing queries, k = {[q.k for q in qs]}")
    ups = make_updates(w.upd_add, w.upd_v1, w.upd_v2, w.upd_label)
    loads = []
    for tok in [x for x in args.loads.split(",") if x]:
        load = float(tok)
        n = int(min(args.serve_seconds * load, w.num_updates))
        tb = time.time()
        m = Matching(args.algo, qs[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                     max_batch=args.batch, time_limit_s=args.time_limit, split_chunk=64,
                     spill_budget=512)
        for q in qs[1:]:
            m.register_query(q)
        m.Preprocessing()
        build = time.time() - tb
        reqs = poisson_requests(ups[:n], load, seed=args.seed)
        srv = Server(m, max_batch=args.batch)
        st = srv.serve(reqs)
        budget = srv.hbm_budget
        srv.close()
        engine_bytes = m.device_bytes()
        m.close()
        qpb = np.asarray(st.query_partial_batches, np.float64)
        loads.append({
            "offered_updates_per_s": load, "updates": st.updates, "requests": st.requests,
            "batches": st.batches, "mean_batch": round(st.mean_batch, 1),
            "updates_per_s": round(st.updates_per_s, 1),
            "achieved_over_offered": round(st.updates_per_s / load, 3),
            "p50_update_latency_ms": round(st.p50_ms, 3),
            "p99_update_latency_ms": round(st.p99_ms, 3),
            "complete_update_fraction": round(1.0 - st.partial_updates / max(st.updates, 1), 5),
            "queries_complete_in_every_batch": int((qpb == 0).sum()),
            "query_batch_completion": [round(1.0 - x / max(st.batches, 1), 4) for x in qpb],
            "matches_counted": st.matches, "overlapped_uploads": st.overlapped_uploads,
            "hbm_budget": budget, "engine_device_bytes": engine_bytes,
            "refused_for_memory": st.refused_for_memory, "engine_build_s": round(build, 2)})
        log(f"load {load:g}: " + json.dumps({k: v for k, v in loads[-1].items()
                                             if k != "query_batch_completion"}))
    out = {"config": d.cfg.description, "mode": "serve", "algo": args.algo, "seed": args.seed,
           "queries": len(qs), "query_sizes": [q.k for q in qs],
           "time_limit_s_per_batch": args.time_limit, "max_batch": args.batch,
           "admission": "none for queries; batches admitted by free HBM", "loads": loads}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
