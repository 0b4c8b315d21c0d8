#!/usr/bin/env python3
"""Benchmark of the incremental match search on one MI355X.

Workload: scale-20 R-MAT (1,048,576 vertices, edge factor 16), Zipf(1.0)
vertex labels over 32, 2 uniform edge labels; 1,000,000 updates (80 %
insertions of held-out edges, 20 % deletions of present edges); 10 standing
8-vertex random-walk queries (5 sparse, 5 dense) admitted by the exact-count
rule of paracosm_amd.workloads.  One step = one batch of updates through the
whole device pipeline (graph update, seeds, search) for all queries; the
update records are resident in HBM before the timed region starts.  The
batch size is cut so that the warm-up and timed batches fit in the stream.

Prints one JSON line: batched updates/s (value) beside the plain per-update
GPU path (one update per launch sequence), the served path (fresh engines,
Poisson arrivals at several offered loads, p50/p99 per-update latency, counts
checked against the batched run), the search kernel's roofline figures, and
the CPU oracle's parallel GraphFlow restatement timed on a bounded sample on
this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "updates/s + p99 update latency, served vs per-update GPU path, exact counts"
HBM_PEAK_GBS = 8000.0


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "k_search_c2_hbm_traffic.md")
COUNTERS_PROFILE = os.path.join(HERE, "profiles", "k_search_c2_counters.md")


def profile_figures(path, config, algo, batch, avg_launch_us):
    """The machine-readable line of a committed rocprofv3 summary, when it
    describes this workload and its average launch duration is within 25 %
    of the live one (otherwise the kernel changed since it was profiled:
    None)."""
    try:
        with open(path) as f:
            line = [x for x in f if x.startswith("machine-readable:")][-1]
    except (OSError, IndexError):
        return None
    kv = dict(tok.split("=", 1) for tok in line.split()[1:])
    if (kv.get("config"), kv.get("algo"), int(kv.get("batch", 0))) != (config, algo, batch):
        return None
    prof_us = float(kv["avg_launch_us"])
    if not avg_launch_us or abs(prof_us - avg_launch_us) > 0.25 * avg_launch_us:
        return None
    return kv


def profiled_traffic(config, algo, batch, avg_launch_us):
    """HBM bytes per search launch from the committed FETCH_SIZE / WRITE_SIZE
    summary (None when it does not describe this kernel and workload)."""
    kv = profile_figures(TRAFFIC_PROFILE, config, algo, batch, avg_launch_us)
    return int(kv["traffic_bytes_per_launch"]) if kv else None


def profiled_l2(config, algo, batch, avg_launch_us):
    """L2 figures of the search from the committed counter summary: what
    bounds the kernel (its HBM traffic is a few % of the algorithmic bytes)."""
    kv = profile_figures(COUNTERS_PROFILE, config, algo, batch, avg_launch_us)
    if not kv:
        return None
    return {"l2_hit_rate": float(kv["l2_hit"]), "l2_busy": float(kv["l2_busy"]),
            "l1_to_l2_latency_cycles": float(kv["l1_l2_latency_cycles"]),
            "l2_read_requests_per_launch": int(kv["l2_read_req_per_launch"]),
            "l2_read_requests_per_candidate": float(kv["l2_read_req_per_candidate"]),
            "wave_wait_fraction": float(kv["wave_wait"]),
            "wave_issue_fraction": float(kv["wave_issue"]),
            "source": os.path.relpath(COUNTERS_PROFILE, HERE)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, p):
    return float(np.percentile(np.asarray(xs, np.float64), p)) if len(xs) else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    # warm-up batches also let the engine's autotuner score its candidate
    # search settings (a settling batch, then one batch searched once per
    # candidate)
    ap.add_argument("--warmup", type=int, default=2)
    # 40,000 x 25 batches = the 10^6-update stream: the same batch size for
    # every warm-up + steps count up to 25
    ap.add_argument("--batch", type=int, default=40000)
    ap.add_argument("--algo", default="parallel_graphflow")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--scale", type=int, default=None, help="override the R-MAT scale (testing)")
    ap.add_argument("--plain-sample", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="wall-clock cap per served offered load")
    ap.add_argument("--serve-loads", default="1e4,1e5,1e6,sat",
                    help="offered loads in updates/s; 'sat' = 2x the batched rate")
    ap.add_argument("--split-chunk", type=int, default=0)
    ap.add_argument("--spill-budget", type=int, default=0)
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) batched leg")
    ap.add_argument("--serve-budget", type=int, default=0,
                    help="served legs: spill budget of the served engines (0 = the one the "
                         "autotuner chose for the batched run)")
    ap.add_argument("--hbm-budget", type=int, default=0,
                    help="served legs: device bytes the engine plus in-flight batches may "
                         "hold (0 = 90%% of the device memory)")
    ap.add_argument("--also", default="parallel_symbi,parallel_turboflux",
                    help="other algorithms timed on the same batches (comma list, '' for none; "
                         "name:ref runs the plugin with the reference's matching orders)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 256 MiB device copy first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2

    # One HIP runtime per process: device memory comes from the engine's
    # runtime (the library links the system HIP; no second runtime is loaded).
    from paracosm_amd import workloads
    from paracosm_amd.engine import Matching, make_updates, lib

    lib()  # fail loudly if the engine is not built
    t0 = time.time()
    d = workloads.build(args.config, seed=args.seed, scale=args.scale)
    w = d.w
    log(f"workload {args.config}: {w.nv} vertices, {len(w.init_src)} initial edges, "
        f"{w.num_updates} updates ({int(w.upd_add.sum())} insertions) in {time.time() - t0:.1f}s")
    t0 = time.time()
    queries = workloads.admit(d, log=log)
    log(f"admitted {len(queries)} queries in {time.time() - t0:.1f}s: "
        f"k={[q.k for q in queries]} edges={[q.num_edges for q in queries]}")

    # every warm-up and timed batch must come from the stream: cut the batch
    # so that they fit, and refuse before any GPU work if even 1 does not
    n_batches = args.warmup + args.steps
    batch = min(args.batch, w.num_updates // max(n_batches, 1))
    if batch < 1 or args.steps < 1:
        print(json.dumps({"error": f"{n_batches} batches do not fit in the "
                                   f"{w.num_updates}-update stream"}))
        return 2
    if batch != args.batch:
        log(f"batch cut from {args.batch} to {batch} so that {n_batches} batches fit in the "
            f"{w.num_updates}-update stream")
    args.batch = batch

    # plain per-update GPU path on the stream's first updates (fresh engine)
    P = min(args.plain_sample, n_batches * args.batch, w.num_updates)
    mp = Matching(args.algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab)
    for q in queries[1:]:
        mp.register_query(q)
    mp.Preprocessing()
    plat = []
    pcounts = np.zeros(P, np.uint64)
    tp0 = time.perf_counter()
    for i in range(P):
        if i % 500 == 0:
            log(f"plain per-update: {i}/{P}")
        ta = time.perf_counter()
        if w.upd_add[i]:
            pcounts[i] = mp.AddEdge(int(w.upd_v1[i]), int(w.upd_v2[i]), int(w.upd_label[i]))
        else:
            pcounts[i] = mp.RemoveEdge(int(w.upd_v1[i]), int(w.upd_v2[i]))
        plat.append(time.perf_counter() - ta)
    plain_elapsed = time.perf_counter() - tp0
    mp.close()
    plain_ups = P / plain_elapsed
    log(f"plain per-update: {P} updates in {plain_elapsed:.3f}s = {plain_ups:.0f} updates/s")

    n_total = n_batches * args.batch
    ups = make_updates(w.upd_add[:n_total], w.upd_v1[:n_total], w.upd_v2[:n_total],
                       w.upd_label[:n_total])
    nq = len(queries)
    m = Matching(args.algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                 max_batch=args.batch, split_chunk=args.split_chunk,
                 spill_budget=args.spill_budget)
    for q in queries[1:]:
        m.register_query(q)
    m.Preprocessing()
    if args.copy_calibration:
        cal = m.device_buffer(512 << 20)
        cal.copy_within(256 << 20, 0, 256 << 20)
        log("copy calibration: one device copy read 268435456 bytes and wrote 268435456")
        cal.close()
    rec = 16
    # the whole timed input resident in HBM before timing starts
    d_ups = m.device_buffer(n_total * rec)
    d_ups.upload(ups.view(np.uint8))
    d_counts = m.device_buffer(n_total * nq * 8)

    def step(i):
        s = i * args.batch
        m.apply_batch_device(d_ups.ptr + s * rec, args.batch, d_counts.ptr + s * nq * 8)
        return args.batch

    for i in range(args.warmup):
        tw = time.perf_counter()
        step(i)
        log(f"warmup step {i}: {time.perf_counter() - tw:.3f}s {m.stats()}")
    m.sync()
    st0 = m.stats()
    lat = []
    n_done = 0
    m.sync()  # barrier: nothing of the warm-up is still running
    t_start = time.perf_counter()
    for i in range(args.warmup, args.warmup + args.steps):
        tb = time.perf_counter()
        n_done += step(i)
        lat.append(time.perf_counter() - tb)
        log(f"step {i}: {lat[-1]:.3f}s")
    m.sync()
    elapsed = time.perf_counter() - t_start
    st1 = m.stats()
    log(f"engine stats after the timed steps: {st1}")
    counts = d_counts.download(np.zeros(n_total * nq, np.uint64)).reshape(n_total, nq)
    timed = counts[args.warmup * args.batch: args.warmup * args.batch + n_done]
    matches = int(timed.sum())
    value = n_done / elapsed
    log(f"batched: {n_done} updates in {elapsed:.3f}s = {value:.0f} updates/s, "
        f"{matches} matches ({matches / elapsed:.3e}/s)")
    exact = bool(np.array_equal(pcounts, counts[:P].sum(axis=1)))
    log(f"plain per-update counts equal to batched: {exact}")

    # the search settings the autotuner chose, reused by the served engines
    tuned = (int(st1["split_chunk"]), int(st1["spill_budget"]))
    served = []
    if not args.no_serve:
        served = serve_legs(w, queries, args, value, tuned, counts, n_total)
    ingest = None if args.no_ingest else ingest_rate(m, w)
    m.close()

    # the other plugins on the same batches: rate, and per-update, per-query
    # counts equal to this run's
    others = {}
    for algo in [a for a in args.also.split(",") if a and a != args.algo]:
        v, c, cpu_ = batched_only(algo, w, queries, ups, args)
        same = bool(np.array_equal(c, counts))
        others[algo] = {"updates_per_s": round(v, 1), "counts_equal": same,
                        "candidates_per_update": round(cpu_, 1)}
        log(f"{algo}: {v:.0f} updates/s, {cpu_:.0f} candidates per update, "
            f"counts equal to {args.algo}: {same}")

    host = None if args.no_host else host_buffers_rate(w, queries, ups, args, counts)
    if host:
        log(f"host buffers (csm_apply_batch): {host}")

    # roofline of the search kernel: algorithmic bytes per launch over its
    # average launch duration (HIP events on the engine stream)
    launches = st1["search_launches"] - st0["search_launches"]
    k_ns = st1["search_ns"] - st0["search_ns"]
    alg = st1["alg_bytes"] - st0["alg_bytes"]
    per_launch_bytes = alg / max(launches, 1)
    avg_ns = k_ns / max(launches, 1)
    achieved = (per_launch_bytes / avg_ns) if avg_ns else 0.0  # bytes/ns == GB/s
    traffic = profiled_traffic(args.config, args.algo, args.batch, avg_ns / 1e3)
    l2 = profiled_l2(args.config, args.algo, args.batch, avg_ns / 1e3)
    cands = st1["candidates"] - st0["candidates"]
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "traffic_source": (os.path.relpath(TRAFFIC_PROFILE, HERE) if traffic else None),
            "kernel": "k_search", "avg_launch_us": round(avg_ns / 1e3, 2),
            "alg_bytes_per_launch": int(per_launch_bytes), "launches": int(launches),
            "candidates_per_launch": int(cands / max(launches, 1)),
            "candidates_per_update": round(cands / max(n_done, 1), 1),
            "traffic_over_alg_bytes": (round(traffic / per_launch_bytes, 4)
                                       if traffic and per_launch_bytes else None),
            "l2": l2,
            "kernel_share_of_step": round(k_ns / 1e9 / elapsed, 4) if elapsed else None}

    sat = next((x for x in served if x.get("saturating")), None)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, queries, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": d.cfg.description, "config": args.config, "seed": args.seed,
                   "algo": args.algo, "batch": args.batch, "queries": nq,
                   "query_fingerprint": workloads.query_fingerprint(queries),
                   "admission": workloads.ADMISSION_RULE,
                   "parallelism": "single-gpu"},
        "p50_update_latency_ms": round(1e3 * pct(lat, 50), 3),
        "p99_update_latency_ms": round(1e3 * pct(lat, 99), 3),
        "latency_note": "closed loop: each update's latency is its batch's latency",
        "served": served,
        # the served rate at a saturating offered load (above the batched
        # rate) over the plain per-update path
        "served_vs_plain": (round(sat["updates_per_s"] / plain_ups, 2) if sat else None),
        "served_saturated_p99_update_latency_ms": (sat["p99_update_latency_ms"] if sat else None),
        "plain_updates_per_s": round(plain_ups, 1),
        "plain_p50_update_latency_ms": round(1e3 * pct(plat, 50), 4),
        "plain_p99_update_latency_ms": round(1e3 * pct(plat, 99), 4),
        "speedup_vs_plain": round(value / plain_ups, 2),
        "matches_per_s": round(matches / elapsed, 1),
        "matches": matches,
        "exact_counts_plain_vs_batched": exact,
        "roofline": roof,
        "other_algorithms": others,
        "host_buffers": host,
        "ingest": ingest,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def serve_legs(w, queries, args, batched_rate, tuned, batched_counts, n_batched):
    """The served path at several offered loads.  Each load gets a fresh
    engine on the initial graph and serves the stream from its first update
    (requests of 1..4096 updates, Poisson arrivals, continuous batching) for
    at most --serve-seconds of arrivals; every served count is compared with
    the batched run's count of the same update and query."""
    from paracosm_amd.engine import Matching, make_updates
    from paracosm_amd.serve import Server, poisson_requests
    out = []
    for tok in [t for t in args.serve_loads.split(",") if t]:
        saturating = tok == "sat"
        load = 2.0 * batched_rate if saturating else float(tok)
        n_srv = int(min(args.serve_seconds * load, w.num_updates))
        if n_srv < 1:
            continue
        m = Matching(args.algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                     max_batch=args.batch, split_chunk=tuned[0],
                     spill_budget=args.serve_budget or tuned[1])
        for q in queries[1:]:
            m.register_query(q)
        m.Preprocessing()
        sl = slice(0, n_srv)
        reqs = poisson_requests(make_updates(w.upd_add[sl], w.upd_v1[sl], w.upd_v2[sl],
                                             w.upd_label[sl]), load, seed=args.seed)
        srv = Server(m, max_batch=args.batch, hbm_budget=args.hbm_budget)
        st = srv.serve(reqs)
        srv.close()
        m.close()
        got = np.concatenate([r.counts for r in reqs])
        k = min(n_srv, n_batched)
        exact = bool(np.array_equal(got[:k], batched_counts[:k]))
        out.append({"offered_updates_per_s": round(load, 1),
                    "updates_per_s": round(st.updates_per_s, 1),
                    "achieved_over_offered": round(st.updates_per_s / load, 3),
                    "saturating": saturating,
                    "p50_update_latency_ms": round(st.p50_ms, 3),
                    "p99_update_latency_ms": round(st.p99_ms, 3),
                    "requests": st.requests, "updates": st.updates,
                    "batches": st.batches, "mean_batch": round(st.mean_batch, 1),
                    "matches_per_s": round(st.matches / st.elapsed, 1),
                    "hbm_budget": srv.hbm_budget,
                    "refused_for_memory": st.refused_for_memory,
                    "counts_equal_batched": exact, "compared_updates": k})
        log(f"served at {load:.0f}/s offered: {out[-1]}")
    return out


def host_buffers_rate(w, queries, ups, args, batched_counts):
    """The batched path through csm_apply_batch with host arrays (records
    copied in, counts copied out on every call): the PCIe-inclusive rate on
    the same warm-up and timed batches, on a fresh engine with the batched
    run's search settings."""
    from paracosm_amd.engine import Matching
    nq = len(queries)
    m = Matching(args.algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                 max_batch=args.batch, split_chunk=args.split_chunk,
                 spill_budget=args.spill_budget)
    for q in queries[1:]:
        m.register_query(q)
    m.Preprocessing()
    counts = np.zeros((len(ups), nq), np.uint64)
    for i in range(args.warmup):
        s = i * args.batch
        pos, neg, _, _ = m.apply_batch(ups[s:s + args.batch])
        counts[s:s + args.batch] = pos + neg
    t = time.perf_counter()
    done = 0
    for i in range(args.warmup, args.warmup + args.steps):
        s = i * args.batch
        pos, neg, _, _ = m.apply_batch(ups[s:s + args.batch])
        counts[s:s + args.batch] = pos + neg
        done += args.batch
    elapsed = time.perf_counter() - t
    m.close()
    return {"updates_per_s": round(done / elapsed, 1),
            "counts_equal": bool(np.array_equal(counts, batched_counts)),
            "note": "csm_apply_batch: host update records in, host counts out, every batch"}


def batched_only(spec, w, queries, ups, args):
    """Batched rate of another plugin on the same warm-up and timed batches
    (spec "name" or "name:ref", the latter with the reference's matching
    orders: for parallel_symbi its search-time order rule); returns
    (updates/s, counts of every batch, candidate entries tested per timed
    update)."""
    from paracosm_amd.engine import Matching
    algo, _, mode = spec.partition(":")
    nq = len(queries)
    n_total = len(ups)
    m = Matching(algo, queries[0], w.vlabels, w.init_src, w.init_dst, w.init_elab,
                 max_batch=args.batch, split_chunk=args.split_chunk,
                 spill_budget=args.spill_budget, reference_orders=(mode == "ref"))
    for q in queries[1:]:
        m.register_query(q)
    m.Preprocessing()
    d_ups = m.device_buffer(n_total * 16)
    d_ups.upload(ups.view(np.uint8))
    d_counts = m.device_buffer(n_total * nq * 8)

    def step(i):
        s = i * args.batch
        m.apply_batch_device(d_ups.ptr + s * 16, args.batch, d_counts.ptr + s * nq * 8)
        return args.batch

    for i in range(args.warmup):
        step(i)
    m.sync()
    c0 = m.stats()["candidates"]
    t = time.perf_counter()
    done = 0
    for i in range(args.warmup, args.warmup + args.steps):
        done += step(i)
        if i % 5 == 0:
            log(f"{spec}: step {i}")
    m.sync()
    elapsed = time.perf_counter() - t
    cands = m.stats()["candidates"] - c0
    counts = d_counts.download(np.zeros(n_total * nq, np.uint64)).reshape(n_total, nq)
    m.close()
    return done / elapsed, counts, cands / max(done, 1)


def ingest_rate(m, w, reps=5):
    """Device parsing of the workload's whole update stream as text
    (`e a b l` / `-e a b l` lines, the Graph::LoadUpdateStream grammar),
    resident in HBM: text bytes and records per second of csm_parse_text_device
    (scratch allocation included), best of `reps`."""
    from paracosm_amd.engine import PARSE_STREAM
    tok = np.where(w.upd_add == 1, "e", "-e")
    lines = np.char.add(np.char.add(np.char.add(np.char.add(tok, " "), w.upd_v1.astype(str)),
                                    np.char.add(" ", w.upd_v2.astype(str))),
                        np.char.add(" ", w.upd_label.astype(str)))
    text = ("\n".join(lines.tolist()) + "\n").encode()
    n = len(lines)
    d_text = m.device_buffer(len(text))
    d_text.upload(np.frombuffer(text, np.uint8))
    d_rec = m.device_buffer(16 * n)
    best = None
    for _ in range(reps):
        m.sync()
        t = time.perf_counter()
        got = m.parse_text_device(d_text.ptr, len(text), PARSE_STREAM, d_rec.ptr, n)
        m.sync()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    assert got == n
    d_text.close()
    d_rec.close()
    out = {"text_bytes": len(text), "records": n, "seconds": round(best, 5),
           "text_GB_per_s": round(len(text) / best / 1e9, 2),
           "records_per_s": round(n / best, 1),
           "note": "csm_parse_text_device on the whole update stream in HBM; host iostream "
                   "parsing (Graph::LoadUpdateStream) is what it replaces"}
    log(f"ingest: {out}")
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, queries, seconds):
    """The oracle's restatement of the reference's parallel GraphFlow
    (serial depth-2 expansion, depth-3 task vector, OpenMP loop) on the first
    S updates of the same stream for every query, S grown until the sample
    takes about `seconds`.  Threads: the host CPUs this process may use,
    capped by OMP_NUM_THREADS where the host sets it (the GPU host's CPU
    share for one GPU)."""
    from oracle import oracle as O
    avail = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    cores = min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail
    t0 = time.time()
    dg = O.OGraph(w.vlabels, w.init_src, w.init_dst, w.init_elab)
    log(f"cpu baseline: oracle graph built in {time.time() - t0:.1f}s")
    qgs = [O.OGraph.from_query(q) for q in queries]
    args = (w.upd_add, w.upd_v1, w.upd_v2, w.upd_label)
    s = 100
    while True:
        tot = 0.0
        for qg in qgs:
            r = O.run("parallel_graphflow", dg, qg, *args, threads=cores, limit=s)
            tot += r.seconds
        log(f"cpu baseline: {s} updates x {len(qgs)} queries took {tot:.2f}s")
        if tot * 2 >= seconds or s >= w.num_updates:
            break
        # at most 2x per step: the cost per update grows along the stream
        s = int(min(w.num_updates, s * min(2.0, max(1.25, 0.8 * seconds / max(tot, 1e-3)))))
    model = cpu_model()
    log(f"cpu baseline: {s} updates x {len(qgs)} queries in {tot:.2f}s on {cores} threads "
        f"({model}, {avail} CPUs available)")
    return {"value": round(s / tot, 1), "unit": "updates/s", "cores": cores, "kind": "port",
            "cpu_model": model, "cpus_available": avail,
            "omp_num_threads_env": env or None,
            "sample": f"first {s} updates of the stream, all {len(qgs)} queries, "
                      f"oracle parallel_graphflow, {cores} OpenMP threads"}


if __name__ == "__main__":
    sys.exit(main())
