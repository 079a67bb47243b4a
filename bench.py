#!/usr/bin/env python3
"""Headline benchmark: out-of-GPU-memory SSSP on one MI355X.

Workload (the metric's single-GPU configuration): directed R-MAT scale 24,
edge factor 16, a/b/c/d = .57/.19/.19/.05 (graph seed 1002), self-loops and
duplicates removed, u32 weights uniform in [1, 63] (seed 2002); device edge
budget = 25% of the CSR edge+weight bytes; Ascetic plan with advisek = 0.8;
source = seeded uniform pick (seed 3002) among vertices with out-degree > 0.
Engine settings (not part of the workload): zero-copy dynamic region,
delta-stepping with light edges of weight < 8 and buckets of width 256.

A step is one full SSSP from that source (graph resident in pinned host
memory + the static prefix in HBM, as prepared once before timing).  The
timed region covers K steps; value = GTEPS = sum of out-degrees of reached
vertices / seconds per step (Graph500 convention).

Prints one JSON line.  Extra fields: roofline of the dominant kernel (HIP
events on its own stream), host-link fraction of the copy stream, the CPU
baseline (oracle Dijkstra on the host, single thread), and --compare runs of
the Subway and USM approaches on the same graph.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "GTEPS (BFS/SSSP/CC/PR) + served BFS queries/s at p99 vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0
LINK_PEAK_GBS = 63.0   # PCIe Gen5 x16, one direction


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pick_source(ptr, E, seed):
    deg = np.diff(np.append(ptr.astype(np.int64), E))
    cand = np.flatnonzero(deg > 0)
    return int(cand[np.random.default_rng(seed).integers(len(cand))]), deg


def kernel_table(st):
    """Per-class algorithmic GB/s from the engine's HIP-event timings.  A
    zero-copy round that relaxes both regions in one launch (k_relax_round)
    is timed as the dynamic class; its row then carries the static bytes too."""
    rows = {}
    fused = st["relax_static_kernel_ms"] == 0 and st["relax_static_bytes"] > 0 and st["relax_dynamic_kernel_ms"] > 0
    for cls, ms_key, n_key, b_key in [
        ("relax_static", "relax_static_kernel_ms", "relax_static_launches", "relax_static_bytes"),
        ("relax_dynamic", "relax_dynamic_kernel_ms", "relax_dynamic_launches", "relax_dynamic_bytes"),
        ("frontier", "frontier_kernel_ms", "frontier_launches", "frontier_bytes"),
    ]:
        ms, n, b = st[ms_key], st[n_key], st[b_key]
        if fused and cls == "relax_dynamic":
            b += st["relax_static_bytes"]
        if n and ms > 0:
            rows[cls] = {"ms": ms, "launches": n, "avg_us": 1e3 * ms / n, "bytes": b,
                         "bytes_per_launch": b / n, "gbs": b / (ms * 1e-3) / 1e9}
            if fused and cls == "relax_dynamic":
                rows[cls]["fused"] = True
    return rows


def cpu_info():
    """CPU model name plus the identifying fields of the first processor
    entry (a virtualised host may report a generic model name only)."""
    want = {"model name": "model", "cpu family": "family", "model": "model_id", "stepping": "stepping",
            "cpu MHz": "mhz", "cache size": "cache"}
    out = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # end of the first processor entry
                k, _, v = line.partition(":")
                k = k.strip()
                if k in want:
                    out[want[k]] = v.strip()
    except OSError:
        pass
    model = out.pop("model", "")
    if out:
        model += " (" + ", ".join(f"{k} {v}" for k, v in out.items()) + ")"
    return model


def host_info():
    """The host the CPU baseline ran on: CPU model, logical CPUs this process
    may use, installed memory and NUMA nodes."""
    info = {"cpu": cpu_info(), "nproc": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else os.cpu_count()}
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    info["memory_gib"] = round(int(line.split()[1]) / (1 << 20), 1)
                    break
    except OSError:
        pass
    try:
        info["numa_nodes"] = len([n for n in os.listdir("/sys/devices/system/node") if n.startswith("node")])
    except OSError:
        pass
    return info


def cpu_record(value, unit, cores, sample, **extra):
    """cpu_baseline: the oracle (a restatement of the reference CPU path, -O3)
    timed over its compute loop only, like the reference's CPUCalTime."""
    rec = {"value": value, "unit": unit, "cores": cores, "kind": "port", "sample": sample,
           "timed": "compute loop only (allocation and initialisation excluded)"}
    rec.update(host_info())
    rec.update(extra)
    return rec


def base_record(metric_value, unit, steps, warmup, ms, dtype, config, higher=True):
    return {"metric": METRIC, "value": metric_value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": round(ms, 3), "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic", "config": config}


KERNEL_NAMES = {"relax_static": "k_relax_group / k_relax_hub (static region)",
                "relax_dynamic": "k_relax_dyn (dynamic region)", "frontier": "k_frontier_count/scan/write",
                "relax_round": "k_relax_round (static bins + zero-copy dynamic region, one launch)"}


def roofline_of(stats_list, zero_copy=False, traffic=None):
    """Roofline of the dominant kernel class (most HIP-event time).

    In HBM (static relax, frontier, a staged dynamic relax reading its slot)
    the bound is HBM: achieved = algorithmic bytes / kernel time against
    8 TB/s.  A zero-copy dynamic relax reads its edges from pinned host memory
    over the link, so its bound is the host link: achieved = edge bytes read
    over the link / kernel time against 63 GB/s; its HBM figures are kept in
    "hbm".  `traffic`: memory-system bytes per launch of that class measured
    by rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy; host
    reads included), when a committed profile of the same workload has them."""
    agg = {}
    for key in ["relax_static_kernel_ms", "relax_static_launches", "relax_static_bytes", "relax_dynamic_kernel_ms",
                "relax_dynamic_launches", "relax_dynamic_bytes", "frontier_kernel_ms", "frontier_launches",
                "frontier_bytes", "copy_ms", "h2d_bytes"]:
        agg[key] = sum(s[key] for s in stats_list)
    kt = kernel_table(agg)
    roof = None
    if kt:
        dom = max(kt, key=lambda c: kt[c]["ms"])
        r = kt[dom]
        hbm = {"achieved": round(r["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(r["gbs"] / HBM_PEAK_GBS, 4), "bytes_per_launch": round(r["bytes_per_launch"], 1),
               "traffic": traffic}
        if dom == "relax_dynamic" and zero_copy and agg["h2d_bytes"]:
            gbs = agg["h2d_bytes"] / (r["ms"] * 1e-3) / 1e9
            name = KERNEL_NAMES["relax_round"] if r.get("fused") else KERNEL_NAMES[dom]
            roof = {"bound": "host-link", "kernel": name, "achieved": round(gbs, 2),
                    "peak": LINK_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / LINK_PEAK_GBS, 4),
                    "traffic": traffic, "link_bytes_per_launch": round(agg["h2d_bytes"] / r["launches"], 1),
                    "avg_launch_us": round(r["avg_us"], 3), "hbm": hbm}
        else:
            roof = {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": hbm["achieved"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": hbm["frac"], "traffic": traffic, "avg_launch_us": round(r["avg_us"], 3),
                    "bytes_per_launch": hbm["bytes_per_launch"]}
    link = None
    if agg["copy_ms"] > 0:
        gbs = agg["h2d_bytes"] / (agg["copy_ms"] * 1e-3) / 1e9
        link = {"achieved": round(gbs, 2), "peak": LINK_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / LINK_PEAK_GBS, 4),
                "measured_on": "copy stream (hipMemcpyAsync)"}
    elif zero_copy and agg["h2d_bytes"] and agg["relax_dynamic_kernel_ms"] > 0:
        gbs = agg["h2d_bytes"] / (agg["relax_dynamic_kernel_ms"] * 1e-3) / 1e9
        link = {"achieved": round(gbs, 2), "peak": LINK_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / LINK_PEAK_GBS, 4),
                "measured_on": ("k_relax_round" if kt.get("relax_dynamic", {}).get("fused") else "k_relax_dyn")
                + " (direct reads)"}
    return roof, link, kt


def plan_flags(args, approach, loader=None):
    """og_plan_params.flags for a workload: the dynamic-region loader and
    the link format (USM reads the CSR in place and takes neither)."""
    import onegraph_amd as og

    if approach == "usm":
        return 0
    loader = loader or args.loader
    return ((og.PLAN_ZERO_COPY if loader == "zerocopy" else 0) | (og.PLAN_NO_PACK if args.no_pack else 0) |
            (og.PLAN_NO_OVERLAP if args.no_overlap else 0) | (og.PLAN_NO_FUSE if args.no_fuse else 0))


def timed(fn, steps, warmup, label):
    for i in range(warmup):
        r = fn()
        log(f"{label}: warmup {i} done")
    stats = []
    t = time.perf_counter()
    for _ in range(steps):
        stats.append(fn())
    wall = time.perf_counter() - t
    log(f"{label}: {steps} steps in {wall:.3f}s")
    return stats, wall


def wl_bfs_hbm(args):
    """Plain BFS, whole graph in HBM (advisek 1.0), R-MAT scale 20 ef 16 seed
    1001, source = lowest vertex with out-degree > 0; bit-exact vs cpu BFS."""
    import onegraph_amd as og
    import oracle

    g = og.Graph.rmat(args.scale, args.edge_factor, seed=1001)
    ptr, cols, _, _ = g.export()
    V, E = len(ptr), len(cols)
    deg = np.diff(np.append(ptr.astype(np.int64), E))
    src = int(np.flatnonzero(deg > 0)[0])
    flags = (og.PLAN_DIRECTION_OPT if args.diropt else 0) | (og.PLAN_NO_FUSE if args.no_fuse else 0)
    plan = g.prepare("ascetic", edge_budget_bytes=E * 4, advisek=1.0, flags=flags)
    stats, wall = timed(lambda: g.bfs(src, want_values=False)[1], args.steps, args.warmup, "bfs-hbm")
    lv, _ = g.bfs(src)
    ms = 1e3 * wall / args.steps
    trav = stats[-1]["traversed_edges"]
    ref, _, _ = oracle.bfs(ptr, cols, src)
    cpu_s = oracle.loop_seconds()
    roof, link, kt = roofline_of(stats)
    out = base_record(round(trav / (ms * 1e-3) / 1e9, 4), "GTEPS", args.steps, args.warmup, ms, "u32",
                      {"workload": f"bfs whole graph in HBM rmat{args.scale} ef{args.edge_factor}", "V": V, "E": E,
                       "source": src, "advisek": 1.0, "approach": "ascetic", "static_edges": plan["static_edges"],
                       "direction": "top-down/bottom-up" if args.diropt else "top-down"})
    out.update({"roofline": roof, "kernels": kt, "iterations": stats[-1]["iterations"],
                "parity": {"bit_exact": bool(np.array_equal(lv, ref)), "reached": int((ref != og.UNREACHED).sum())},
                "cpu_baseline": cpu_record(round(trav / cpu_s / 1e9, 5), "GTEPS", 1,
                                           f"oracle level-synchronous BFS (runCPUBFS restated), one full run, "
                                           f"{cpu_s:.3f} s")})
    print(json.dumps(out), flush=True)


def er_graph(args):
    import onegraph_amd as og

    t0 = time.time()
    g = og.Graph.er(1 << args.scale, 32, seed=1003)
    log(f"ER n=2^{args.scale} E={g.num_edges} generated in {time.time() - t0:.1f}s")
    return g


def wl_cc_ooc(args):
    """CC out-of-memory: ER n=2^25 avg degree 32 (seed 1003), Ascetic
    advisek 0.8, edge budget 25% of the CSR edge bytes; bit-exact vs serial."""
    import onegraph_amd as og
    import oracle

    g = er_graph(args)
    ptr, cols, _, _ = g.export()
    V, E = len(ptr), len(cols)
    plan = g.prepare("ascetic", edge_budget_bytes=E, advisek=0.8, flags=plan_flags(args, "ascetic"))
    if args.dyn_grid is not None:
        g.set_tuning(dyn_grid=args.dyn_grid)
    stats, wall = timed(lambda: g.cc(want_values=False)[1], args.steps, args.warmup, "cc-ooc")
    lab, _ = g.cc()
    ms = 1e3 * wall / args.steps
    cmp = {}
    if args.compare:
        for ap_name in ["subway", "usm"]:
            g.prepare(ap_name, edge_budget_bytes=E if ap_name == "subway" else 0,
                      flags=plan_flags(args, ap_name, "copy"))
            s2, w2 = timed(lambda: g.cc(want_values=False)[1], 1, 0, f"cc-{ap_name}")
            cmp[ap_name] = {"ms": round(1e3 * w2, 2), "gteps": round(E / w2 / 1e9, 4), "h2d_bytes": s2[0]["h2d_bytes"]}
    ref, _ = oracle.cc(ptr, cols)
    cpu_s = oracle.loop_seconds()
    roof, link, kt = roofline_of(stats, zero_copy=args.loader == "zerocopy")
    out = base_record(round(E / (ms * 1e-3) / 1e9, 4), "GTEPS", args.steps, args.warmup, ms, "u32",
                      {"workload": f"cc out-of-memory er n=2^{args.scale} avg 32", "V": V, "E": E,
                       "edge_budget_bytes": E, "advisek": 0.8, "approach": "ascetic", "loader": args.loader,
                       "static_edges": plan["static_edges"], "staging_edges": plan["staging_edges"]})
    out.update({"roofline": roof, "host_link": link, "kernels": kt, "iterations": stats[-1]["iterations"],
                "h2d_bytes_per_step": stats[-1]["h2d_bytes"], "compare": cmp,
                "parity": {"bit_exact": bool(np.array_equal(lab, ref)),
                           "components": int(len(np.unique(ref)))},
                "cpu_baseline": cpu_record(round(E / cpu_s / 1e9, 5), "GTEPS", 1,
                                           f"oracle serial CC (CC/serial restated), one full run, {cpu_s:.2f} s")})
    print(json.dumps(out), flush=True)


def wl_pr_ooc(args):
    """PageRank out-of-memory on the same ER graph (pull layout), Ascetic
    advisek 0.8, budget 25%; FP32-Kahan and FP64, errors vs the FP64 oracle."""
    import onegraph_amd as og
    import oracle

    g0 = er_graph(args)
    g = g0.to_pull()
    g0.close()
    ptr, cols, _, od = g.export()
    V, E = len(ptr), len(cols)
    plan = g.prepare("ascetic", edge_budget_bytes=E, advisek=0.8)
    res = {}
    threads = min(16, os.cpu_count() or 1)
    # the reference's CPU PageRank is single-threaded: that is the baseline;
    # the all-core run (same results bit for bit) is reported beside it
    ref, ref_it, ref_l1 = oracle.pagerank(od, ptr, cols, parallel=False)
    cpu_s = oracle.loop_seconds()
    ref_p, it_p, _ = oracle.pagerank(od, ptr, cols, threads=threads, parallel=True)
    cpu_s_par = oracle.loop_seconds()
    assert it_p == ref_it and np.array_equal(ref_p, ref)
    rec = None
    for fp64 in (False, True):
        name = "fp64" if fp64 else "fp32"
        stats, wall = timed(lambda: g.pagerank(fp64=fp64, want_values=False)[1], args.steps, args.warmup, f"pr-{name}")
        r, st = g.pagerank(fp64=fp64)
        ms = 1e3 * wall / args.steps
        it = stats[-1]["iterations"]
        err = np.abs(r.astype(np.float64) - ref)
        roof, link, kt = roofline_of([s["run"] for s in stats])
        res[name] = {"ms": round(ms, 3), "iterations": it, "gteps": round(E * it / (ms * 1e-3) / 1e9, 4),
                     "max_abs": float(err.max()), "max_rel": float((err / ref).max()), "l1": float(err.sum()),
                     "final_l1": stats[-1]["final_l1"], "roofline": roof, "host_link": link, "kernels": kt,
                     "h2d_bytes_per_step": stats[-1]["run"]["h2d_bytes"]}
        if not fp64:
            rec = base_record(res[name]["gteps"], "GTEPS", args.steps, args.warmup, ms, "f32",
                              {"workload": f"pagerank out-of-memory er n=2^{args.scale} avg 32", "V": V, "E": E,
                               "edge_budget_bytes": E, "advisek": 0.8, "approach": "ascetic",
                               "static_edges": plan["static_edges"]})
            rec.update({"roofline": roof, "host_link": link})
    rec.update({"pagerank": res,
                "cpu_baseline": cpu_record(round(E * ref_it / cpu_s / 1e9, 5), "GTEPS", 1,
                                           f"oracle FP64 PageRank, one thread, {ref_it} iterations, {cpu_s:.2f} s",
                                           all_cores={"value": round(E * ref_it / cpu_s_par / 1e9, 5),
                                                      "cores": threads, "seconds": round(cpu_s_par, 3),
                                                      "label": "all-core OpenMP variant (identical ranks)"}),
                "parity": {"tolerance_max_abs": 1e-6, "fp32_ok": res["fp32"]["max_abs"] < 1e-6,
                           "fp64_ok": res["fp64"]["max_abs"] < 1e-6}})
    print(json.dumps(rec), flush=True)


def wl_msbfs(args):
    """Batched multi-source BFS: R-MAT scale 26 ef 16 (seed 1004), edge budget
    50%, Ascetic advisek 0.8, 1024 sources (seed 3004) in batches of 64,
    against single-source launches (a timed sample, extrapolated)."""
    import onegraph_amd as og
    import oracle

    t0 = time.time()
    g = og.Graph.rmat(args.scale, args.edge_factor, seed=1004)
    V, E = g.num_vertices, g.num_edges
    log(f"rmat scale {args.scale} E={E} generated in {time.time() - t0:.1f}s")
    ptr, cols, _, _ = g.export()
    deg = np.diff(np.append(ptr.astype(np.int64), E))
    srcs = np.random.default_rng(3004).choice(np.flatnonzero(deg > 0), args.sources, replace=False).astype(np.uint32)
    plan = g.prepare("ascetic", edge_budget_bytes=E * 4 // 2, advisek=0.8, flags=plan_flags(args, "ascetic"))
    if args.dyn_grid is not None:
        g.set_tuning(dyn_grid=args.dyn_grid)
    log(f"plan {plan}")
    tuned = None
    if args.autotune:  # bins, lanes per vertex, hub tile and dynamic grid for 64-source batches
        best, best_ms = g.autotune(og.TUNE_MSBFS, 3004, runs=1)
        tuned = {"tuning": best, "batch_ms": round(best_ms, 3)}
        log(f"msbfs autotuned {best} ({best_ms:.1f} ms per 64-source batch)")
    if args.warmup:
        g.msbfs(srcs[:64], want_levels=False)
    t = time.perf_counter()
    _, cs, rc, st = g.msbfs(srcs, want_levels=False)
    ms_wall = time.perf_counter() - t
    log(f"msbfs {len(srcs)} sources in {ms_wall:.2f}s")
    sample = srcs[: args.plain_sample]
    plain = []
    t = time.perf_counter()
    for s in sample:
        _, s1 = g.bfs(int(s), want_values=False)
        plain.append(s1)
    plain_wall = time.perf_counter() - t
    per_plain = plain_wall / len(sample)
    digests_equal = all(p["checksum"] == int(cs[i]) for i, p in enumerate(plain))
    # full oracle check of two sources
    ok = True
    for i in range(2):
        ref, _, _ = oracle.bfs(ptr, cols, int(srcs[i]))
        ok &= og.levels_checksum(ref) == int(cs[i])
    trav = st["traversed_edges"]
    qps = len(srcs) / ms_wall
    # CPU baseline: the oracle BFS (runCPUBFS restated, one thread) on a
    # bounded sample of the same sources, compute loop only
    cpu_n = min(args.cpu_sample, len(srcs))
    cpu_s = 0.0
    for s in srcs[:cpu_n]:
        oracle.bfs(ptr, cols, int(s))
        cpu_s += oracle.loop_seconds()
    out = base_record(round(qps, 3), "queries/s", 1, args.warmup, 1e3 * ms_wall, "u64-mask",
                      {"workload": f"multi-source bfs rmat{args.scale} ef{args.edge_factor} {len(srcs)} sources",
                       "V": V, "E": E, "edge_budget_bytes": E * 4 // 2, "advisek": 0.8, "approach": "ascetic",
                       "loader": args.loader, "batch": 64, "plain_sample": len(sample),
                       "autotuned": bool(args.autotune)})
    out["aggregate_gteps"] = {"value": round(trav / ms_wall / 1e9, 4),
                              "note": "sum over the sources of each traversal's edges / wall time; traversals of one "
                                      "batch share edge reads, so this is not comparable to single-source GTEPS"}
    if tuned:
        out["autotune"] = tuned
    out["cpu_baseline"] = cpu_record(round(cpu_n / cpu_s, 4) if cpu_s else None, "queries/s", 1,
                                     f"oracle BFS from the first {cpu_n} of the same sources, {cpu_s:.2f} s")
    roof, link, kt = roofline_of([st], zero_copy=args.loader == "zerocopy")
    out.update({"queries_per_s": round(qps, 2), "plain_queries_per_s": round(1 / per_plain, 3),
                "speedup_vs_plain": round(qps * per_plain, 2),
                "plain_sample": len(sample), "plain_gteps": round(np.mean([p["traversed_edges"] for p in plain]) /
                                                                   per_plain / 1e9, 4),
                "roofline": roof, "host_link": link, "kernels": kt, "iterations": st["iterations"],
                "h2d_bytes": st["h2d_bytes"],
                "parity": {"digests_equal_plain_sample": digests_equal, "oracle_two_sources": bool(ok)}})
    print(json.dumps(out), flush=True)


def _pct(xs, p):
    return float(np.percentile(np.asarray(xs), p)) if len(xs) else None


def wl_serve(args):
    """Served BFS/SSSP queries on three device-resident R-MAT graphs (scales
    22, 24, 26; ef 16; weights [1,63]; seeds 1005+i / 2005+i).  Poisson
    arrivals at each rate (seed 3005), graph uniform, source uniform over
    vertices with out-degree > 0; each mix in --serve-mix gives the share of
    BFS queries (0.7: the 70% BFS / 30% SSSP workload; 1.0: BFS only).  The
    server batches continuously; the plain path serves one query at a time
    with og_bfs / og_sssp, its latency obtained by replaying the same arrival
    trace through a FIFO queue with service times measured on a sample of its
    queries (graphs prepared with their in-edges, so plain BFS is
    direction-optimising like the server's; the top-down plain rate is
    reported beside it).  value = served queries/s of the first mix at the
    highest rate whose p99 latency is within --p99-target-ms (else at the
    lowest rate)."""
    import threading

    import onegraph_amd as og

    scales = [int(x) for x in (args.serve_scales.split(","))]
    graphs, cands = [], []
    for i, sc in enumerate(scales):
        t0 = time.time()
        g = og.Graph.rmat(sc, args.edge_factor, seed=1005 + i, weight_seed=2005 + i)
        ptr, _, _, _ = g.export()
        deg = np.diff(np.append(ptr.astype(np.int64), g.num_edges))
        cands.append(np.flatnonzero(deg > 0))
        graphs.append(g)
        log(f"graph {i}: rmat scale {sc} V={g.num_vertices} E={g.num_edges} ready in {time.time() - t0:.1f}s")
    rates = [float(r) for r in args.rates.split(",")]
    mixes = [float(x) for x in args.serve_mix.split(",")]

    def trace(rate, seed, bfs_share):
        rng = np.random.default_rng(seed)
        n = max(1, int(rate * args.duration))
        t = np.cumsum(rng.exponential(1.0 / rate, n))
        gi = rng.integers(0, len(graphs), n)
        kind = (rng.random(n) >= bfs_share).astype(int)  # BFS (0) with probability bfs_share, else SSSP (1)
        src = np.array([int(cands[g][rng.integers(len(cands[g]))]) for g in gi])
        return t, gi, kind, src

    # plain one-at-a-time path: service times on a sample, per (graph, kind);
    # first top-down BFS (the reference's order), then with the graphs'
    # in-edges on the device (direction-optimising BFS), which the server
    # also uses: the served/plain ratio is taken against the latter
    def sample_plain(tag, flags):
        svc = {}
        rng = np.random.default_rng(3005)
        for gi, g in enumerate(graphs):
            g.prepare("ascetic", edge_budget_bytes=g.num_edges * 8, advisek=1.0, flags=flags,
                      split_weight=args.split_weight)
            g.set_tuning(delta=args.delta)  # SSSP near-far width for the plain path and the lanes alike
            if args.autotune:  # shapes for SSSP, shared by the plain path and the server's lanes
                best, best_ms = g.autotune(og.TUNE_SSSP, int(cands[gi][0]), runs=1)
                log(f"{tag} g{gi}: autotuned {best} ({best_ms:.2f} ms)")
            for kind in (0, 1):
                ts = []
                for _ in range(args.plain_sample // 2 if kind else args.plain_sample):
                    s = int(cands[gi][rng.integers(len(cands[gi]))])
                    t = time.perf_counter()
                    _, st = g.bfs(s, want_values=False) if kind == 0 else g.sssp(s, want_values=False)
                    ts.append(time.perf_counter() - t)
                svc[(gi, kind)] = np.array(ts)
                log(f"plain {tag} g{gi} {'sssp' if kind else 'bfs'}: mean {1e3 * np.mean(ts):.2f} ms")
        return svc

    service_td = sample_plain("top-down", 0)
    service = sample_plain("diropt", og.PLAN_DIRECTION_OPT)

    # per-query state (og_server_create's admission rule): per graph the BFS
    # engine (40 V bytes), per weighted graph with batching the 16-slot SSSP
    # engine (92 V + 8 + the relax's chunk table), then as many single-source
    # SSSP lanes (up to --sssp-lanes per weighted graph) as the rest holds;
    # by default the budget holds the engines and all the lanes
    batch = args.sssp_batch or args.no_sssp_lanes
    engines = sum(40 * g.num_vertices + (92 * g.num_vertices + 8 + 4 * (16 * g.num_edges // 1024 + 2)
                                         if g.info()["kind"] == og.KIND_WEIGHTED and batch else 0) for g in graphs)
    lanes = sum(args.sssp_lanes * 19 * g.num_vertices for g in graphs if g.info()["kind"] == og.KIND_WEIGHTED)
    budget = int(args.serve_budget_gib * (1 << 30)) if args.serve_budget_gib else engines + lanes
    srv_flags = (og.SERVER_NO_SSSP_LANES if args.no_sssp_lanes else 0) | (og.SERVER_SSSP_BATCH if args.sssp_batch
                                                                         else 0)

    def replay(t_arr, gis, kinds, svc, seed):
        """FIFO one-at-a-time service of the trace with sampled service times."""
        srng = np.random.default_rng(seed)
        free_at, plat = 0.0, []
        for k in range(len(t_arr)):
            sv = svc[(int(gis[k]), int(kinds[k]))]
            start = max(free_at, t_arr[k])
            free_at = start + float(sv[srng.integers(len(sv))])
            plat.append(1e3 * (free_at - t_arr[k]))
        return len(t_arr) / max(free_at, t_arr[-1]), plat

    def run_rate(rate, seed, bfs_share):
        t_arr, gis, kinds, srcs = trace(rate, seed, bfs_share)
        plain_qps, plat = replay(t_arr, gis, kinds, service, seed + 4)
        td_qps, _ = replay(t_arr, gis, kinds, service_td, seed + 4)
        srv = og.Server(graphs, state_budget_bytes=budget, sssp_delta=args.serve_sssp_delta, flags=srv_flags,
                        sssp_lanes=args.sssp_lanes, sssp_running_max=args.sssp_running_max)
        info0 = srv.info()
        lat, done, submit_at = [], {}, {}
        t0 = time.perf_counter()

        def submitter():
            for k in range(len(t_arr)):
                dt = t0 + t_arr[k] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                qid = srv.submit(int(gis[k]), int(kinds[k]), int(srcs[k]))
                submit_at[qid] = (k, time.perf_counter() - t0 - t_arr[k])

        th = threading.Thread(target=submitter)
        th.start()
        last = time.perf_counter()
        while len(done) < len(t_arr):
            for r in srv.poll(4096, timeout_ms=100):
                done[r["id"]] = r
                last = time.perf_counter()
            if time.perf_counter() - last > 120:
                raise RuntimeError("server stalled")
        th.join()
        wall = last - t0
        info = srv.info()
        srv.close()
        bad = 0
        lat_kind = {0: [], 1: []}
        lat_graph = {}
        for qid, r in done.items():
            k, lag = submit_at[qid]
            lt = r["latency_ms"] + 1e3 * lag  # from the scheduled arrival, not from the submit call
            lat.append(lt)
            lat_kind[r["kind"]].append(lt)
            lat_graph.setdefault(f"{'sssp' if r['kind'] else 'bfs'}_scale{scales[r['graph']]}", []).append(
                (lt, r["queue_ms"] + 1e3 * lag, r["rounds"]))
            if r["status"] != 0:
                bad += 1
        # spot-check against the single-query path (graphs are free again)
        ids = list(done)
        pick = np.random.default_rng(seed + 8).choice(len(ids), min(args.verify, len(ids)), replace=False)
        mism = 0
        for p in pick:
            r = done[ids[p]]
            g = graphs[r["graph"]]
            _, st = g.bfs(r["source"], want_values=False) if r["kind"] == 0 else g.sssp(r["source"], want_values=False)
            mism += st["checksum"] != r["checksum"]
        qps = len(t_arr) / wall
        res = {"queries": len(t_arr), "served_qps": round(qps, 1), "p50_ms": round(_pct(lat, 50), 3),
               "p99_ms": round(_pct(lat, 99), 3), "plain_qps": round(float(plain_qps), 1),
               "plain_p50_ms": round(_pct(plat, 50), 3), "plain_p99_ms": round(_pct(plat, 99), 3),
               "speedup_qps": round(float(qps / plain_qps), 2), "plain_topdown_qps": round(float(td_qps), 1),
               "speedup_vs_topdown": round(float(qps / td_qps), 2), "errors": bad, "rounds": info["rounds"],
               "checked": int(len(pick)), "mismatches": int(mism),
               "slots": {"bfs": info0["bfs_slots"], "sssp": info0["sssp_slots"]}, "state_bytes": info0["state_bytes"]}
        for kk, name in ((0, "bfs"), (1, "sssp")):
            if lat_kind[kk]:
                res[f"{name}_p50_ms"] = round(_pct(lat_kind[kk], 50), 3)
                res[f"{name}_p99_ms"] = round(_pct(lat_kind[kk], 99), 3)
        res["per_graph"] = {
            k: {"n": len(v), "p50_ms": round(_pct([x[0] for x in v], 50), 3),
                "p99_ms": round(_pct([x[0] for x in v], 99), 3),
                "queue_p99_ms": round(_pct([x[1] for x in v], 99), 3),
                "mean_rounds": round(float(np.mean([x[2] for x in v])), 1),
                "ms_per_round": round(float(np.mean([(x[0] - x[1]) / max(x[2], 1) for x in v])), 3)}
            for k, v in sorted(lat_graph.items())}
        return res

    per_mix = {}
    for mi, share in enumerate(mixes):
        results = {}
        for ri, rate in enumerate(rates):
            results[str(int(rate))] = run_rate(rate, 3005 + 10 * mi + ri, share)
            log(f"bfs share {share} rate {rate}: {results[str(int(rate))]}")
        ok = [r for r in results.values() if r["p99_ms"] <= args.p99_target_ms]
        best = max(ok, key=lambda x: x["served_qps"]) if ok else results[str(int(rates[0]))]
        per_mix[f"bfs{int(round(100 * share))}"] = {"bfs_share": share, "rates": results,
                                                    "p99_target_met": bool(ok), "best": best}
    first = per_mix[next(iter(per_mix))]
    best = first["best"]
    out = base_record(best["served_qps"], "queries/s", len(rates) * len(mixes), 0,
                      1e3 / max(best["served_qps"], 1e-9), "u32",
                      {"workload": "served bfs/sssp queries, rmat " + "/".join(map(str, scales)) + " ef16",
                       "rates": rates, "mixes": mixes, "duration_s": args.duration, "state_budget_bytes": budget,
                       "p99_target_ms": args.p99_target_ms, "sssp_delta": args.delta,
                       "served_sssp_delta": args.serve_sssp_delta or args.delta,
                       "sssp_lanes": 0 if args.no_sssp_lanes else args.sssp_lanes,
                       "sssp_batch": bool(args.sssp_batch or args.no_sssp_lanes),
                       "sssp_running_max": args.sssp_running_max,
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                       "split_weight": args.split_weight})
    out.update({"serve": per_mix,
                "parity": [{"mix": k, "rate": r, "checked": v["checked"], "mismatches": v["mismatches"]}
                           for k, m in per_mix.items() for r, v in m["rates"].items()],
                "plain_service_ms": {f"g{k[0]}_{'sssp' if k[1] else 'bfs'}": round(1e3 * float(np.mean(v)), 3)
                                     for k, v in service.items()},
                "plain_topdown_service_ms": {f"g{k[0]}_{'sssp' if k[1] else 'bfs'}": round(1e3 * float(np.mean(v)), 3)
                                             for k, v in service_td.items()}})
    print(json.dumps(out), flush=True)
    for g in graphs:
        g.close()


def wl_sssp_batch(args):
    """The query server's batched SSSP engine against single-source runs:
    16 SSSP queries submitted at once to a server over one device-resident
    R-MAT graph (scale 24 by default, ef 16, seeds 1006 / 2006, sources seed
    5), for several bucket widths with and without the light/heavy split,
    against 16 consecutive og_sssp runs (split 8, delta 256).  Every result's
    checksum is compared with the single-source run's."""
    import onegraph_amd as og

    g = og.Graph.rmat(args.scale, args.edge_factor, seed=1006, weight_seed=2006)
    ptr, to, w, _ = g.export()
    E = len(to)
    cand = np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), E)) > 0)
    srcs = [int(x) for x in np.random.default_rng(5).choice(cand, 16, replace=False)]
    log(f"rmat scale {args.scale} E={E} ready")
    g.prepare("ascetic", edge_budget_bytes=E * 8, advisek=1.0, split_weight=8)
    g.set_tuning(delta=256)
    g.sssp(srcs[0], want_values=False)
    ref, t = {}, time.perf_counter()
    for s in srcs:
        ref[s] = g.sssp(s, want_values=False)[1]["checksum"]
    plain_ms = 1e3 * (time.perf_counter() - t)
    rows = []
    for split, sd in ((0, 64), (8, 64), (8, 128), (8, 256), (8, 512)):
        g.prepare("ascetic", edge_budget_bytes=E * 8, advisek=1.0, split_weight=split)
        with og.Server([g], sssp_delta=sd, flags=og.SERVER_NO_SSSP_LANES) as srv:
            first = srv.submit(0, og.QUERY_SSSP, srcs[0])
            while not srv.poll(16, timeout_ms=100):
                pass
            r0 = srv.info()["rounds"]
            t = time.perf_counter()
            ids = [srv.submit(0, og.QUERY_SSSP, s) for s in srcs]
            got = {}
            while len(got) < len(ids):
                for r in srv.poll(16, timeout_ms=100):
                    got[r["id"]] = r
                if time.perf_counter() - t > 120:
                    raise RuntimeError("batched SSSP did not finish")
            ms = 1e3 * (time.perf_counter() - t)
            rows.append({"split_weight": split, "delta": sd, "ms": round(ms, 2), "rounds": srv.info()["rounds"] - r0,
                         "checksums_equal": all(got[i]["checksum"] == ref[s] for i, s in zip(ids, srcs))})
        log(f"batched split {split} delta {sd}: {rows[-1]}")
    best = min(rows, key=lambda r: r["ms"])
    out = base_record(round(16 / (best["ms"] * 1e-3), 2), "queries/s", 1, 1, best["ms"], "u32",
                      {"workload": f"16 concurrent sssp queries, batched server engine, rmat{args.scale} ef16",
                       "V": g.num_vertices, "E": E, "graph_in_hbm": True})
    out.update({"batched": rows, "plain_16_ms": round(plain_ms, 2),
                "plain_queries_per_s": round(16 / (plain_ms * 1e-3), 2),
                "speedup_vs_plain": round(plain_ms / best["ms"], 2),
                "parity": {"checksums_equal": all(r["checksums_equal"] for r in rows)}})
    print(json.dumps(out), flush=True)
    g.close()


WORKLOADS = {"bfs-hbm": wl_bfs_hbm, "cc-ooc": wl_cc_ooc, "pr-ooc": wl_pr_ooc, "msbfs": wl_msbfs, "serve": wl_serve,
             "sssp-batch": wl_sssp_batch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=None,
                    help="log2 V (default: 24 sssp-ooc, 20 bfs-hbm, 25 cc/pr-ooc, 26 msbfs)")
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--budget", type=float, default=0.25, help="edge budget / CSR edge bytes")
    ap.add_argument("--advisek", type=float, default=0.8)
    ap.add_argument("--approach", default="ascetic", choices=["ascetic", "subway", "usm"])
    ap.add_argument("--compare", action="store_true", help="also time subway and usm")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--delta", type=int, default=None,
                    help="SSSP near-far bucket width (0: plain rounds; default 256 with the light/heavy split, else 32)")
    ap.add_argument("--defer", type=int, default=None, choices=[0, 1, 2],
                    help="dynamic-region order: 1 held back while static vertices are active, "
                         "2 lagged behind the static part of each near-far bucket")
    ap.add_argument("--autotune", action="store_true", help="run the autotuner before timing")
    ap.add_argument("--loader", default="zerocopy", choices=["zerocopy", "copy"],
                    help="dynamic region: relax reads pinned host memory in place, or host gather + copy stream")
    ap.add_argument("--no-pack", action="store_true", help="8-byte weighted edges in HBM and on the link")
    ap.add_argument("--diropt", action="store_true", help="bfs-hbm: direction-optimising BFS (in-edge CSR in HBM)")
    ap.add_argument("--no-overlap", action="store_true", help="serialise copies, static and dynamic relax")
    ap.add_argument("--no-fuse", action="store_true",
                    help="zero-copy rounds: static and dynamic relax as two launches on two streams")
    ap.add_argument("--sweep-delta", default="", help="comma list: time SSSP at each delta and report")
    ap.add_argument("--source-seed", type=int, default=3002, help="sssp-ooc: seed of the source pick")
    ap.add_argument("--dyn-grid", type=int, default=None, help="workgroups of the dynamic-region relax (0: 8192)")
    ap.add_argument("--split-weight", type=int, default=None,
                    help="SSSP light/heavy split: edges of weight < this are light (0: off; default 8 for "
                         "sssp-ooc, off elsewhere)")
    ap.add_argument("--sweep-dyn-grid", default="", help="comma list: time SSSP at each dyn_grid and report")
    ap.add_argument("--link-probe", action="store_true",
                    help="only measure shader reads of pinned host memory and hipMemcpy (host-link bound)")
    ap.add_argument("--gather-probe", action="store_true",
                    help="only measure random 4-byte reads from device tables of 4 MiB .. 4 GiB")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--workload", default="sssp-ooc",
                    choices=["sssp-ooc", "bfs-hbm", "cc-ooc", "pr-ooc", "msbfs", "serve", "sssp-batch"],
                    help="sssp-ooc is the headline line; the others measure the remaining configurations")
    ap.add_argument("--sources", type=int, default=1024, help="msbfs: number of sources")
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="msbfs: single-source launches timed (1024: all of the sources)")
    ap.add_argument("--cpu-sample", type=int, default=8, help="msbfs: sources of the CPU baseline")
    ap.add_argument("--rates", default="200,1000,5000", help="serve: arrival rates (queries/s)")
    ap.add_argument("--duration", type=float, default=3.0, help="serve: seconds of arrivals per rate")
    ap.add_argument("--serve-scales", default="22,24,26", help="serve: R-MAT scales of the graphs")
    ap.add_argument("--serve-budget-gib", type=float, default=0.0, help="serve: per-query state budget")
    ap.add_argument("--sssp-lanes", type=int, default=2, help="serve: single-source SSSP lanes per weighted graph")
    ap.add_argument("--sssp-running-max", type=int, default=0,
                    help="serve: SSSP lanes running at once over all graphs (0: no cap)")
    ap.add_argument("--sssp-batch", action="store_true",
                    help="serve: the batched SSSP engine beside the lanes (admits when the lanes are busy and 8 wait)")
    ap.add_argument("--no-sssp-lanes", action="store_true",
                    help="serve: no single-source SSSP lanes (every SSSP query in the batched engine)")
    ap.add_argument("--verify", type=int, default=100, help="serve: served results re-run on the plain path")
    ap.add_argument("--serve-mix", default="0.7,1.0", help="serve: BFS shares of the query mixes to run")
    ap.add_argument("--p99-target-ms", type=float, default=50.0, help="serve: p99 latency target")
    ap.add_argument("--serve-sssp-delta", type=int, default=0,
                    help="serve: bucket width of the server's batched SSSP slots (0: the graphs' tuning delta, "
                         "256 with the light/heavy split)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="serve: GPU_MAX_HW_QUEUES for the process (each engine's and lane's stream gets its own "
                         "hardware queue instead of sharing one of HIP's default 4; at most 32; 0: one per stream "
                         "of the three graphs, their engines and lanes)")
    args = ap.parse_args()
    if args.scale is None:
        args.scale = {"sssp-ooc": 24, "bfs-hbm": 20, "cc-ooc": 25, "pr-ooc": 25, "msbfs": 26, "serve": 0,
                      "sssp-batch": 24}[args.workload]
    if args.workload == "serve":
        # before the first HIP call: the graphs' own streams (3 each), the
        # server's engine streams (BFS and batched SSSP per graph) and its
        # SSSP lanes' streams (one per lane), 21 for three graphs with two
        # lanes each, each get a hardware queue, so one engine's kernels do
        # not wait behind another's in a shared queue
        ng = len(args.serve_scales.split(","))
        batch = args.sssp_batch or args.no_sssp_lanes
        hwq = args.hw_queues or 3 * ng + ng * (2 if batch else 1) + ng * (0 if args.no_sssp_lanes else args.sssp_lanes)
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, hwq))
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the engine runs on one GPU; sharding is not implemented); --gpus {args.gpus} refused",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration_mib:
        # A copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel give the
        # factor between the counters and bytes actually moved.
        import torch

        n = args.copy_calibration_mib * (1 << 20) // 4
        a = torch.arange(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "reps": 3}))
        return

    import onegraph_amd as og

    if og.device_count() < 1:
        print("bench.py needs a GPU: no HIP device is visible", file=sys.stderr)
        sys.exit(1)
    if args.gather_probe:
        # random 4-byte reads from tables the size of the relax's value arrays
        # (C1 4 MiB, C2 64 MiB, C4 256 MiB) and from HBM proper (4 GiB)
        print(json.dumps({"gather_probe": [og.probe_hbm_gather(b) for b in (4 << 20, 64 << 20, 256 << 20, 4 << 30)],
                          "hbm_peak_gbs": HBM_PEAK_GBS}))
        return
    if args.link_probe:
        # what the host link gives shader reads (the zero-copy loader's bound)
        print(json.dumps({"link_probe": og.probe_host_link(1 << 30), "peak_gbs": LINK_PEAK_GBS}))
        return
    if args.workload != "sssp-ooc":
        if args.split_weight is None:  # the served SSSP lanes take the headline's split too
            args.split_weight = 8 if args.workload == "serve" else 0
        if args.delta is None:
            args.delta = 256 if args.split_weight else 32
        WORKLOADS[args.workload](args)
        return
    # headline SSSP: delta-stepping with light edges (weight < 8) relaxed in
    # the buckets' rounds and heavy ones once per settled vertex; with the
    # split the work hardly depends on the bucket width, so wide buckets
    # (fewer rounds) pay
    if args.split_weight is None:
        args.split_weight = 8
    if args.delta is None:
        args.delta = 256 if args.split_weight else 32

    t0 = time.time()
    g = og.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=1002, weight_seed=2002)
    V, E = g.num_vertices, g.num_edges
    log(f"graph: rmat scale {args.scale} V={V} E={E} generated in {time.time() - t0:.1f}s")
    ptr, to, w, _ = g.export()
    src, deg = pick_source(ptr, E, args.source_seed)
    budget = int(E * 8 * args.budget)

    def run_approach(approach, steps, warmup, loader=None):
        k = args.advisek if approach == "ascetic" else 0.0
        flags = plan_flags(args, approach, loader)
        sw = args.split_weight if approach != "usm" else 0
        plan = g.prepare(approach, edge_budget_bytes=budget if approach != "usm" else 0, advisek=k, flags=flags,
                         split_weight=sw)
        log(f"{approach}: plan {plan}")
        if args.delta is not None:
            g.set_tuning(delta=args.delta)
        if args.defer is not None:
            g.set_tuning(defer=args.defer)
        if args.dyn_grid is not None:
            g.set_tuning(dyn_grid=args.dyn_grid)
        if args.autotune:
            best, best_ms = g.autotune(og.TUNE_SSSP, src, runs=1)
            log(f"{approach}: autotuned {best} ({best_ms:.1f} ms)")
        if args.sweep_dyn_grid:
            for dg in [int(x) for x in args.sweep_dyn_grid.split(",")]:
                g.set_tuning(dyn_grid=dg)
                ts = [g.sssp(src, want_values=False)[1] for _ in range(3)]
                log(f"{approach}: dyn_grid {dg}: {min(t['total_ms'] for t in ts):.2f} ms")
            g.set_tuning(dyn_grid=args.dyn_grid or 0)
        if args.sweep_delta:
            for dl in [int(x) for x in args.sweep_delta.split(",")]:
                g.set_tuning(delta=dl)
                ts = [g.sssp(src, want_values=False)[1] for _ in range(3)]
                log(f"{approach}: delta {dl}: {min(t['total_ms'] for t in ts):.2f} ms, "
                    f"{ts[-1]['iterations']} rounds, h2d {ts[-1]['h2d_bytes'] / 1e9:.2f} GB, "
                    f"work {ts[-1]['work_edges'] / 1e6:.0f} M")
            if args.delta is not None:
                g.set_tuning(delta=args.delta)
        log(f"{approach}: tuning {g.tuning()}")
        for i in range(warmup):
            _, st = g.sssp(src, want_values=False)
            log(f"{approach}: warmup {i} {st['total_ms']:.1f} ms, {st['iterations']} iterations")
        stats = []
        t = time.perf_counter()
        for _ in range(steps):
            _, st = g.sssp(src, want_values=False)
            stats.append(st)
        wall = time.perf_counter() - t
        log(f"{approach}: {steps} steps in {wall:.3f}s")
        return plan, stats, wall

    plan, stats, wall = run_approach(args.approach, args.steps, args.warmup)
    # one checked run: distances must equal the oracle's (bit-exact)
    dist, st_chk = g.sssp(src)
    ms_per_step = 1e3 * wall / args.steps
    # the timed steps leave the distances in HBM; the reference's sssp_opt
    # copies them back inside its total time, so the same steps with the
    # device-to-host copy of the 4V-byte distance array are timed beside them
    cb_steps = max(1, min(args.steps, 5))
    t = time.perf_counter()
    for _ in range(cb_steps):
        g.sssp(src)
    ms_with_copyback = 1e3 * (time.perf_counter() - t) / cb_steps
    trav = stats[-1]["traversed_edges"]
    gteps = trav / (ms_per_step * 1e-3) / 1e9

    # weighted edges in HBM and on the link: packed (to << 6 | w) or the CSR's {to, w}
    edge_format = "8-byte {to,w}" if args.no_pack else "4-byte to<<6|w"
    agg = {}
    for key in ["iterations", "work_edges", "h2d_bytes", "chunks", "gather_ms", "copy_ms"]:
        agg[key] = sum(s[key] for s in stats)
    # HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE (calibrated against a plain copy) on this
    # workload and setting, when a committed profile holds them
    traffic = traffic_src = None
    fused = (sum(x["relax_static_kernel_ms"] for x in stats) == 0 and sum(x["relax_static_bytes"] for x in stats) > 0
             and sum(x["relax_dynamic_kernel_ms"] for x in stats) > 0)
    kclass = "relax_round" if fused else "relax_dynamic"
    tpath = os.path.join(HERE, "profiles", f"k_{kclass.replace('relax_dynamic', 'relax_dyn')}_sssp_rmat24_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        want = {"scale": args.scale, "approach": args.approach, "loader": args.loader, "delta": args.delta,
                "split_weight": args.split_weight, "kernel_class": kclass, "edge_format": edge_format}
        if tr.get("config") == want:
            traffic, traffic_src = tr["traffic_bytes_per_launch"], tr["source"]
    roof, link, kt = roofline_of(stats, zero_copy=args.loader == "zerocopy" and args.approach != "usm",
                                 traffic=traffic)
    if roof and traffic_src:
        roof["traffic_source"] = traffic_src
    if link:
        link["bytes_per_step"] = agg["h2d_bytes"] / args.steps
    # the in-HBM static relax beside it (its own roofline)
    static = None
    if kt.get("relax_dynamic", {}).get("fused"):
        static = {"note": "relaxed inside k_relax_round beside the dynamic region (one launch), so not timed "
                          "apart; its bytes are in the roofline's hbm figure"}
    elif "relax_static" in kt:
        r = kt["relax_static"]
        static = {"bound": "hbm", "achieved": round(r["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(r["gbs"] / HBM_PEAK_GBS, 4), "avg_launch_us": round(r["avg_us"], 3),
                  "note": "overlaps the dynamic relax on another stream, so its launch times include sharing "
                          "the CUs with it"}

    out = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"sssp-out-of-memory rmat{args.scale} ef{args.edge_factor} {args.approach}",
            "graph": f"R-MAT scale {args.scale} ef {args.edge_factor} a/b/c/d .57/.19/.19/.05 seed 1002, "
                     f"dedup, no self-loops, u32 weights [1,63] seed 2002",
            "V": V, "E": E, "source": src,
            "edge_budget_bytes": budget, "budget_fraction": args.budget, "advisek": args.advisek,
            "approach": args.approach, "static_edges": plan["static_edges"], "staging_edges": plan["staging_edges"],
            "vsplit": plan["vsplit"], "loader": args.loader, "delta": args.delta, "split_weight": args.split_weight,
            "edge_format": edge_format,
        },
        "result_copy_back": {"in_value": False, "bytes": 4 * V, "ms_per_step_with_copy_back": round(ms_with_copyback, 3),
                             "gteps_with_copy_back": round(trav / (ms_with_copyback * 1e-3) / 1e9, 4),
                             "steps": cb_steps,
                             "note": "value times steps whose distances stay in HBM; these steps also copy them "
                                     "to a host array, as the reference's total time does"},
        "roofline": roof,
        "static_region": static,
        "host_link": link,
        "kernels": kt,
        "per_step": {"iterations": agg["iterations"] / args.steps, "work_edges": agg["work_edges"] / args.steps,
                     "h2d_bytes": agg["h2d_bytes"] / args.steps, "chunks": agg["chunks"] / args.steps,
                     "gather_ms": agg["gather_ms"] / args.steps, "copy_ms": agg["copy_ms"] / args.steps,
                     "traversed_edges": trav, "reached": stats[-1]["reached"]},
    }

    if args.compare:
        cmp = {}
        for ap_name in ["ascetic", "subway", "usm"]:
            if ap_name == args.approach:
                continue
            # Subway as designed: the active subgraph is gathered and copied
            p2, s2, w2 = run_approach(ap_name, max(1, args.steps // 2), 1, loader="copy")
            ms2 = 1e3 * w2 / len(s2)
            cmp[ap_name] = {"ms_per_step": round(ms2, 3), "gteps": round(trav / (ms2 * 1e-3) / 1e9, 4),
                            "h2d_bytes": s2[-1]["h2d_bytes"], "iterations": s2[-1]["iterations"],
                            "loader": "copy" if ap_name == "subway" else "mapped host memory"}
        out["compare"] = cmp

    # correctness of the measured path + CPU baseline (oracle, single thread)
    import oracle

    ref = oracle.sssp(ptr, to, w, src)
    cpu_s = oracle.loop_seconds()
    out["parity"] = {"bit_exact": bool(np.array_equal(dist, ref)), "reached": int((ref != og.UNREACHED).sum())}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_record(round(trav / cpu_s / 1e9, 5), "GTEPS", 1,
                                         f"one full Dijkstra from the same source on the same graph ({cpu_s:.2f} s)")
    print(json.dumps(out), flush=True)
    g.close()
    if not out["parity"]["bit_exact"]:
        sys.exit(3)


if __name__ == "__main__":
    main()
