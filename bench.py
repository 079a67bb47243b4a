#!/usr/bin/env python3
"""Headline benchmark: out-of-GPU-memory SSSP on one MI355X.

Workload (the metric's single-GPU configuration): directed R-MAT scale 24,
edge factor 16, a/b/c/d = .57/.19/.19/.05 (graph seed 1002), self-loops and
duplicates removed, u32 weights uniform in [1, 63] (seed 2002); device edge
budget = 25% of the CSR edge+weight bytes; Ascetic plan with advisek = 0.8;
source = seeded uniform pick (seed 3002) among vertices with out-degree > 0.

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
    """Per-class algorithmic GB/s from the engine's HIP-event timings."""
    rows = {}
    for cls, ms_key, n_key, b_key in [
        ("relax_static", "relax_static_kernel_ms", "relax_static_launches", "relax_static_bytes"),
        ("relax_dynamic", "relax_dynamic_kernel_ms", "relax_dynamic_launches", "relax_dynamic_bytes"),
        ("frontier", "frontier_kernel_ms", "frontier_launches", "frontier_bytes"),
    ]:
        ms, n, b = st[ms_key], st[n_key], st[b_key]
        if n and ms > 0:
            rows[cls] = {"ms": ms, "launches": n, "avg_us": 1e3 * ms / n, "bytes": b,
                         "bytes_per_launch": b / n, "gbs": b / (ms * 1e-3) / 1e9}
    return rows


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--budget", type=float, default=0.25, help="edge budget / CSR edge bytes")
    ap.add_argument("--advisek", type=float, default=0.8)
    ap.add_argument("--approach", default="ascetic", choices=["ascetic", "subway", "usm"])
    ap.add_argument("--compare", action="store_true", help="also time subway and usm")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); --gpus {args.gpus} refused",
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

    t0 = time.time()
    g = og.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=1002, weight_seed=2002)
    V, E = g.num_vertices, g.num_edges
    log(f"graph: rmat scale {args.scale} V={V} E={E} generated in {time.time() - t0:.1f}s")
    ptr, to, w, _ = g.export()
    src, deg = pick_source(ptr, E, 3002)
    budget = int(E * 8 * args.budget)

    def run_approach(approach, steps, warmup):
        k = args.advisek if approach == "ascetic" else 0.0
        plan = g.prepare(approach, edge_budget_bytes=budget if approach != "usm" else 0, advisek=k)
        log(f"{approach}: plan {plan}")
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
    trav = stats[-1]["traversed_edges"]
    gteps = trav / (ms_per_step * 1e-3) / 1e9

    agg = {}
    for key in ["relax_static_kernel_ms", "relax_static_launches", "relax_static_bytes", "relax_dynamic_kernel_ms",
                "relax_dynamic_launches", "relax_dynamic_bytes", "frontier_kernel_ms", "frontier_launches",
                "frontier_bytes", "copy_ms", "h2d_bytes", "gather_ms", "chunks", "iterations", "work_edges"]:
        agg[key] = sum(s[key] for s in stats)
    kt = kernel_table(agg)
    dom = max(kt, key=lambda c: kt[c]["ms"]) if kt else None
    roof = None
    if dom:
        r = kt[dom]
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(r["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(r["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(r["avg_us"], 3), "bytes_per_launch": round(r["bytes_per_launch"], 1)}
        # HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
        # (calibrated against a plain copy) for this kernel and workload.
        tpath = os.path.join(HERE, "profiles", "k_relax_dyn_sssp_rmat24_traffic.json")
        if dom == "relax_dynamic" and args.scale == 24 and args.approach == "ascetic" and os.path.exists(tpath):
            with open(tpath) as f:
                tr = json.load(f)
            roof["traffic"] = tr["traffic_bytes_per_launch"]
            roof["traffic_source"] = tr["source"]
    link = None
    if agg["copy_ms"] > 0:
        gbs = agg["h2d_bytes"] / (agg["copy_ms"] * 1e-3) / 1e9
        link = {"achieved": round(gbs, 2), "peak": LINK_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / LINK_PEAK_GBS, 4),
                "bytes_per_step": agg["h2d_bytes"] / args.steps}

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
            "vsplit": plan["vsplit"],
        },
        "roofline": roof,
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
            p2, s2, w2 = run_approach(ap_name, max(1, args.steps // 2), 1)
            ms2 = 1e3 * w2 / len(s2)
            cmp[ap_name] = {"ms_per_step": round(ms2, 3), "gteps": round(trav / (ms2 * 1e-3) / 1e9, 4),
                            "h2d_bytes": s2[-1]["h2d_bytes"], "iterations": s2[-1]["iterations"]}
        out["compare"] = cmp

    # correctness of the measured path + CPU baseline (oracle, single thread)
    import oracle

    t = time.perf_counter()
    ref = oracle.sssp(ptr, to, w, src)
    cpu_s = time.perf_counter() - t
    out["parity"] = {"bit_exact": bool(np.array_equal(dist, ref)), "reached": int((ref != og.UNREACHED).sum())}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = {"value": round(trav / cpu_s / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
                               "sample": f"one full Dijkstra from the same source on the same graph ({cpu_s:.2f} s)",
                               "cpu": cpu_info(), "nproc": os.cpu_count()}
    print(json.dumps(out), flush=True)
    g.close()
    if not out["parity"]["bit_exact"]:
        sys.exit(3)


if __name__ == "__main__":
    main()
