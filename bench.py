#!/usr/bin/env python3
"""Benchmark of the gfx950 semi-naive engine.

Default workload (one GPU): multi-source Trop cheapest cost on an RMAT graph
(scale 20: 1,048,576 nodes, 16,777,216 draws -> ~16.1M distinct edges,
weights uniform [1, 1000], 64 sources):
    cost(S,Y) :- src(S), edge(S,Y).   cost(S,Z) :- cost(S,Y), edge(Y,Z).
A step is one complete semi-naive fixpoint over the EDB already resident in
HBM (relations reset to the loaded state before each step).  value = derived
tuples (|cost|) per second.

Also: --config 1|3|4|5 for the other workloads (TC, triangles, batched,
served) used by the documentation; they print their own JSON line.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "strata-k_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def roofline_of(top, traffic=None):
    if not top or top["ms"] <= 0:
        return None
    per_launch_bytes = top["bytes"] / top["launches"]
    avg_ms = top["ms"] / top["launches"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "kernel": top["name"],
            "launches": top["launches"], "avg_launch_ms": round(avg_ms, 4),
            "alg_bytes_per_launch": round(per_launch_bytes)}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprof PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline_cfg2(wl_full, k_sources=1):
    """The oracle (single-threaded C restatement of strata-eval semi-naive)
    on the same graph restricted to the first k sources: cost(S,.) depends on
    S only, so this is an exact sub-workload."""
    from oracle import oracle as O
    from strata_k import workloads as wl
    ws = wl.restrict_sources(wl_full, k_sources)
    r = O.run(ws.program, bulk=ws.bulk, want=["cost"])
    n = len(r.relations["cost"])
    return {"value": round(n / r.seconds, 1), "unit": "tuples/s", "cores": 1, "kind": "port",
            "sample": f"same RMAT-20 graph, first {k_sources} of 64 sources: {n} cost tuples, "
                      f"{r.iterations[0]} rounds, {r.seconds:.1f} s of semi-naive evaluation (EDB load excluded)",
            "host_cpu": platform.processor() or platform.machine()}


def run_steps(ctx, run, steps, warmup):
    """Times `steps` fixpoints.  Every launch in the timed steps is bracketed by
    HIP events on the engine's stream (no host synchronisation between
    launches); the per-kernel records are summed over the timed steps."""
    for _ in range(warmup):
        run.reset()
        run.execute()
    ctx.synchronize()
    ctx.set_profiling(True)
    acc = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        run.reset()
        run.execute()
        for r in run.kernel_records():
            a = acc.setdefault(r["name"], dict(r, launches=0, ms=0.0, bytes=0.0, items=0))
            for k in ("launches", "ms", "bytes", "items"):
                a[k] += r[k]
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(1, steps)
    ctx.set_profiling(False)
    return ms, list(acc.values())


def copy_probe(mib):
    """One device-to-device pass of `mib` MiB through an elementwise kernel
    (int64 negation: reads and writes exactly mib MiB each), the known-size
    launch that calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.int64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    log(f"copy probe: {mib} MiB read, {mib} MiB written")


def bench_cfg2(args):
    import strata_k as sk
    from strata_k import workloads as wl
    t = time.time()
    w = wl.config2(seed=args.seed, scale=args.scale)
    log(f"generated {w.params} in {time.time() - t:.1f}s")
    ctx = sk.Context(0)
    run = ctx.prepare(w.program)
    for p, (cols, wt, inf) in w.bulk.items():
        run.add_sym(p, cols, wt, inf)
    t = time.time()
    run.load()
    log(f"EDB resident in HBM ({time.time() - t:.2f}s incl. H2D + device sort)")
    ms, recs = run_steps(ctx, run, args.steps, args.warmup)
    n_cost = run.size("cost")
    iters = run.iterations()
    top = max(recs, key=lambda r: r["ms"]) if recs else None
    roof = roofline_of(top, measured_traffic(top["name"]) if top else None)
    stats = {}
    for r in recs:
        c = stats.setdefault(r["cls"], {"ms": 0.0, "bytes": 0.0})
        c["ms"] += r["ms"] / args.steps
        c["bytes"] += r["bytes"] / args.steps
    all_kernel_ms = sum(c["ms"] for c in stats.values())
    all_kernel_bytes = sum(c["bytes"] for c in stats.values())
    line = {
        "metric": "derived tuples/s", "value": round(n_cost / (ms / 1e3), 1), "unit": "tuples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded RMAT a=.57 b=.19 c=.19, weights uniform [1,1000])",
        "config": {"workload": "trop multi-source cheapest cost, RMAT scale 20 (1,048,576 nodes, "
                               f"{w.params['n_edges']} edges), 64 sources",
                   "derived_tuples": n_cost, "semi_naive_rounds": iters[0], "seed": args.seed},
        "roofline": roof,
        "kernel_ms_per_step": {k: round(v["ms"], 3) for k, v in stats.items()},
        "kernel_ms_total": round(all_kernel_ms, 3),
        "hbm_gbps_all_kernels": round(all_kernel_bytes / (all_kernel_ms * 1e-3) / 1e9, 1) if all_kernel_ms else None,
    }
    if args.kernels:
        log(json.dumps(sorted(recs, key=lambda r: -r["ms"]), indent=1))
    run.free()
    ctx.close()
    if not args.no_cpu:
        log("timing the CPU baseline (oracle, 1 source) ...")
        line["cpu_baseline"] = cpu_baseline_cfg2(w, 1)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def bench_cfg1(args):
    import strata_k as sk
    from strata_k import workloads as wl
    w = wl.config1()
    ctx = sk.Context(0)
    run = ctx.prepare(w.program)
    run.add_sym("edge", *w.bulk["edge"][:1], None)
    run.load()
    ms, recs = run_steps(ctx, run, args.steps, args.warmup)
    n = run.size("path")
    top = max(recs, key=lambda r: r["ms"]) if recs else None
    print(json.dumps({"config": "bool_tc_uniform_2000_10000", "derived_tuples": n, "rounds": run.iterations()[0],
                      "ms_per_step": round(ms, 3), "tuples_per_s": round(n / (ms / 1e3), 1),
                      "kernel_ms_per_step": round(sum(r["ms"] for r in recs) / args.steps, 3),
                      "launches_per_step": sum(r["launches"] for r in recs) // args.steps,
                      "roofline": roofline_of(top)}), flush=True)


def bench_cfg3(args):
    import strata_k as sk
    from strata_k import workloads as wl
    t = time.time()
    w = wl.config3(scale=args.scale if args.scale != 20 else 22)
    log(f"generated {w.params} in {time.time() - t:.1f}s")
    ctx = sk.Context(0)
    run = ctx.prepare(w.program)
    run.add_sym("e", *w.bulk["e"][:1], None)
    run.load()
    ms, recs = run_steps(ctx, run, args.steps, args.warmup)
    n = run.size("tri")
    ntri = run.relation("ntri").rows()
    top = max(recs, key=lambda r: r["ms"]) if recs else None
    if args.kernels:
        log(json.dumps(sorted(recs, key=lambda r: -r["ms"]), indent=1))
    print(json.dumps({"config": "wcoj_triangles_rmat", "params": w.params, "tri": n, "ntri": str(ntri),
                      "ms_per_step": round(ms, 3), "tuples_per_s": round(n / (ms / 1e3), 1),
                      "roofline": roofline_of(top)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one known-size copy kernel (PMC counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (per-GPU shards would be "
              "independent source sets); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    {1: bench_cfg1, 2: bench_cfg2, 3: bench_cfg3}[args.config](args)


if __name__ == "__main__":
    main()
