#!/usr/bin/env python3
"""Benchmark of the gfx950 semi-naive engine.

Default workload (one GPU): multi-source Trop cheapest cost on an RMAT graph
(scale 20: 1,048,576 nodes, 16,777,216 draws -> ~16.1M distinct edges,
weights uniform [1, 1000], 64 sources):
    cost(S,Y) :- src(S), edge(S,Y).   cost(S,Z) :- cost(S,Y), edge(Y,Z).
A step is one complete semi-naive fixpoint over the EDB already resident in
HBM (relations reset to the loaded state before each step).  value = derived
tuples (|cost|) per second.

The timed steps drive the rounds from the host (--graph 0), so that every
launch, the pull kernel's included, is bracketed by its own HIP events.

The same JSON line also carries, as extra keys:
  device_driven  the same fixpoints with the engine's default round driving
               (device-driven rounds in hipGraph launches), wall clock only;
  end_to_end   config 2 timed as the reference times its benchmark (host EDB
               in, host result out), and with the result left in HBM;
  batched      config 4: 512 programs in one batch against 512 plain runs
               (execute-phase and end-to-end speedups, bit-exact check);
  served       config 5: Poisson queries on a resident RMAT-24 graph (p50/p99,
               queries/s, batched vs one at a time, answer contents compared
               with plain runs);
  bool_tc      config 1: Bool transitive closure, 2,000 nodes / 10,000 edges;
  triangles    config 3: triangles of RMAT scale 22, listed and counted;
  sort         the radix sort primitive on 2^26 device-resident keys, alone
               and with i64 weights, at the autotuner's pick (checked against
               torch.sort);
  cpu_baseline the oracle (C restatement of strata-eval) on host threads, one
               source per run; cpu_scan_config1 the faithful scan-mode
               interpreter's labelled extrapolation to config 1.
--no-extras leaves out batched, served, bool_tc, triangles and sort; --no-cpu
the CPU baselines.

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

import numpy as np

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


def host_info():
    """CPU model and the cores this process may use (the CPU baselines' host)."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline_cfg2(wl_full, threads):
    """The oracle (the single-threaded C restatement of strata-eval's indexed
    semi-naive evaluation) on the same graph, one run per source, `threads`
    runs at a time on host threads: cost(S,.) depends on S only, so each run
    is an exact sub-workload and the runs together are the multi-threaded CPU
    engine's evaluation of that many sources."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from strata_k import workloads as wl
    src = wl_full.bulk["src"][0][0][:threads]

    def one(s0):
        bulk = dict(wl_full.bulk)
        bulk["src"] = ([np.array([s0], np.uint32)], np.zeros(1, np.int64), None)
        r = O.run(wl_full.program, bulk=bulk, want=["cost"])
        return len(r.relations["cost"]), r.iterations[0], r.seconds

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, src.tolist()))
    wall = time.perf_counter() - t0
    n = sum(r[0] for r in res)
    single = res[0][0] / res[0][2]
    info = host_info()
    return {"value": round(n / wall, 1), "unit": "tuples/s", "cores": threads, "kind": "port",
            "sample": f"same RMAT-20 graph, {len(src)} of 64 sources, one oracle run (indexed semi-naive "
                      f"restatement of strata-eval) per source on {threads} host threads: {n} cost tuples in "
                      f"{wall:.1f} s wall, EDB conversion included",
            "single_thread_tuples_per_s": round(single, 1), **info}


def cpu_scan_extrapolation_cfg1():
    """The oracle's scan mode (the reference's nested-loop join order, one
    core) on config 1's program over a 400-node / 2,000-edge sample, where its
    unify-step count is checked against |E| * (|path| + rounds + 2); full
    config 1's count by the same formula (from the indexed run's |path| and
    rounds) divided by the sample's steps/s is the labelled extrapolation of
    the faithful interpreter on config 1."""
    from oracle import oracle as O
    from strata_k import workloads as wl
    ws = wl.config1(seed=1, n_nodes=400, n_edges=2000)
    r = O.run(ws.program, bulk=ws.bulk, want=["path"], indexed=False)
    m = len(ws.bulk["edge"][0][0])
    formula_ok = r.steps == m * (len(r.relations["path"]) + r.iterations[0] + 2)
    rate = r.steps / r.seconds
    wf = wl.config1()
    ri = O.run(wf.program, bulk=wf.bulk, want=["path"])
    mf = len(wf.bulk["edge"][0][0])
    steps_full = mf * (len(ri.relations["path"]) + ri.iterations[0] + 2)
    return {"kind": "port, extrapolated", "cores": 1, "steps_per_s": round(rate, 1),
            "sample_steps_match_formula": bool(formula_ok), "config1_unify_steps": steps_full,
            "config1_seconds_extrapolated": round(steps_full / rate, 1),
            "config1_tuples_per_s_extrapolated": round(len(ri.relations["path"]) / (steps_full / rate), 1),
            "config1_indexed_seconds": round(ri.seconds, 2)}


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


def time_fixpoints(ctx, run, steps, warmup):
    """Wall-clock ms per fixpoint with no per-launch events (the engine as a
    caller runs it)."""
    for _ in range(warmup):
        run.reset()
        run.execute()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.reset()
        run.execute()
    ctx.synchronize()
    return (time.perf_counter() - t0) * 1e3 / max(1, steps)


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
    run.set_graph(args.graph != 0)
    run.set_pack(args.pack != 0)
    for p, (cols, wt, inf) in w.bulk.items():
        run.add_sym(p, cols, wt, inf)
    t = time.time()
    run.load()
    log(f"EDB resident in HBM ({time.time() - t:.2f}s incl. H2D + device sort)")
    ms, recs = run_steps(ctx, run, args.steps, args.warmup)
    n_cost = run.size("cost")
    iters = run.iterations()
    top = max((r for r in recs if r["bytes"] > 0), key=lambda r: r["ms"], default=None)
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
    # the engine's default round driving (device-driven rounds in hipGraphs
    # once a stratum's rounds all pull), without per-launch events: the
    # timed steps above drive rounds from the host so that every pull launch
    # has its own events for the roofline
    run.set_graph(True)
    ms_dd = time_fixpoints(ctx, run, args.steps, args.warmup)
    line["device_driven"] = {
        "ms_per_step": round(ms_dd, 3), "tuples_per_s": round(run.size("cost") / (ms_dd / 1e3), 1),
        "derived_tuples": run.size("cost"), "semi_naive_rounds": run.iterations()[0],
        "same_result": run.size("cost") == n_cost and run.iterations()[0] == iters[0],
        "note": "the engine's default (sk_run_set_graph on): pull and push rounds device-driven in hipGraph "
                "launches after the first pull round, no per-launch events; same workload and steps"}
    run.free()
    ctx.close()
    line["end_to_end"] = cfg2_end_to_end(w, n_cost)
    if not args.no_extras:
        line["batched"] = cfg4_summary(args)
        line["served"] = cfg5_summary(args)
        line["bool_tc"] = cfg1_summary(args)
        line["triangles"] = cfg3_summary(args)
        line["sort"] = sort_summary(args)
    if not args.no_cpu:
        threads = max(1, min(16, host_info()["usable_cpus"]))
        log(f"timing the CPU baseline (oracle, {threads} sources on {threads} threads) ...")
        line["cpu_baseline"] = cpu_baseline_cfg2(w, threads)
        log("timing the scan-mode extrapolation of config 1 ...")
        line["cpu_scan_config1"] = cpu_scan_extrapolation_cfg1()
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def cfg2_end_to_end(w, n_cost):
    """Config 2 as the reference times its benchmark (examples/bench.rs:53-63):
    host EDB in, fixpoint, host result out, on a fresh context; also the same
    without the result read (the answer left in HBM)."""
    import strata_k as sk
    ctx = sk.Context(0)
    t0 = time.perf_counter()
    run = ctx.prepare(w.program)
    for p, (cols, wt, inf) in w.bulk.items():
        run.add_sym(p, cols, wt, inf)
    run.load()
    run.execute()
    ctx.synchronize()
    t1 = time.perf_counter()
    (sa, za), cw, _ = run.relation_sym("cost")
    t2 = time.perf_counter()
    assert len(sa) == n_cost
    run.free()
    ctx.close()
    return {"seconds": round(t2 - t0, 4), "tuples_per_s": round(n_cost / (t2 - t0), 1),
            "seconds_result_in_hbm": round(t1 - t0, 4), "tuples_per_s_result_in_hbm": round(n_cost / (t1 - t0), 1),
            "includes": "prepare, H2D of the host EDB, device sort, fixpoint, D2H of the result"}


def cfg1_summary(args):
    """Config 1 (Bool TC, 2,000 nodes / 10,000 edges) as extra keys."""
    log("config 1: Bool TC ...")
    o = bench_cfg1(argparse.Namespace(**vars(args)), emit=False)
    return {k: o[k] for k in ("config", "derived_tuples", "rounds", "ms_per_step", "tuples_per_s",
                              "ms_per_step_event_timed", "kernel_ms_per_step", "launches_per_step")}


def cfg3_summary(args):
    """Config 3 (triangles, RMAT scale 22) as extra keys: listing and count
    per run, the dominant kernel's roofline."""
    log("config 3: triangles on RMAT scale 22 ...")
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup, a.scale = 3, 1, 22
    o = bench_cfg3(a, emit=False)
    return {k: o[k] for k in ("config", "params", "tri", "ntri", "tri_equals_ntri", "ms_per_step", "tuples_per_s",
                              "kernel_ms_per_step", "roofline")}


def sort_summary(args):
    """The radix sort primitive on 2^26 device-resident keys (44 significant
    bits), alone and with i64 weights, at the tuner's pick."""
    log("radix sort, 2^26 keys ...")
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup, a.sort_log2 = 5, 1, 26
    r = bench_sort(a, emit=False)["results"]
    return [{k: x[k] for k in ("weighted", "n", "key_bits", "candidate", "digit_bits", "items_per_thread", "block",
                               "single_sweep", "passes", "ms_per_sort", "alg_gbps", "frac_of_hbm_peak",
                               "sorted_equals_torch")} for x in r]


def cfg4_summary(args):
    """Config 4 (512 batched programs against 512 plain runs) as extra keys."""
    a = argparse.Namespace(**vars(args))
    a.programs, a.seed = 512, 4
    log("config 4: 512 programs batched and plain ...")
    r = bench_cfg4(a, emit=False)
    keep = ("value", "value_is", "end_to_end_speedup", "end_to_end_speedup_results_in_hbm", "end_to_end_note",
            "plain_s", "batched_s", "plain_read_s", "plain_device_ms", "batched_device_ms", "batched_phases",
            "bit_exact_vs_plain")
    out = {"workload": r["config"]["workload"], "derived_tuples": r["config"]["derived_tuples"]}
    out.update({("execute_speedup" if k == "value" else k): r[k] for k in keep})
    return out


def cfg5_summary(args):
    """Config 5 (served queries on a resident RMAT-24 graph) as extra keys."""
    a = argparse.Namespace(**vars(args))
    # the stated shape: source sets log-uniform in [1, 4096]
    a.scale, a.queries, a.max_sources, a.rate, a.max_batch, a.plain_queries = 24, 16, 4096, 2.0, 16, 2
    log("config 5: served queries on RMAT scale 24 ...")
    r = bench_cfg5(a, emit=False)
    out = {"workload": r["config"]["workload"]}
    for k in ("p50_ms", "p99_ms", "served_qps", "capacity_qps_batched", "capacity_qps_one_at_a_time",
              "batched_vs_one_at_a_time", "capacity_sources_per_s_batched", "plain_run_sources_per_s",
              "served_vs_plain_runs", "server_rounds", "same_rows_and_rounds_alone_batched_and_plain",
              "answers_compared_with_plain_runs", "answer_contents_equal"):
        out[k] = r[k]
    return out


def bench_cfg1(args, emit=True):
    """Config 1: Bool transitive closure of a uniform 2,000-node / 10,000-edge
    graph, fixpoints on the resident EDB."""
    import strata_k as sk
    from strata_k import workloads as wl
    w = wl.config1()
    ctx = sk.Context(0)
    run = ctx.prepare(w.program)
    run.add_sym("edge", *w.bulk["edge"][:1], None)
    run.load()
    # about 90 launches per fixpoint of a few microseconds each: per-launch
    # events would add a large share of the wall clock, so the fixpoints are
    # timed without them and a second, event-timed pass gives the kernels
    ms_plain = time_fixpoints(ctx, run, args.steps, args.warmup)
    ms, recs = run_steps(ctx, run, args.steps, 0)
    n = run.size("path")
    top = max((r for r in recs if r["bytes"] > 0), key=lambda r: r["ms"], default=None)
    out = {"config": "bool_tc_uniform_2000_10000", "derived_tuples": n, "rounds": run.iterations()[0],
           "ms_per_step": round(ms_plain, 3), "tuples_per_s": round(n / (ms_plain / 1e3), 1),
           "ms_per_step_event_timed": round(ms, 3),
           "kernel_ms_per_step": round(sum(r["ms"] for r in recs) / args.steps, 3),
           "launches_per_step": sum(r["launches"] for r in recs) // args.steps,
           "roofline": roofline_of(top)}
    run.free()
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_cfg3(args, emit=True):
    """Config 3: triangles of an RMAT scale-22 graph (64 M undirected draws),
    oriented by (degree, id): the tri relation listed and counted."""
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
    top = max((r for r in recs if r["bytes"] > 0), key=lambda r: r["ms"], default=None)
    if args.kernels:
        log(json.dumps(sorted(recs, key=lambda r: -r["ms"]), indent=1))
    count = int(ntri[0][0][0].value) if ntri else None
    out = {"config": "wcoj_triangles_rmat", "params": w.params, "tri": n, "ntri": count,
           "tri_equals_ntri": count == n, "ms_per_step": round(ms, 3), "tuples_per_s": round(n / (ms / 1e3), 1),
           "kernel_ms_per_step": {r["name"]: round(r["ms"] / args.steps, 2)
                                  for r in sorted(recs, key=lambda r: -r["ms"])[:8]},
           "roofline": roofline_of(top)}
    run.free()
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_cfg4(args, emit=True):
    """Batched: 512 independent programs (256 Bool TC, 256 Trop cost) through
    one batch, against 512 sequential plain runs of the same programs.  value
    = the execute-phase speedup (every fixpoint over EDB already resident in
    HBM, wall clock); both sides are also timed end to end (host EDB in, host
    results out).  Every batched result is compared with its plain run."""
    import strata_k as sk
    from strata_k import workloads as wl
    t = time.time()
    progs = wl.batch_programs(seed=args.seed if args.seed != 2 else 4, count=args.programs, cost_sources=1,
                              tc_node_cap=2000, tc_edge_cap=10_000)
    log(f"generated {len(progs)} programs in {time.time() - t:.1f}s")
    ctx = sk.Context(0)
    # plain: one prepared run per program
    plain, plain_exec_ms, plain_read_s = [], 0.0, 0.0
    t0 = time.perf_counter()
    for p in progs:
        run = ctx.prepare(p.program)
        for pred, (cols, w) in p.bulk.items():
            run.add_sym(pred, cols, w, None)
        run.load()
        run.execute()
        plain_exec_ms += run.execute_ms
        tr = time.perf_counter()
        plain.append((run.relation_sym(p.output), run.iterations()))
        plain_read_s += time.perf_counter() - tr
        run.free()
    ctx.synchronize()
    plain_s = time.perf_counter() - t0
    # batched
    t0 = time.perf_counter()
    b = sk.Batch(ctx)
    for p in progs:
        b.add(p.program, p.bulk)
    t_add = time.perf_counter()
    b.load()
    t_load = time.perf_counter()
    b.execute()
    t_exec = time.perf_counter()
    batched = [(b.relation_sym(i, p.output), b.iterations(i)) for i, p in enumerate(progs)]
    ctx.synchronize()
    t_read = time.perf_counter()
    batch_s = t_read - t0
    phases = {"add_s": round(t_add - t0, 3), "load_s": round(t_load - t_add, 3), "execute_s": round(t_exec - t_load, 3),
              "read_s": round(t_read - t_exec, 3)}
    batch_exec_ms = b.execute_ms
    b.free()
    exact = True
    n_out = 0
    for (pa, pi), (ba, bi) in zip(plain, batched):
        (pc, pw, _), (bc, bw) = pa, ba
        n_out += len(pc[0])
        same = pi == bi and len(pc[0]) == len(bc[0]) and all((x == y).all() for x, y in zip(pc, bc))
        if pw is not None:
            same = same and (pw == bw).all()
        exact = exact and same
    ctx.close()
    out = {
        "metric": "batched-vs-plain speedup",
        "value": round(plain_exec_ms / batch_exec_ms, 2) if batch_exec_ms else None, "unit": "x",
        "value_is": "execute phase (fixpoints on resident EDB, wall clock): batched vs 512 plain runs",
        "end_to_end_speedup": round(plain_s / batch_s, 2),
        "end_to_end_speedup_results_in_hbm": round((plain_s - plain_read_s) / (batch_s - (t_read - t_exec)), 2),
        "end_to_end_note": "host EDB in; end_to_end_speedup includes copying every result to host arrays "
                           "(512 reads each side), the results_in_hbm figure stops with the results in HBM",
        "n_gpus": 1, "higher_is_better": True, "dtype": "int64", "data": "synthetic (seeded ER / RMAT)",
        "config": {"workload": f"{len(progs)} programs: {len(progs) // 2} Bool TC (n <= 2000, m <= 10000) + "
                               f"{len(progs) - len(progs) // 2} Trop cost (n in [1e3, 2e5], m in [4e3, 2e6], 1 source)",
                   "derived_tuples": n_out},
        "plain_s": round(plain_s, 3), "batched_s": round(batch_s, 3), "plain_read_s": round(plain_read_s, 3),
        "plain_tuples_per_s": round(n_out / plain_s, 1), "batched_tuples_per_s": round(n_out / batch_s, 1),
        "plain_device_ms": round(plain_exec_ms, 1), "batched_device_ms": round(batch_exec_ms, 1),
        "batched_phases": phases,
        "bit_exact_vs_plain": bool(exact)}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_cfg5(args, emit=True):
    """Served: resident RMAT graph, Poisson query stream (arrival times from
    the workload generator), continuous batching; p50/p99 latency and
    queries/s.  Every query's answer is taken source by source as its
    instances finish (sk_server_poll_source / take_source): the answers of
    the checked queries are copied to the host, the others are released on
    the device once compacted (a 4096-source answer is ~2e10 rows, more than
    host memory).  Capacity: the same queries offered at once, batched
    versus one query at a time on the same resident graph.  The checked
    queries (the first ones with at most 64 sources) are also run as plain
    programs (EDB loaded per run): rows, weights and round counts must match."""
    import strata_k as sk
    from strata_k import workloads as wl
    t = time.time()
    scale = args.scale if args.scale != 20 else 24
    s, d, w, n = wl.config5_graph(seed=5, scale=scale)
    log(f"graph scale {scale}: {len(s)} edges in {time.time() - t:.1f}s")
    qs = wl.config5_queries(seed=5, n_queries=args.queries, n_nodes=n, max_sources=args.max_sources,
                            n_targets=64, filter_frac=0.1, rate_per_s=args.rate)
    ctx = sk.Context(0)
    small = [i for i, q in enumerate(qs) if len(q["sources"]) <= 64]
    n_plain = min(len(small), args.plain_queries)
    check = set(small[:min(n_plain, getattr(args, "check_queries", 4))])
    checked = None
    stats = {}

    def serve(max_batch, open_loop, keep=()):
        srv = sk.Server(ctx, s, d, w, max_batch=max_batch)
        pending = list(range(len(qs)))
        tickets, done, answers, parts = {}, {}, {}, {}
        t0 = time.perf_counter()
        last = t0
        while len(done) < len(qs):
            now = time.perf_counter() - t0
            while pending and (not open_loop or qs[pending[0]]["arrival_s"] <= now):
                i = pending.pop(0)
                tickets[i] = srv.submit(qs[i]["sources"], qs[i]["targets"])
                parts[i] = []
            active, waiting = srv.step()
            for i, tk in list(tickets.items()):
                if i in done:
                    continue
                while True:
                    p = srv.poll_source(tk)
                    if p is None:
                        break
                    if p["index"] == len(qs[i]["sources"]):
                        done[i] = srv.poll(tk)
                        if i in keep:
                            answers[i] = parts[i]
                        break
                    if i in keep:
                        z, ww = srv.take_source(tk, p["rows"])
                        parts[i].append((p["source"], z, ww))
                    else:
                        srv.take_source(tk, p["rows"], discard=True)
            if active == 0 and waiting == 0 and pending and open_loop:
                time.sleep(max(0.0, min(0.002, qs[pending[0]]["arrival_s"] - (time.perf_counter() - t0))))
            if time.perf_counter() - last > 30:
                last = time.perf_counter()
                log(f"  served {len(done)}/{len(qs)} queries, {time.perf_counter() - t0:.0f}s")
        total = time.perf_counter() - t0
        stats[max_batch if not open_loop else "open"] = srv.stats()
        log(f"  server counters: {srv.stats()}")
        srv.free()
        return total, done, answers

    t_open, res, answers = serve(args.max_batch, True, keep=check)
    lat = np.array([res[i]["latency_ms"] for i in range(len(qs))])
    rows = sum(res[i]["rows"] for i in range(len(qs)))
    n_src = sum(len(q["sources"]) for q in qs)
    log(f"open loop: {t_open:.1f}s, {rows} answer rows over {n_src} sources")
    t_cap, res_b, _ = serve(args.max_batch, False)
    log(f"batched capacity: {t_cap:.1f}s")
    t_seq, res1, _ = serve(1, False)
    log(f"one at a time: {t_seq:.1f}s")
    exact = all(res1[i]["rows"] == res[i]["rows"] == res_b[i]["rows"] and
                res1[i]["iterations"] == res[i]["iterations"] == res_b[i]["iterations"] for i in range(len(qs)))
    # plain runs (EDB loaded per run) of the first queries with <= 64 sources
    t0 = time.perf_counter()
    plain_src = 0
    for i in small[:n_plain]:
        q = qs[i]
        plain_src += len(q["sources"])
        filt = len(q["targets"]) > 0
        bulk = {"edge": ([s, d], w, None), "src": ([q["sources"]], np.zeros(len(q["sources"]), np.int64), None)}
        if filt:
            bulk["tgt"] = ([q["targets"]], np.zeros(len(q["targets"]), np.int64), None)
        out_rel = "ans" if filt else "cost"
        db = ctx.run_semi_naive(wl.query_program(filt), bulk=bulk, want=[out_rel])
        exact = exact and len(db.relation(out_rel)) == res[i]["rows"] and db.iterations[0] == res[i]["iterations"]
        if i in answers:
            # the served answer's rows and weights against the plain run's
            # (the engine's dense cells and pull rounds, not the server's)
            rel = db.relation(out_rel)
            ps, pd_, pw = rel.vals[:, 0], rel.vals[:, 1], rel.w
            got = answers[i]
            gs = np.concatenate([np.full(len(z), src, np.uint32) for src, z, _ in got]) if got else np.empty(0)
            gd = np.concatenate([z for _, z, _ in got]) if got else np.empty(0)
            gw = np.concatenate([ww for _, _, ww in got]) if got else np.empty(0)
            same = len(gs) == len(ps) and bool((gs == ps).all() and (gd == pd_).all() and (gw == pw).all())
            checked = same if checked is None else (checked and same)
    t_plain = time.perf_counter() - t0
    ctx.close()
    st = stats["open"]
    out = {
        "metric": "p99 query ms", "value": round(float(np.percentile(lat, 99)), 2), "unit": "ms",
        "n_gpus": 1, "higher_is_better": False, "dtype": "int64", "data": "synthetic (seeded RMAT, Poisson arrivals)",
        "config": {"workload": f"served cheapest cost, RMAT scale {scale} ({n} nodes, {len(s)} edges), "
                               f"{len(qs)} queries, sources log-uniform [1, {args.max_sources}], 10% filtered",
                   "rate_per_s": args.rate, "max_batch": args.max_batch, "answer_rows": int(rows),
                   "sources": int(n_src), "largest_source_set": int(max(len(q["sources"]) for q in qs)),
                   "lanes": st["lanes"]},
        "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "served_qps": round(len(qs) / t_open, 2),
        "capacity_qps_batched": round(len(qs) / t_cap, 2), "capacity_qps_one_at_a_time": round(len(qs) / t_seq, 2),
        "batched_vs_one_at_a_time": round(t_seq / t_cap, 2),
        "capacity_sources_per_s_batched": round(n_src / t_cap, 1),
        "plain_run_sources_per_s": round(plain_src / t_plain, 1) if n_plain else None,
        "plain_queries_timed": n_plain,
        "served_vs_plain_runs": round((n_src / t_cap) / (plain_src / t_plain), 2) if n_plain else None,
        "server_rounds": {k: st[k] for k in ("steps", "pull_rounds", "push_joins", "pull_pairs", "push_pairs")},
        "same_rows_and_rounds_alone_batched_and_plain": bool(exact),
        "answers_compared_with_plain_runs": len(answers) if checked is not None else 0,
        "answer_contents_equal": checked,
        "answers": "taken per source as instances finish; checked queries copied to the host, the others "
                   "released in HBM after compaction"}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_sort(args, emit=True):
    """Device-resident LSD radix sort (the engine's sort primitive) of 2^26
    packed two-column keys (44 significant bits: 22-bit ids), keys alone and
    with i64 weights, under the context's autotuner.  Algorithmic bytes of a
    sort of n records of r bytes with k-byte keys in P passes:
    n*k + P*2*n*r (one histogram read of the keys, a read and a write of
    every record per pass), whichever candidate runs."""
    import torch
    import strata_k as sk
    n = 1 << args.sort_log2
    ctx = sk.Context(0)
    ctx.set_autotune(2)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(7)
    out = []
    for weighted in (False, True):
        keys = torch.randint(0, 1 << 44, (n,), dtype=torch.int64, device="cuda:0", generator=g)
        alt = torch.empty_like(keys)
        w = torch.randint(1, 1000, (n,), dtype=torch.int64, device="cuda:0", generator=g) if weighted else None
        walt = torch.empty_like(w) if weighted else None
        src = keys.clone()
        torch.cuda.synchronize()
        choice = ctx.radix_choice(n, 44, weighted=weighted)
        passes = -(-44 // choice["digit_bits"])
        for _ in range(args.warmup):
            keys.copy_(src)
            torch.cuda.synchronize()
            ctx.radix_sort_device(keys.data_ptr(), alt.data_ptr(), n, 44, w.data_ptr() if weighted else 0,
                                  walt.data_ptr() if weighted else 0)
            ctx.synchronize()
        ms = []
        for _ in range(args.steps):
            keys.copy_(src)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            in_alt = ctx.radix_sort_device(keys.data_ptr(), alt.data_ptr(), n, 44, w.data_ptr() if weighted else 0,
                                           walt.data_ptr() if weighted else 0)
            ctx.synchronize()
            ms.append((time.perf_counter() - t0) * 1e3)
        res = alt if in_alt else keys
        ok = bool(torch.equal(res, torch.sort(src).values))
        t = float(np.median(ms))
        # every candidate pinned in turn, the same keys (stderr)
        per_cand = {}
        for c in range(sk.radix_candidates()):  # every tuner candidate (sk_ctx_force_radix)
            ctx.force_radix(c)
            cm = []
            for _ in range(3):
                keys.copy_(src)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ctx.radix_sort_device(keys.data_ptr(), alt.data_ptr(), n, 44, w.data_ptr() if weighted else 0,
                                      walt.data_ptr() if weighted else 0)
                ctx.synchronize()
                cm.append((time.perf_counter() - t1) * 1e3)
            per_cand[c] = round(float(np.median(cm)), 3)
        ctx.force_radix(-1)
        log(f"weighted={weighted}: ms per sort by candidate {per_cand}")
        rec = 8 + (8 if weighted else 0)
        alg_bytes = n * 8 + passes * 2 * rec * n
        gbs = alg_bytes / (t * 1e-3) / 1e9
        out.append({"weighted": weighted, "n": n, "key_bits": 44, "candidate": choice["candidate"],
                    "digit_bits": choice["digit_bits"], "items_per_thread": choice["items_per_thread"],
                    "block": choice["block"], "single_sweep": choice["sweep"], "passes": passes,
                    "ms_per_sort": round(t, 3),
                    "keys_per_s": round(n / (t * 1e-3), 1), "alg_gbps": round(gbs, 1),
                    "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBPS, 4), "sorted_equals_torch": ok,
                    "ms_per_sort_by_candidate": per_cand})
        del keys, alt, w, walt, src
    ctx.close()
    line = {"metric": "radix sort keys/s", "unit": "keys/s", "results": out}
    if emit:
        print(json.dumps(line), flush=True)
    return line


# The reference's published chain-union figures (crates/strata-gpu/README.md:56-58,
# RTX-4090-class card, CUDA 11.8; examples/bench.rs:22-90 times the whole host
# call): (semiring, edges, chain length, seconds).
CHAIN_UNION_REF = [("bool", 10_000_000, 5, 1.08), ("trop", 10_000_000, 5, 2.07),
                   ("bool", 100_000_000, 2, 3.2), ("trop", 100_000_000, 2, 7.9),
                   ("bool", 100_000_000, 5, 9.1)]


def bench_chain_union(args):
    """The reference's own benchmark workload: K disjoint chains of C edges
    (chain k on nodes k(C+1) .. k(C+1)+C), closure K·C(C+1)/2, through the
    crate functions transitive_closure_bool / shortest_paths_trop (unit
    weights: distance = d - s), timed as the reference times them: the whole
    host call, host edges in and host result out (H2D, fixpoint, D2H).
    Checked: the closure size, and for Trop that every distance is d - s."""
    import ctypes as C
    import strata_k as sk
    from strata_k.engine import _take
    ctx = sk.Context(0)
    L = ctx.L
    out = []
    for sem, m, c, ref_s in CHAIN_UNION_REF:
        k = m // c
        base = np.arange(k, dtype=np.int64) * (c + 1)
        src = (base[:, None] + np.arange(c)[None, :]).reshape(-1).astype(np.uint32)
        dst = src + np.uint32(1)
        w = np.ones(len(src), dtype=np.int64)
        expected = k * c * (c + 1) // 2
        times = []
        for rep in range(2):  # the first call also pays one-time set-up (autotuner classes, pool growth)
            pa, pb, pw, n = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_uint64()
            t0 = time.perf_counter()
            if sem == "bool":
                code = L.sk_transitive_closure_bool(ctx.h, src.ctypes.data, dst.ctypes.data, len(src), C.byref(pa),
                                                    C.byref(pb), C.byref(n))
            else:
                code = L.sk_shortest_paths_trop(ctx.h, src.ctypes.data, dst.ctypes.data, w.ctypes.data, len(src),
                                                C.byref(pa), C.byref(pb), C.byref(pw), C.byref(n))
            times.append(time.perf_counter() - t0)
            if code:
                raise RuntimeError(f"chain union {sem} {m}: status {code}")
            a = _take(L, pa, n.value, np.uint32)
            b = _take(L, pb, n.value, np.uint32)
            ok = n.value == expected
            if sem == "trop":
                ww = _take(L, pw, n.value, np.int64)
                ok = ok and bool(np.array_equal(ww, b.astype(np.int64) - a.astype(np.int64)))
            del a, b
        t = min(times)
        log(f"chain union {sem} {m} edges, C={c}: {n.value} pairs, calls {['%.3f' % x for x in times]} s")
        out.append({"semiring": sem, "edges": m, "chain_len": c, "closure": expected, "exact": bool(ok),
                    "seconds": round(t, 4), "closure_tuples_per_s": round(expected / t, 1),
                    "first_call_s": round(times[0], 4), "reference_seconds": ref_s,
                    "reference_tuples_per_s": round(expected / ref_s, 1), "speedup_vs_reference": round(ref_s / t, 2)})
    ctx.close()
    print(json.dumps({"metric": "chain-union closure tuples/s (the reference's bench workload, whole host call)",
                      "unit": "tuples/s", "reference_hardware": "RTX-4090-class, CUDA 11.8 (published)",
                      "results": out}), flush=True)


def bench_tri_hubs(args):
    """The reference's skewed triangle example (crates/strata-gpu/examples/
    tri.rs, defaults: 16 hubs, 100,000 bridges, 500,000 random neighbours per
    hub below 10 M ids: 9.6 M edges, 12,000,560 triangles) through
    count_triangles, timed as the example times it (the host call); the
    reference publishes ~0.54 s (crates/strata-gpu/README.md:73-77)."""
    import ctypes as C
    import strata_k as sk
    h, b, d, n = 16, 100_000, 500_000, 10_000_000
    ctx = sk.Context(0)
    L = ctx.L
    cap = h * (h - 1) // 2 + h * b + h * d
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    m = int(L.sk_wl_tri_hubs(h, b, d, n, src.ctypes.data, dst.ctypes.data))
    expected = b * (h * (h - 1) // 2) + h * (h - 1) * (h - 2) // 6
    times, got = [], None
    for _ in range(3):  # the first call also pays one-time set-up
        cnt = C.c_uint64()
        t0 = time.perf_counter()
        code = L.sk_count_triangles(ctx.h, src.ctypes.data, dst.ctypes.data, m, C.byref(cnt))
        times.append(time.perf_counter() - t0)
        if code:
            raise RuntimeError(f"count_triangles: status {code}")
        got = cnt.value
    t = min(times)
    ctx.close()
    print(json.dumps({"metric": "triangle count, the reference's skewed hub graph (whole host call)", "unit": "s",
                      "edges": m, "triangles": got, "exact": got == expected, "seconds": round(t, 4),
                      "calls_s": [round(x, 4) for x in times], "reference_seconds": 0.54,
                      "speedup_vs_reference": round(0.54 / t, 2),
                      "reference_hardware": "RTX-4090-class, CUDA 11.8 (published)"}), flush=True)


def asp_program(rng, n_rules, n_atoms, max_body):
    """A random ground program with the shape of the reference's large
    simplify-filter case (crates/strata-gpu/tests/asp_diff.rs:24-58,72-77):
    1/8 constraints, bodies of 0..max_body signed literals, ~75% of atoms
    possible, a third of those certain (drawn with numpy, not its LCG)."""
    head = rng.integers(0, n_atoms, n_rules, dtype=np.int32)
    head[rng.integers(0, 8, n_rules) == 0] = -1
    blen = rng.integers(0, max_body + 1, n_rules)
    bs = np.zeros(n_rules + 1, np.uint32)
    np.cumsum(blen, out=bs[1:])
    n_lits = int(bs[-1])
    lits = (rng.integers(0, n_atoms, n_lits, dtype=np.int32) + 1) * np.where(rng.integers(0, 2, n_lits) == 1, 1, -1)
    poss = (rng.integers(0, 4, n_atoms) != 0).astype(np.uint8)
    cert = (poss.astype(bool) & (rng.integers(0, 3, n_atoms) == 0)).astype(np.uint8)
    return head, bs, lits.astype(np.int32), poss, cert


def bench_asp(args):
    """The ASP grounding simplify filter over a device-resident ground program
    of 4M rules on 4096 atoms (bodies up to 6 literals); value = rules
    filtered per second.  The dominant kernel's roofline uses its algorithmic
    bytes (per rule 9: end offset read, alive flag and kept count written;
    per literal 5: literal and atom state read)."""
    import torch
    import strata_k as sk
    from oracle import oracle as O
    n_rules, n_atoms, max_body = args.asp_rules, 4096, 6
    rng = np.random.default_rng(args.seed)
    head, bs, lits, poss, cert = asp_program(rng, n_rules, n_atoms, max_body)
    n_lits = int(bs[-1])
    ctx = sk.Context(0)
    dev = "cuda:0"
    t = {k: torch.from_numpy(a).to(dev) for k, a in
         (("head", head), ("bs", bs.view(np.int32)), ("lits", lits), ("poss", poss), ("cert", cert))}
    oh = torch.empty(n_rules, dtype=torch.int32, device=dev)
    ob = torch.empty(n_rules + 1, dtype=torch.int32, device=dev)
    ol = torch.empty(max(1, n_lits), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def call():
        return ctx.asp_simplify_filter_device(t["head"].data_ptr(), t["bs"].data_ptr(), t["lits"].data_ptr(), n_rules,
                                              n_lits, t["poss"].data_ptr(), t["cert"].data_ptr(), n_atoms,
                                              oh.data_ptr(), ob.data_ptr(), ol.data_ptr())
    for _ in range(args.warmup):
        call()
    ctx.synchronize()
    ctx.kernel_reset()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nr, nl = call()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(1, args.steps)
    ctx.set_profiling(False)
    recs = ctx.kernel_records()
    top = max((r for r in recs if r["bytes"] > 0), key=lambda r: r["ms"], default=None)
    # exactness on a sample of the same program: the first 200K rules
    k = min(n_rules, 200_000)
    sample = (head[:k], bs[:k + 1], lits[:int(bs[k])], poss, cert)
    got = sk.asp_simplify_filter(*sample, ctx=ctx)
    want = O.asp_simplify_filter(*sample)
    exact = got[0].tolist() == want[0] and got[1].tolist() == want[1] and got[2].tolist() == want[2]
    # PCIe-inclusive rate of the host-buffer entry point (stderr only)
    t1 = time.perf_counter()
    sk.asp_simplify_filter(head, bs, lits, poss, cert, ctx=ctx)
    log(f"host-buffer entry point (copies in and out included): {n_rules / (time.perf_counter() - t1) / 1e6:.1f}M rules/s")
    if args.kernels:
        for r in sorted(recs, key=lambda r: -r["ms"]):
            log(f"{r['name']:<24} launches {r['launches']:>4}  ms {r['ms']:.3f}  GB {r['bytes'] / 1e9:.3f}")
    cpu = None
    if not args.no_cpu:
        m = min(n_rules, 4_000_000)
        sample = (head[:m], bs[:m + 1], lits[:int(bs[m])], poss, cert)
        t2 = time.perf_counter()
        O.asp_simplify_filter(*sample)
        cs = time.perf_counter() - t2
        cpu = {"value": round(m / cs, 1), "unit": "rules/s", "cores": 1, "kind": "port",
               "sample": f"the first {m} rules of the same program through the oracle's restatement of "
                         f"simplify_filter_cpu: {cs:.1f} s",
               "host_cpu": platform.processor() or platform.machine()}
    line = {"metric": "ASP simplify-filter rules/s", "value": round(n_rules / (ms * 1e-3), 1), "unit": "rules/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i32",
            "data": "synthetic", "exact_vs_oracle_sample": exact,
            "config": {"workload": "ASP simplify filter, device-resident ground program", "rules": n_rules,
                       "atoms": n_atoms, "max_body": max_body, "literals": n_lits, "kept_rules": nr,
                       "kept_literals": nl},
            "roofline": roofline_of(top, measured_traffic(top["name"]) if top else None), "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="config 2: leave out the batched (config 4), served (config 5), config 1, config 3 and sort "
                         "measurements")
    ap.add_argument("--kernels", action="store_true")
    ap.add_argument("--programs", type=int, default=512, help="config 4: number of programs")
    ap.add_argument("--queries", type=int, default=64, help="config 5: number of queries")
    ap.add_argument("--max-sources", type=int, default=16, help="config 5: largest source set")
    ap.add_argument("--rate", type=float, default=20.0, help="config 5: arrivals per second")
    ap.add_argument("--max-batch", type=int, default=16, help="config 5: queries in flight")
    ap.add_argument("--plain-queries", type=int, default=8, help="config 5: queries timed as plain runs")
    ap.add_argument("--graph", type=int, default=0,
                    help="config 2: device-driven rounds captured in a hipGraph (1/0).  Off by default so that "
                         "every launch of the dominant kernel is timed by its own HIP events; on, the pull "
                         "rounds inside graph launches are timed as one record per launch")
    ap.add_argument("--pack", type=int, default=1,
                    help="config 2: packed pull deltas (1, the default) or the unpacked weights and lanes (0)")
    ap.add_argument("--sort", action="store_true", help="time the device radix sort primitive")
    ap.add_argument("--sort-log2", type=int, default=26)
    ap.add_argument("--chain-union", action="store_true",
                    help="the reference's chain-union benchmark through the crate functions")
    ap.add_argument("--tri-hubs", action="store_true",
                    help="the reference's skewed triangle example through count_triangles")
    ap.add_argument("--asp", action="store_true", help="time the ASP grounding simplify filter")
    ap.add_argument("--asp-rules", type=int, default=4_000_000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one known-size copy kernel (PMC counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (per-GPU shards would be "
              "independent source sets); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.sort or args.copy_probe or (args.config == 2 and not args.no_extras):
        # torch (device buffers and the reference sort of the sort leg) brings
        # its own HIP runtime, which finds the device only when it starts
        # before the engine's: start it first
        import torch
        torch.zeros(1, device="cuda:0")
        torch.cuda.synchronize()
    if args.copy_probe:
        copy_probe(args.copy_probe)
    if not args.sort:
        # the radix autotuner times every size class once per device and
        # shares the timings with every later context of the process: do it
        # here, before any timed region
        import strata_k as sk
        t = time.perf_counter()
        tctx = sk.Context(0)
        tctx.set_autotune(2)
        tctx.close()
        log(f"radix autotuner: every class timed in {time.perf_counter() - t:.2f}s")
    if args.sort:
        bench_sort(args)
        return
    if args.asp:
        bench_asp(args)
        return
    if args.chain_union:
        bench_chain_union(args)
        return
    if args.tri_hubs:
        bench_tri_hubs(args)
        return
    {1: bench_cfg1, 2: bench_cfg2, 3: bench_cfg3, 4: bench_cfg4, 5: bench_cfg5}[args.config](args)


if __name__ == "__main__":
    main()
