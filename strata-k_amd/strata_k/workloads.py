"""Seeded synthetic workloads of the five benchmark configurations.

All randomness comes from the library's counter-based generators
(sk_wl_* in csrc/sk_workload.cpp), so the CPU oracle and the device engine
read identical EDB bytes for a given seed and size.  Every function takes the
full-size parameters as defaults and smaller sizes for parity tests.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .core_ir import BOOL, TROP, cost_program, tc_program, triangle_program, CoreProgram
from .engine import load_library

RMAT_ABC = (0.57, 0.19, 0.19)


def _L():
    return load_library()


def uniform_graph(seed: int, n_nodes: int, n_edges: int):
    """Distinct directed edges drawn uniformly, no self-loops (acceptance order)."""
    s = np.empty(n_edges, np.uint32)
    d = np.empty(n_edges, np.uint32)
    m = _L().sk_wl_uniform_graph(seed, n_nodes, n_edges, s.ctypes.data, d.ctypes.data)
    return s[:m], d[:m]


def rmat_graph(seed: int, scale: int, n_draws: int, undirected: bool = False, n_limit: int = 0):
    """RMAT draws (a,b,c,d) = (.57,.19,.19,.05), self-loops and duplicates removed, sorted."""
    s = np.empty(n_draws, np.uint32)
    d = np.empty(n_draws, np.uint32)
    L = _L()
    m = L.sk_wl_rmat(seed, scale, n_draws, RMAT_ABC[0], RMAT_ABC[1], RMAT_ABC[2], n_limit, s.ctypes.data,
                     d.ctypes.data)
    m = L.sk_wl_normalize(s.ctypes.data, d.ctypes.data, m, 1, 1 if undirected else 0)
    return s[:m].copy(), d[:m].copy()


def weights(seed: int, n: int, lo: int = 1, hi: int = 1000) -> np.ndarray:
    w = np.empty(n, np.int64)
    _L().sk_wl_weights(seed, n, lo, hi, w.ctypes.data)
    return w


def sample_distinct(seed: int, n_nodes: int, k: int) -> np.ndarray:
    out = np.empty(k, np.uint32)
    m = _L().sk_wl_sample_distinct(seed, n_nodes, k, out.ctypes.data)
    return out[:m]


def orient_by_degree(a: np.ndarray, b: np.ndarray, n_nodes: int):
    a = np.ascontiguousarray(a, np.uint32).copy()
    b = np.ascontiguousarray(b, np.uint32).copy()
    _L().sk_wl_orient_by_degree(a.ctypes.data, b.ctypes.data, len(a), n_nodes)
    return a, b


def log_uniform(seed: int, i: int, lo: int, hi: int) -> int:
    return int(_L().sk_wl_log_uniform(seed, i, lo, hi))


def uniform01(seed: int, i: int) -> float:
    return float(_L().sk_wl_uniform01(seed, i))


@dataclass
class Workload:
    name: str
    program: CoreProgram
    bulk: dict                      # pred -> (columns, weights, inf)
    outputs: List[str]              # derived (IDB) relations
    params: dict = field(default_factory=dict)


# ------------------------------------------------------------------ config 1

def config1(seed: int = 1, n_nodes: int = 2000, n_edges: int = 10_000) -> Workload:
    """Plain Bool transitive closure on a uniform graph (2,000 nodes, 10,000 edges)."""
    s, d = uniform_graph(seed, n_nodes, n_edges)
    return Workload("bool_tc_uniform", tc_program(BOOL), {"edge": ([s, d], None, None)}, ["path"],
                    dict(n_nodes=n_nodes, n_edges=int(len(s)), seed=seed))


# ------------------------------------------------------------------ config 2

def config2(seed: int = 2, scale: int = 20, edge_factor: int = 16, n_sources: int = 64) -> Workload:
    """Plain Trop multi-source cheapest cost on RMAT (scale 20, 16M draws, 64 sources)."""
    n = 1 << scale
    s, d = rmat_graph(seed, scale, edge_factor * n)
    w = weights(seed, len(s))
    src = np.sort(sample_distinct(seed, n, n_sources))
    return Workload("trop_cost_rmat", cost_program(),
                    {"edge": ([s, d], w, None), "src": ([src], np.zeros(len(src), np.int64), None)}, ["cost"],
                    dict(scale=scale, n_nodes=n, n_edges=int(len(s)), n_sources=int(len(src)), seed=seed))


def restrict_sources(wl: Workload, k: int) -> Workload:
    """The same workload with only the first k sources (cost(S,.) partitions by S)."""
    src = wl.bulk["src"][0][0][:k]
    bulk = dict(wl.bulk)
    bulk["src"] = ([src], np.zeros(len(src), np.int64), None)
    p = dict(wl.params)
    p["n_sources"] = int(len(src))
    return Workload(wl.name + f"_src{k}", wl.program, bulk, wl.outputs, p)


# ------------------------------------------------------------------ config 3

def config3(seed: int = 3, scale: int = 22, edge_factor: int = 16) -> Workload:
    """Bool WCOJ triangles on RMAT (scale 22, 64M undirected draws), (degree, id)-oriented."""
    n = 1 << scale
    a, b = rmat_graph(seed, scale, edge_factor * n, undirected=True)
    a, b = orient_by_degree(a, b, n)
    return Workload("wcoj_triangles_rmat", triangle_program(), {"e": ([a, b], None, None)}, ["tri", "ntri"],
                    dict(scale=scale, n_nodes=n, n_edges=int(len(a)), seed=seed))


# ------------------------------------------------------------------ config 4

@dataclass
class BatchProgram:
    kind: str          # "tc" or "cost"
    gen: str           # "er" or "rmat"
    n_nodes: int
    n_edges: int
    seed: int
    program: CoreProgram
    bulk: dict
    output: str


def batch_programs(seed: int = 4, count: int = 512, n_lo: int = 1000, n_hi: int = 200_000, m_lo: int = 4000,
                   m_hi: int = 2_000_000, cost_sources: int = 1, tc_node_cap: Optional[int] = None,
                   tc_edge_cap: Optional[int] = None) -> List[BatchProgram]:
    """Independent programs: half Bool TC, half Trop cheapest cost, each on its
    own seeded ER or RMAT graph with log-uniform node and edge counts.

    ``tc_node_cap`` / ``tc_edge_cap`` / ``cost_sources`` bound the output
    volume (the Bool closures of the uncapped configuration total ~10^11
    tuples)."""
    out = []
    for i in range(count):
        ps = seed * 1_000_003 + i
        kind = "tc" if i % 2 == 0 else "cost"
        gen = "er" if (uniform01(ps, 0) < 0.5) else "rmat"
        n = log_uniform(ps, 1, n_lo, n_hi)
        if kind == "tc" and tc_node_cap:
            n = min(n, tc_node_cap)
        m = log_uniform(ps, 2, m_lo, m_hi)
        if kind == "tc" and tc_edge_cap:
            m = min(m, tc_edge_cap)
        m = min(m, n * (n - 1) // 2)
        if gen == "er":
            s, d = uniform_graph(ps, n, m)
        else:
            scale = max(1, int(np.ceil(np.log2(n))))
            s, d = rmat_graph(ps, scale, m, n_limit=n)
        if kind == "tc":
            prog = tc_program(BOOL)
            bulk = {"edge": ([s, d], None)}
            output = "path"
        else:
            w = weights(ps, len(s))
            src = np.sort(sample_distinct(ps, n, cost_sources))
            prog = cost_program()
            bulk = {"edge": ([s, d], w), "src": ([src], np.zeros(len(src), np.int64))}
            output = "cost"
        out.append(BatchProgram(kind, gen, n, int(len(s)), ps, prog, bulk, output))
    return out


# ------------------------------------------------------------------ config 5

def config5_graph(seed: int = 5, scale: int = 24, edge_factor: int = 16):
    """Resident EDB: RMAT (scale 24, 256M draws), Trop weights uniform [1, 1000]."""
    n = 1 << scale
    s, d = rmat_graph(seed, scale, edge_factor * n)
    w = weights(seed, len(s))
    return s, d, w, n


def config5_queries(seed: int, n_queries: int, n_nodes: int, max_sources: int = 4096, n_targets: int = 64,
                    filter_frac: float = 0.1, rate_per_s: float = 50.0):
    """Poisson arrivals at `rate_per_s`; source sets log-uniform in [1, max_sources];
    a `filter_frac` share adds a 64-node target filter."""
    qs = []
    t = 0.0
    for i in range(n_queries):
        u = uniform01(seed, 3 * i)
        t += -np.log(max(1e-12, 1.0 - u)) / rate_per_s
        k = log_uniform(seed, 3 * i + 1, 1, max_sources)
        src = np.sort(sample_distinct(seed * 7919 + i, n_nodes, k))
        tg = np.empty(0, np.uint32)
        if uniform01(seed, 3 * i + 2) < filter_frac:
            tg = np.sort(sample_distinct(seed * 104729 + i, n_nodes, n_targets))
        qs.append(dict(arrival_s=t, sources=src, targets=tg))
    return qs


def query_program(filtered: bool) -> CoreProgram:
    """The served query: cost closure from src, optionally ans(S,Z) :- cost(S,Z), tgt(Z)."""
    from .core_ir import CoreProgram, atom, pos, pred, rule, v
    p = cost_program()
    if not filtered:
        return p
    preds = list(p.predicates) + [pred("tgt", 1, TROP, 0), pred("ans", 2, TROP, 1)]
    rules = list(p.rules) + [rule(atom("ans", v(0), v(1)), [pos("cost", v(0), v(1)), pos("tgt", v(1))], 2, 1)]
    return CoreProgram(preds, rules, 2)
