"""CPU: pins for the oracle functions the reference's vectors leave open.

The reference's own vectors cover count and sum (interp.rs:327-402); the
checks here pin the rest of the aggregate folds, every EvalError status the
oracle can return, and the oriented triangle counter, each against something
computed another way:
  * aggregates: a brute-force enumeration of the body's solutions in Python
    (nested loops over the facts), grouped with a dict and folded with
    Python's len/sum/min/max (naive.rs:300-403);
  * errors: hand-written programs (tests/helpers.py ERROR_CASES) whose
    outcome is read off the reference's raise points (naive.rs:26-54 and
    their call sites), each "raised when reached" case beside a twin where
    the failing literal or head is never reached, so no error is raised;
  * orc_tri_count_oriented: brute-force triples, C(n,3) on complete graphs
    and the hub/bridge closed form b*C(h,2)+C(h,3) (crates/strata-gpu/examples/tri.rs:72),
    under id and (degree, id) orientations (wcoj.rs:27-43, wcoj_diff.rs:56-106).
"""
import math
import random
from collections import defaultdict

import numpy as np
import pytest

from helpers import ERROR_CASES, brute_triangles, orient_simple
from oracle import oracle as O
from strata_k import workloads as wl
from strata_k.core_ir import (BOOL, Agg, CoreAtom, CoreProgram, CoreRule, IntV, Pos, Sym, atom, neg, pos, pred,
                              rule, v)

I64 = 2**63


def wrap64(x: int) -> int:
    """i64 addition as the reference's release build does it (Rust's Sum over
    i64 adds without overflow checks in the release profile, which the
    workspace's Cargo.toml leaves at its default)."""
    return (x + I64) % (2 * I64) - I64


def fold(op, vals):
    if op == "count":
        return len(vals)
    if op == "sum":
        return wrap64(sum(vals))
    return min(vals) if op == "min" else max(vals)


def run_both(prog, edb):
    """Semi-naive and naive oracle runs, which must agree on everything."""
    s = O.run(prog, edb, semi=True)
    n = O.run(prog, edb, semi=False)
    assert s.status == n.status, (s.status, n.status)
    if s.ok:
        for k in s.relations:
            assert s.relations[k].rows() == n.relations[k].rows(), k
    else:
        assert (s.error["slot"], s.error["pred"]) == (n.error["slot"], n.error["pred"])
    return s


def int_rows(rel):
    out = []
    for t, _ in rel.rows():
        out.append(tuple(x if k == "int" else ("s", x) for k, x in t))
    return out


# ------------------------------------------------------------- aggregates

OPS = ("count", "sum", "min", "max")


def _w_facts(rng, n_keys, n_rows, lo, hi):
    seen, facts = set(), []
    while len(facts) < n_rows:
        k, x = rng.randrange(n_keys), rng.randint(lo, hi)
        if (k, x) not in seen:
            seen.add((k, x))
            facts.append((k, x))
    return facts


@pytest.mark.parametrize("op", OPS)
def test_fold_single_literal_vs_groupby(op):
    """r(X, op<V>) :- w(X, V).  Negative values, single-member groups, and
    groups of many values."""
    rng = random.Random(hash(op) & 0xffff)
    prog = CoreProgram([pred("w", 2, BOOL, 0), pred("r", 2, BOOL, 1)],
                       [rule(atom("r", v(0), Agg(op, 1)), [pos("w", v(0), v(1))], 2, 1)], 2)
    for trial in range(40):
        facts = _w_facts(rng, rng.randint(1, 6), rng.randint(1, 25), -50, 50)
        edb = [("w", (Sym(k), IntV(x)), None) for k, x in facts]
        r = run_both(prog, edb)
        assert r.ok
        groups = defaultdict(list)
        for k, x in facts:
            groups[k].append(x)
        want = sorted((("s", k), fold(op, xs)) for k, xs in groups.items())
        assert int_rows(r.relations["r"]) == want, trial
        assert all(t[1][0] == "int" for t, _ in r.relations["r"].rows())


@pytest.mark.parametrize("op", OPS)
def test_fold_counts_solutions_not_values(op):
    """r(X, op<V>) :- e(X, Y), w(Y, V).  The fold runs over every solution of
    the body (naive.rs:344-363), so a value reached through two Y's counts
    twice in count and sum; min/max see the same multiset."""
    rng = random.Random(0x5eed + len(op))
    prog = CoreProgram([pred("e", 2, BOOL, 0), pred("w", 2, BOOL, 0), pred("r", 2, BOOL, 1)],
                       [rule(atom("r", v(0), Agg(op, 2)), [pos("e", v(0), v(1)), pos("w", v(1), v(2))], 3, 1)], 2)
    for trial in range(40):
        n = rng.randint(2, 6)
        es = sorted({(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(1, 12))})
        ws = _w_facts(rng, n, rng.randint(1, 10), -9, 9)
        edb = [("e", (Sym(a), Sym(b)), None) for a, b in es] + [("w", (Sym(k), IntV(x)), None) for k, x in ws]
        r = run_both(prog, edb)
        assert r.ok
        groups = defaultdict(list)
        for a, b in es:
            for k, x in ws:
                if k == b:
                    groups[a].append(x)
        want = sorted((("s", k), fold(op, xs)) for k, xs in groups.items())
        assert int_rows(r.relations["r"]) == want, trial


@pytest.mark.parametrize("op", OPS)
def test_fold_aggregate_first_and_two_keys(op):
    """r(op<V>, X, Y) :- t(X, Y, V): the Int result takes the aggregate's own
    position, the group key the other head positions in order."""
    rng = random.Random(0xa99 + len(op))
    prog = CoreProgram([pred("t", 3, BOOL, 0), pred("r", 3, BOOL, 1)],
                       [rule(atom("r", Agg(op, 2), v(0), v(1)), [pos("t", v(0), v(1), v(2))], 3, 1)], 2)
    for trial in range(30):
        rows = sorted({(rng.randrange(3), rng.randrange(3), rng.randint(-1000, 1000)) for _ in range(rng.randint(1, 20))})
        edb = [("t", (Sym(a), Sym(b), IntV(x)), None) for a, b, x in rows]
        r = run_both(prog, edb)
        groups = defaultdict(list)
        for a, b, x in rows:
            groups[(a, b)].append(x)
        want = sorted((fold(op, xs), ("s", a), ("s", b)) for (a, b), xs in groups.items())
        assert int_rows(r.relations["r"]) == want, trial


@pytest.mark.parametrize("op", OPS)
def test_fold_global_group_and_empty_body(op):
    """r(op<V>) :- w(X, V): one tuple over every solution; an empty body
    yields no tuple at all, not a zero (naive.rs:365-378)."""
    prog = CoreProgram([pred("w", 2, BOOL, 0), pred("r", 1, BOOL, 1)],
                       [rule(atom("r", Agg(op, 1)), [pos("w", v(0), v(1))], 2, 1)], 2)
    r = run_both(prog, [])
    assert r.ok and len(r.relations["r"]) == 0
    vals = [7, -3, 7, 12]
    edb = [("w", (Sym(i), IntV(x)), None) for i, x in enumerate(vals)]
    r = run_both(prog, edb)
    assert int_rows(r.relations["r"]) == [(fold(op, vals),)]


def test_sum_wraps_like_the_release_build():
    prog = CoreProgram([pred("w", 2, BOOL, 0), pred("r", 2, BOOL, 1)],
                       [rule(atom("r", v(0), Agg("sum", 1)), [pos("w", v(0), v(1))], 2, 1)], 2)
    vals = [2**62, 2**62 + 1, 2**62 - 1, -5]  # distinct: equal facts collapse
    edb = [("w", (Sym(0), IntV(x)), None) for x in vals]
    r = run_both(prog, edb)
    assert int_rows(r.relations["r"]) == [(("s", 0), wrap64(sum(vals)))]
    assert wrap64(sum(vals)) < 0


# ------------------------------------------------------------------ errors

@pytest.mark.parametrize("case", ERROR_CASES, ids=[c.label for c in ERROR_CASES])
def test_error_status(case):
    """Each hand-written program's outcome, read off the reference's raise
    points (tests/helpers.py ERROR_CASES names them), in both the
    semi-naive and the naive restatement."""
    r = run_both(case.prog, case.edb)
    assert r.status == case.status, (r.status, r.error)
    if case.slot is not None:
        assert r.error["slot"] == case.slot
    if case.pred is not None:
        assert r.error["pred"] == case.pred
    if case.status == "ok":
        for name, rows in (case.want or {}).items():
            assert int_rows(r.relations[name]) == rows, name


# --------------------------------------------------------------- triangles

def _oriented(edges, n, by_degree):
    if not by_degree:
        o = orient_simple(edges)
        return (np.array([a for a, _ in o], np.uint32), np.array([b for _, b in o], np.uint32))
    s = orient_simple(edges)
    a = np.array([x for x, _ in s], np.uint32)
    b = np.array([y for _, y in s], np.uint32)
    x, y = wl.orient_by_degree(a, b, n)
    k = np.unique(x.astype(np.uint64) << np.uint64(32) | y.astype(np.uint64))
    return (k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xffffffff)).astype(np.uint32)


@pytest.mark.parametrize("by_degree", [False, True])
def test_tri_count_oriented_vs_brute_force(by_degree):
    rng = random.Random(0x7c + by_degree)
    for trial in range(150):
        n = rng.randint(3, 24)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 4 * n))]
        a, b = _oriented(edges, n, by_degree)
        assert O.tri_count_oriented(a, b, n) == brute_triangles(edges, n), trial


@pytest.mark.parametrize("by_degree", [False, True])
def test_tri_count_oriented_closed_forms(by_degree):
    for n in (3, 4, 5, 9, 17, 40):
        edges = [(x, y) for x in range(n) for y in range(x + 1, n)]
        a, b = _oriented(edges, n, by_degree)
        assert O.tri_count_oriented(a, b, n) == math.comb(n, 3)
    h, br = 7, 60  # crates/strata-gpu/examples/tri.rs:72
    edges = [(i, j) for i in range(h) for j in range(i + 1, h)] + [(hub, h + k) for k in range(br) for hub in range(h)]
    a, b = _oriented(edges, h + br, by_degree)
    assert O.tri_count_oriented(a, b, h + br) == br * math.comb(h, 2) + math.comb(h, 3)
    # a triangle-free bipartite graph
    edges = [(x, 10 + y) for x in range(10) for y in range(10)]
    a, b = _oriented(edges, 20, by_degree)
    assert O.tri_count_oriented(a, b, 20) == 0
    assert O.tri_count_oriented(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 5) == 0


# ------------------------------------------- multi-source Dijkstra checker

from helpers import hop_bellman_ford  # noqa: E402
from strata_k.core_ir import cost_program  # noqa: E402

NONE = np.iinfo(np.int64).max


def test_cost_dijkstra_vs_hop_bellman_ford():
    """Every source of small random graphs (zero weights, duplicates with
    different weights, self-loops, cycles through the source): weights and
    the fewest hops among cheapest walks, against the hop-aware
    Bellman-Ford in tests/helpers.py."""
    rng = random.Random(0xd1a)
    for trial in range(150):
        n = rng.randint(2, 12)
        edges = [(rng.randrange(n), rng.randrange(n), rng.randint(0, 9)) for _ in range(rng.randint(0, 3 * n))]
        s = np.array([a for a, _, _ in edges], np.uint32)
        d = np.array([b for _, b, _ in edges], np.uint32)
        w = np.array([x for _, _, x in edges], np.int64)
        W, H, mh = O.cost_dijkstra(n, s, d, w, np.arange(n))
        cost, maxh = hop_bellman_ford(edges)
        got = {(a, b): int(W[a, b]) for a in range(n) for b in range(n) if W[a, b] != NONE}
        assert got == cost, trial
        assert mh == maxh, trial
        assert ((W == NONE) == (H == np.iinfo(np.uint32).max)).all()


def test_cost_dijkstra_vs_semi_naive_rounds():
    """The checker's rows and round count equal the semi-naive restatement's
    on random multi-source programs (cost_program, src facts of weight 0)."""
    rng = random.Random(0x5a11)
    for trial in range(60):
        n = rng.randint(3, 30)
        m = rng.randint(1, 4 * n)
        edges = [(rng.randrange(n), rng.randrange(n), rng.randint(1, 50)) for _ in range(m)]
        srcs = sorted(rng.sample(range(n), rng.randint(1, min(5, n))))
        s = np.array([a for a, _, _ in edges], np.uint32)
        d = np.array([b for _, b, _ in edges], np.uint32)
        w = np.array([x for _, _, x in edges], np.int64)
        W, _, mh = O.cost_dijkstra(n, s, d, w, srcs)
        bulk = {"edge": ([s, d], w, None), "src": ([np.array(srcs, np.uint32)], np.zeros(len(srcs), np.int64), None)}
        r = O.run(cost_program(), bulk=bulk)
        assert r.ok
        rel = r.relations["cost"]
        S, Z, WW = O.cost_rows(W, srcs)
        assert (rel.vals[:, 0] == S).all() and (rel.vals[:, 1] == Z).all() and (rel.w == WW).all(), trial
        assert r.iterations == [mh], trial


def test_cost_dijkstra_closed_forms():
    """Weighted chain (trop_diff.rs:103-118): cost(0,j) = 1+..+j in j hops; a
    unit ring of n nodes: cost(s,s) = n in n hops, the round count n."""
    n = 300
    W, H, mh = O.cost_dijkstra(n + 1, np.arange(n), np.arange(1, n + 1), np.arange(1, n + 1), [0])
    j = np.arange(1, n + 1)
    assert (W[0, 1:] == j * (j + 1) // 2).all() and W[0, 0] == NONE
    assert (H[0, 1:] == j).all() and mh == n
    n = 50
    W, H, mh = O.cost_dijkstra(n, np.arange(n), (np.arange(n) + 1) % n, np.ones(n), [0, 7])
    assert W[0, 0] == n and W[1, 7] == n and H[1, 7] == n and mh == n
    assert W[1, 6] == n - 1 and W[0, 3] == 3
    # an equal-weight tie: 0->1->2 (1+1) against 0->2 (2): the fewer hops win
    W, H, mh = O.cost_dijkstra(3, [0, 1, 0], [1, 2, 2], [1, 1, 2], [0])
    assert W[0, 2] == 2 and H[0, 2] == 1 and mh == 1
