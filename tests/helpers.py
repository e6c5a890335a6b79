"""Shared test helpers: reference fixtures, independent algorithms, random programs.

The independent algorithms here (BFS closure, hop-aware Bellman-Ford,
brute-force triangles / 4-cliques) are not the algorithm under test; they pin
the oracle (and through it the engine) by computing the same answers another
way.
"""
from __future__ import annotations

import json
import os
import random
from itertools import combinations
from typing import Dict, List, Tuple

import numpy as np

from strata_k.core_ir import (BOOL, TROP, Agg, Const, CoreAtom, CorePred, CoreProgram, CoreRule, Int, IntV, Neg,
                              Pos, Sym, Var, atom, neg, pos, pred, rule, v)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


class SymbolDict:
    """First-seen interning (crates/strata-ir/src/dict.rs:35-43)."""

    def __init__(self):
        self.by_str: Dict[str, int] = {}
        self.by_id: List[str] = []

    def intern(self, s: str) -> int:
        if s not in self.by_str:
            self.by_str[s] = len(self.by_id)
            self.by_id.append(s)
        return self.by_str[s]

    def sym(self, s: str) -> Sym:
        return Sym(self.intern(s))


# ------------------------------------------------------------ programs

def interp_pred(name, sem, arity=2, stratum=0):
    return CorePred(name, arity, sem, stratum)


def negation_program():
    """reach(Y) :- edge(X,Y), reach(X).   unreach(X) :- node(X), not reach(X)."""
    return CoreProgram(
        [pred("node", 1, BOOL, 0), pred("edge", 2, BOOL, 0), pred("reach", 1, BOOL, 0), pred("unreach", 1, BOOL, 1)],
        [rule(atom("reach", v(1)), [pos("edge", v(0), v(1)), pos("reach", v(0))], 2, 0),
         rule(atom("unreach", v(0)), [pos("node", v(0)), neg("reach", v(0))], 1, 1)],
        2)


def neg_same_stratum_program():
    """p(X) :- q(X), not p(X)."""
    return CoreProgram([pred("q", 1, BOOL, 0), pred("p", 1, BOOL, 0)],
                       [rule(atom("p", v(0)), [pos("q", v(0)), neg("p", v(0))], 1, 0)], 1)


def outdeg_program():
    """outdeg(X, count<Y>) :- edge(X,Y)."""
    return CoreProgram([pred("edge", 2, BOOL, 0), pred("outdeg", 2, BOOL, 1)],
                       [rule(atom("outdeg", v(0), Agg("count", 1)), [pos("edge", v(0), v(1))], 2, 1)], 2)


def sum_program():
    """total(X, sum<V>) :- w(X,V)."""
    return CoreProgram([pred("w", 2, BOOL, 0), pred("total", 2, BOOL, 1)],
                       [rule(atom("total", v(0), Agg("sum", 1)), [pos("w", v(0), v(1))], 2, 1)], 2)


# ------------------------------------------------------ independent checks

def bfs_closure(edges, n=None) -> Tuple[set, int]:
    """Pairs (x, y) joined by a path of >= 1 edge, and the largest shortest
    path length (the semi-naive round count of the left-linear closure)."""
    adj: Dict[int, List[int]] = {}
    nodes = set()
    for a, b in edges:
        adj.setdefault(a, []).append(b)
        nodes.add(a)
        nodes.add(b)
    out = set()
    maxd = 0
    for s in nodes:
        dist = {}
        frontier = [s]
        d = 0
        while frontier:
            d += 1
            nxt = []
            for x in frontier:
                for y in adj.get(x, ()):
                    if y not in dist:
                        dist[y] = d
                        nxt.append(y)
            frontier = nxt
        for y, dd in dist.items():
            out.add((s, y))
            maxd = max(maxd, dd)
    return out, maxd


def hop_bellman_ford(edges) -> Tuple[Dict[Tuple[int, int], int], int]:
    """All-pairs cheapest cost over paths of >= 1 edge, with the fewest edges
    among optimal paths; returns (cost map, max of those hop counts) — the
    max is the semi-naive round count of the Trop closure."""
    best: Dict[Tuple[int, int], int] = {}
    for a, b, w in edges:
        best[(a, b)] = min(best.get((a, b), w), w)
    adj: Dict[int, List[Tuple[int, int]]] = {}
    for (a, b), w in best.items():
        adj.setdefault(a, []).append((b, w))
    nodes = sorted({x for e in best for x in e})
    cost: Dict[Tuple[int, int], int] = {}
    maxh = 0
    for s in nodes:
        dist: Dict[int, Tuple[int, int]] = {}
        for b, w in adj.get(s, ()):
            if b not in dist or (w, 1) < dist[b]:
                dist[b] = (w, 1)
        for _ in range(len(nodes) + 1):
            changed = False
            for x, (dw, dh) in list(dist.items()):
                for y, w in adj.get(x, ()):
                    cand = (dw + w, dh + 1)
                    if y not in dist or cand < dist[y]:
                        dist[y] = cand
                        changed = True
            if not changed:
                break
        for y, (dw, dh) in dist.items():
            cost[(s, y)] = dw
            maxh = max(maxh, dh)
    return cost, maxh


def floyd_warshall(edges, n) -> Dict[Tuple[int, int], int]:
    """crates/strata-gpu/tests/trop_diff.rs:14-44 style oracle (no free self-paths)."""
    INF = 1 << 62
    d = [[INF] * n for _ in range(n)]
    for a, b, w in edges:
        d[a][b] = min(d[a][b], w)
    for k in range(n):
        for i in range(n):
            if d[i][k] == INF:
                continue
            for j in range(n):
                if d[k][j] != INF and d[i][k] + d[k][j] < d[i][j]:
                    d[i][j] = d[i][k] + d[k][j]
    return {(i, j): d[i][j] for i in range(n) for j in range(n) if d[i][j] < INF}


def orient_simple(edges):
    s = set()
    for a, b in edges:
        if a == b:
            continue
        s.add((min(a, b), max(a, b)))
    return sorted(s)


def brute_triangles(edges, n) -> int:
    es = set(orient_simple(edges))
    t = 0
    for a in range(n):
        for b in range(a + 1, n):
            if (a, b) not in es:
                continue
            for c in range(b + 1, n):
                if (a, c) in es and (b, c) in es:
                    t += 1
    return t


def brute_4cliques(edges, n) -> int:
    es = set(orient_simple(edges))
    e = lambda x, y: (min(x, y), max(x, y)) in es
    return sum(1 for q in combinations(range(n), 4) if all(e(x, y) for x, y in combinations(q, 2)))


def tri_bulk(edges):
    o = orient_simple(edges)
    a = np.array([x for x, _ in o], dtype=np.uint32)
    b = np.array([y for _, y in o], dtype=np.uint32)
    return {"e": ([a, b], None, None)}


# -------------------------------------------------------- random programs

def random_program(seed: int, big: bool = False):
    """A random safe Core-IR program and EDB, in the shape of the reference's
    fuzzer (crates/strata-eval/tests/fuzz.rs:47-155): Bool or Trop arity-2
    predicates, 2-5 rules of 1-3 body atoms over a few variable slots, head
    variables drawn from body variables, optionally a stratum-1 Bool rule with
    a negated stratum-0 literal.  `big` widens the domain and the fact count
    and adds constants, arity-1 atoms and repeated variables."""
    r = random.Random(seed)
    trop = r.random() < 0.5
    sem = TROP if trop else BOOL
    domain = r.randint(2, 4) if not big else r.randint(4, 24)
    n_pred = r.randint(2, 3)
    nvars = r.randint(2, 3) if not big else r.randint(2, 4)
    preds = [pred(f"p{i}", 2, sem, 0) for i in range(n_pred)]
    rules = []
    for _ in range(r.randint(2, 5)):
        body = []
        bvars = []
        for _ in range(r.randint(1, 3)):
            p = r.randrange(n_pred)
            args = []
            for _ in range(2):
                if big and r.random() < 0.1:
                    args.append(Const(r.randrange(domain)))
                else:
                    s = r.randrange(nvars)
                    args.append(Var(s))
                    bvars.append(s)
            body.append(Pos(CoreAtom(f"p{p}", args)))
        if not bvars:
            bvars = [0]
            body.append(Pos(CoreAtom("p0", [Var(0), Var(0)])))
        h1 = bvars[r.randrange(len(bvars))]
        h2 = bvars[r.randrange(len(bvars))]
        rules.append(CoreRule(CoreAtom(f"p{r.randrange(n_pred)}", [Var(h1), Var(h2)]), body, 0, nvars,
                              neg_weight_cycle_check=trop))
    num_strata = 1
    if not trop and r.random() < 0.35:
        q = f"p{n_pred}"
        preds.append(pred(q, 2, BOOL, 1))
        p = r.randrange(n_pred)
        ng = r.randrange(n_pred)
        rules.append(CoreRule(CoreAtom(q, [Var(0), Var(1)]),
                              [Pos(CoreAtom(f"p{p}", [Var(0), Var(1)])), Neg(CoreAtom(f"p{ng}", [Var(0), Var(1)]))],
                              1, 2))
        num_strata = 2
    n_facts = r.randint(2, 8) if not big else r.randint(8, 60)
    edb = []
    for _ in range(n_facts):
        p = f"p{r.randrange(n_pred)}"
        t = (Sym(r.randrange(domain)), Sym(r.randrange(domain)))
        a = r.randint(0, 4) if trop else None
        edb.append((p, t, a))
    return CoreProgram(preds, rules, num_strata), edb


def rows_of(rel) -> list:
    """Comparable rows of an oracle relation or engine relation."""
    return rel.rows()


def oracle_rows_as_engine(orel):
    """Oracle relation rows -> (tuple of Sym/IntV, ann None/int/'inf')."""
    out = []
    for t, a in orel.rows():
        tt = tuple(Sym(x) if k == "sym" else IntV(x) for k, x in t)
        out.append((tt, a))
    return out


# ------------------------------------------------------- error programs

class ErrorCase:
    """A program, its EDB and the outcome the reference gives it: a status
    (the oracle's names), the error's slot or predicate where it has one, and
    for "ok" the expected rows of some relations (Int values as ints, Sym
    values as ("s", id))."""

    def __init__(self, label, prog, edb, status, slot=None, pred=None, want=None):
        self.label, self.prog, self.edb, self.status = label, prog, edb, status
        self.slot, self.pred, self.want = slot, pred, want


def _error_cases():
    P = lambda *specs: [pred(*s) for s in specs]  # noqa: E731
    q_a = [("q", (Sym(3),), None)]
    cases = []
    # EDB fact naming no declared predicate (seminaive.rs:30-34)
    prog = CoreProgram(P(("q", 1, BOOL)), [], 1)
    cases += [ErrorCase("unknown_fact", prog, [("nope", (Sym(0),), None)], "unknown_pred", pred="nope"),
              ErrorCase("known_fact", prog, [("q", (Sym(0),), None)], "ok")]
    # p(X) :- q(X), [not] nope(X): the relation lookup (seminaive.rs:254-257)
    # and the negated literal's stratum lookup (naive.rs:545-547) happen when
    # solve reaches the literal; with q empty it never does
    for negated in (False, True):
        lit = neg("nope", v(0)) if negated else pos("nope", v(0))
        prog = CoreProgram(P(("q", 1, BOOL, 0), ("p", 1, BOOL, 1)),
                           [rule(atom("p", v(0)), [pos("q", v(0)), lit], 1, 1)], 2)
        tag = "neg" if negated else "pos"
        cases += [ErrorCase(f"unknown_body_{tag}", prog, q_a, "unknown_pred", pred="nope"),
                  ErrorCase(f"unknown_body_{tag}_unreached", prog, [], "ok", want={"p": []})]
    # the head relation is looked up before any body is solved
    # (seminaive.rs:128-131, naive.rs:193-195), whatever the data
    prog = CoreProgram(P(("q", 1, BOOL)), [rule(atom("nope", v(0)), [pos("q", v(0))], 1)], 1)
    cases.append(ErrorCase("unknown_head", prog, [], "unknown_pred", pred="nope"))
    # p(X, Y) :- q(X): slot 1 never bound; build_head fails on the first
    # solution (naive.rs:485-490)
    prog = CoreProgram(P(("q", 1, BOOL), ("p", 2, BOOL)), [rule(atom("p", v(0), v(1)), [pos("q", v(0))], 2)], 1)
    cases += [ErrorCase("unsafe_head", prog, q_a, "unsafe_head_var", slot=1),
              ErrorCase("unsafe_head_unreached", prog, [], "ok", want={"p": []})]
    # aggregand slot unbound (naive.rs:360); group-key slot unbound
    # (head_arg_value, naive.rs:513-522)
    for tag, head in (("aggregand", atom("p", v(0), Agg("sum", 1))), ("group_key", atom("p", v(1), Agg("sum", 0)))):
        prog = CoreProgram(P(("q", 1, BOOL, 0), ("p", 2, BOOL, 1)), [rule(head, [pos("q", v(0))], 2, 1)], 2)
        cases += [ErrorCase(f"unsafe_{tag}", prog, [("q", (IntV(4),), None)], "unsafe_head_var", slot=1),
                  ErrorCase(f"unsafe_{tag}_unreached", prog, [], "ok", want={"p": []})]
    # p(X) :- q(X), not r(X, Y): Y unbound when the negation is reached
    # (ground_atom with negated = true, naive.rs:485-488)
    prog = CoreProgram(P(("q", 1, BOOL, 0), ("r", 2, BOOL, 0), ("p", 1, BOOL, 1)),
                       [rule(atom("p", v(0)), [pos("q", v(0)), neg("r", v(0), v(1))], 2, 1)], 2)
    cases += [ErrorCase("unsafe_neg", prog, [("q", (Sym(0),), None)], "unsafe_neg_var", slot=1),
              ErrorCase("unsafe_neg_unreached", prog, [("r", (Sym(0), Sym(0)), None)], "ok", want={"p": []})]
    # p(X) :- q(X), not p(X): the stratum check sits in negation_holds
    # (naive.rs:548-553), reached only through a solution of q
    prog = CoreProgram(P(("q", 1, BOOL), ("p", 1, BOOL)), [rule(atom("p", v(0)), [pos("q", v(0)), neg("p", v(0))], 1)],
                       1)
    cases += [ErrorCase("neg_same_stratum", prog, q_a, "negation_not_stratified", pred="p"),
              ErrorCase("neg_same_stratum_unreached", prog, [], "ok", want={"p": []})]
    # c(count<X>) :- d(X) with d in the rule's own stratum: the body strata
    # are checked before any solving (naive.rs:307-318), whatever the data
    prog = CoreProgram(P(("d", 1, BOOL, 0), ("c", 1, BOOL, 0)), [rule(atom("c", Agg("count", 0)), [pos("d", v(0))], 1)],
                       1)
    cases.append(ErrorCase("aggregate_in_recursion", prog, [], "aggregate_in_recursion", pred="d"))
    prog = CoreProgram(P(("d", 1, BOOL, 0), ("c", 1, BOOL, 1)),
                       [rule(atom("c", Agg("count", 0)), [pos("d", v(0))], 1, 1)], 2)
    cases.append(ErrorCase("aggregate_lower_stratum", prog, [("d", (Sym(0),), None), ("d", (Sym(1),), None)], "ok",
                           want={"c": [(2,)]}))
    # two aggregate terms in one head (naive.rs:320-330), checked after the
    # body strata: a recursive malformed rule reports AggregateInRecursion
    prog = CoreProgram(P(("q", 2, BOOL, 0), ("m", 2, BOOL, 1)),
                       [rule(atom("m", Agg("count", 0), Agg("sum", 1)), [pos("q", v(0), v(1))], 2, 1)], 2)
    cases.append(ErrorCase("malformed_head", prog, [], "malformed_aggregate_head"))
    prog = CoreProgram(P(("q", 2, BOOL, 0), ("m", 2, BOOL, 0)),
                       [rule(atom("m", Agg("count", 0), Agg("sum", 1)), [pos("q", v(0), v(1))], 2, 0)], 1)
    cases.append(ErrorCase("malformed_head_recursive", prog, [], "aggregate_in_recursion", pred="q"))
    # sum/min/max over a Sym value fail (naive.rs:384-403); count does not
    # look at the values; no group, no fold, no error
    mixed = [("q", (Sym(0), IntV(5)), None), ("q", (Sym(0), Sym(7)), None)]
    ints = [("q", (Sym(0), IntV(5)), None), ("q", (Sym(1), IntV(-2)), None)]
    for op in ("sum", "min", "max", "count"):
        prog = CoreProgram(P(("q", 2, BOOL, 0), ("s", 2, BOOL, 1)),
                           [rule(atom("s", v(0), Agg(op, 1)), [pos("q", v(0), v(1))], 2, 1)], 2)
        if op == "count":
            cases.append(ErrorCase("count_over_sym", prog, mixed, "ok", want={"s": [(("s", 0), 2)]}))
            continue
        cases += [ErrorCase(f"{op}_over_sym", prog, mixed, "non_integer_aggregand"),
                  ErrorCase(f"{op}_no_group", prog, [], "ok", want={"s": []}),
                  ErrorCase(f"{op}_over_ints", prog, ints, "ok", want={"s": [(("s", 0), 5), (("s", 1), -2)]})]
    # prob-or is refused inside fold_aggregate (naive.rs:401): only once a
    # group exists
    prog = CoreProgram(P(("q", 2, BOOL, 0), ("s", 2, BOOL, 1)),
                       [rule(atom("s", v(0), Agg("prob_or", 1)), [pos("q", v(0), v(1))], 2, 1)], 2)
    cases += [ErrorCase("prob_or", prog, [("q", (Sym(0), IntV(1)), None)], "unsupported"),
              ErrorCase("prob_or_no_group", prog, [], "ok", want={"s": []})]
    return cases


ERROR_CASES = _error_cases()
