"""CPU: the oracle (C restatement of strata-eval) pinned before it is trusted.

Each check is chosen so that a plausible slip in the restatement (a dropped
delta position, the cap off by one, ⊕ keeping the wrong weight, an iteration
counted at the wrong time, a join column swapped) fails at least one of them:
  * the reference's own known-answer vectors (tests/golden/reference_vectors.json);
  * closed forms (chains, chain unions, complete graphs, hub graphs);
  * independent algorithms (BFS closure, hop-aware Bellman-Ford, Floyd-Warshall,
    brute-force triangles);
  * naive == semi-naive on random programs (the reference's own I5 invariant,
    crates/strata-eval/tests/fuzz.rs:157-176);
  * indexed == full-scan evaluation (the index must not change what unifies).
"""
import math
import random

import numpy as np
import pytest

from helpers import (SymbolDict, bfs_closure, brute_triangles, floyd_warshall, golden, hop_bellman_ford,
                     neg_same_stratum_program, negation_program, orient_simple, outdeg_program, random_program,
                     sum_program, tri_bulk)
from oracle import oracle as O
from strata_k.core_ir import BOOL, TROP, IntV, Sym, cost_program, tc_program, triangle_program

G = golden()


def syms(rel):
    return [tuple(x for _, x in t) for t, _ in rel.rows()]


# -------------------------------------------------------------- interp.rs

def test_interp_transitive_closure():
    g = G["interp"]["transitive_closure_bool"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), None) for a, b in g["edges"]]
    r = O.run(tc_program(BOOL), edb)
    assert r.ok
    want = sorted((d.intern(a), d.intern(b)) for a, b in g["path"])
    assert syms(r.relations["path"]) == want
    assert r.iterations == [3]  # longest shortest path a->d


def test_interp_sssp():
    g = G["interp"]["sssp_trop_min_plus"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), w) for a, b, w in g["edges"]]
    r = O.run(tc_program(TROP, path="reach"), edb)
    assert r.ok
    got = {(t[0][1], t[1][1]): a for t, a in r.relations["reach"].rows()}
    for a, b, w in g["reach"]:
        assert got[(d.intern(a), d.intern(b))] == w


@pytest.mark.parametrize("semi", [True, False])
def test_interp_overflow(semi):
    g = G["interp"]["trop_overflow_is_error_not_wrap"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), w) for a, b, w in g["edges"]]
    r = O.run(tc_program(TROP, path="reach"), edb, semi=semi)
    assert r.status == "overflow"
    assert (r.error["lhs"], r.error["rhs"]) == (2**63 - 1, 1)


@pytest.mark.parametrize("semi", [True, False])
def test_interp_negative_cycle(semi):
    g = G["interp"]["negative_weight_cycle_is_detected"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), w) for a, b, w in g["edges"]]
    r = O.run(tc_program(TROP, path="reach"), edb, semi=semi)
    assert r.status == "negative_weight_cycle"
    # cap = universe(2) + 2; the error is raised when the round count exceeds it
    assert r.iterations[0] == 5


def test_interp_negation():
    g = G["interp"]["stratified_negation_complement"]
    d = SymbolDict()
    edb = [("node", (d.sym(x),), None) for x in g["node"]]
    edb += [("edge", (d.sym(a), d.sym(b)), None) for a, b in g["edge"]]
    edb += [("reach", (d.sym(x),), None) for x in g["reach"]]
    for semi in (True, False):
        r = O.run(negation_program(), edb, semi=semi)
        assert r.ok
        assert syms(r.relations["unreach"]) == [(d.intern(x),) for x in g["unreach"]]
        assert len(r.relations["reach"]) == g["reach_size"]


def test_interp_negation_same_stratum():
    d = SymbolDict()
    edb = [("q", (d.sym("a"),), None)]
    for semi in (True, False):
        assert O.run(neg_same_stratum_program(), edb, semi=semi).status == "negation_not_stratified"


def test_interp_count_and_sum():
    g = G["interp"]["count_aggregate_outdegree"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), None) for a, b in g["edges"]]
    r = O.run(outdeg_program(), edb)
    got = {t[0][1]: t[1][1] for t, _ in r.relations["outdeg"].rows()}
    assert got == {d.intern(k): v for k, v in g["outdeg"].items()}
    assert all(t[1][0] == "int" for t, _ in r.relations["outdeg"].rows())
    g = G["interp"]["sum_aggregate"]
    d = SymbolDict()
    edb = [("w", (d.sym(a), IntV(x)), None) for a, x in g["w"]]
    r = O.run(sum_program(), edb)
    got = {t[0][1]: t[1][1] for t, _ in r.relations["total"].rows()}
    assert got == {d.intern(k): v for k, v in g["total"].items()}


# ------------------------------------------------------ tc_diff / trop_diff

def test_tc_fixed_cases():
    for case in G["tc_diff"]["fixed_cases"]["cases"]:
        edb = [("edge", (Sym(a), Sym(b)), None) for a, b in case["edges"]]
        r = O.run(tc_program(BOOL), edb)
        assert syms(r.relations["path"]) == [tuple(p) for p in case["path"]]


def test_tc_chain_400_closed_form():
    n = G["tc_diff"]["chain_scale_semi_naive"]["n"]
    bulk = {"edge": ([np.arange(n), np.arange(1, n + 1)], None, None)}
    r = O.run(tc_program(BOOL), bulk=bulk)
    rel = r.relations["path"]
    assert len(rel) == G["tc_diff"]["chain_scale_semi_naive"]["closure_size"]
    want = np.array([(i, j) for i in range(n + 1) for j in range(i + 1, n + 1)], dtype=np.int64)
    assert (rel.vals == want).all()
    assert r.iterations == [n]


def test_trop_fixed_and_weighted_chain():
    g = G["trop_diff"]["fixed_cases"]
    edb = [("edge", (Sym(a), Sym(b)), w) for a, b, w in g["edges"]]
    r = O.run(tc_program(TROP), edb)
    got = [(t[0][1], t[1][1], a) for t, a in r.relations["path"].rows()]
    assert got == [tuple(x) for x in g["reach"]]
    n = G["trop_diff"]["weighted_chain"]["n"]
    bulk = {"edge": ([np.arange(n), np.arange(1, n + 1)], np.arange(1, n + 1), None)}
    r = O.run(tc_program(TROP), bulk=bulk)
    rel = r.relations["path"]
    want = [(i, j, sum(range(i + 1, j + 1))) for i in range(n + 1) for j in range(i + 1, n + 1)]
    got = list(zip(rel.vals[:, 0].tolist(), rel.vals[:, 1].tolist(), rel.w.tolist()))
    assert got == want
    assert r.iterations == [n]


def test_chain_union_closed_form():
    # crates/strata-gpu/examples/bench.rs:22-42,65-76
    K, Cn = 50, 6
    src = np.array([k * (Cn + 1) + i for k in range(K) for i in range(Cn)])
    bulk = {"edge": ([src, src + 1], np.ones(len(src), np.int64), None)}
    r = O.run(tc_program(TROP), bulk=bulk)
    rel = r.relations["path"]
    assert len(rel) == K * Cn * (Cn + 1) // 2
    assert (rel.w == rel.vals[:, 1] - rel.vals[:, 0]).all()


# ---------------------------------------------------------- independent

def test_tc_random_vs_bfs():
    rng = random.Random(0x1234)
    for trial in range(120):
        n = rng.randint(3, 12)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 2 * n + 2))]
        r = O.run(tc_program(BOOL), [("edge", (Sym(a), Sym(b)), None) for a, b in edges])
        want, maxd = bfs_closure(edges)
        assert set(syms(r.relations["path"])) == want, trial
        assert r.iterations == [maxd if edges else 0], trial


def test_trop_random_vs_bellman_ford_and_floyd_warshall():
    rng = random.Random(0xbeef)
    for trial in range(120):
        n = rng.randint(3, 9)
        edges = [(rng.randrange(n), rng.randrange(n), rng.randint(0, 9)) for _ in range(rng.randint(0, 2 * n + 2))]
        r = O.run(tc_program(TROP), [("edge", (Sym(a), Sym(b)), w) for a, b, w in edges])
        assert r.ok
        cost, maxh = hop_bellman_ford(edges)
        got = {(t[0][1], t[1][1]): a for t, a in r.relations["path"].rows()}
        assert got == cost, trial
        assert got == floyd_warshall(edges, n), trial
        assert r.iterations == [maxh if edges else 0], trial


def test_multi_source_cost_vs_bellman_ford():
    rng = random.Random(77)
    for trial in range(40):
        n = rng.randint(4, 14)
        edges = [(rng.randrange(n), rng.randrange(n), rng.randint(1, 20)) for _ in range(rng.randint(1, 3 * n))]
        srcs = sorted(rng.sample(range(n), rng.randint(1, 3)))
        edb = [("edge", (Sym(a), Sym(b)), w) for a, b, w in edges] + [("src", (Sym(s),), 0) for s in srcs]
        r = O.run(cost_program(), edb)
        cost, _ = hop_bellman_ford(edges)
        want = {k: w for k, w in cost.items() if k[0] in srcs}
        got = {(t[0][1], t[1][1]): a for t, a in r.relations["cost"].rows()}
        assert got == want, trial


def test_triangles_known_answers_and_brute_force():
    for case in G["wcoj_diff"]["fixed_cases"]["cases"]:
        r = O.run(triangle_program(), bulk=tri_bulk(case["edges"]))
        assert len(r.relations["tri"]) == case["triangles"]
        cnt = r.relations["ntri"].rows()
        assert cnt == ([((("int", case["triangles"]),), None)] if case["triangles"] else [])
    for n in G["wcoj_diff"]["complete_graphs"]["n"][:3]:
        edges = [(a, b) for a in range(n) for b in range(a + 1, n)]
        r = O.run(triangle_program(), bulk=tri_bulk(edges))
        assert len(r.relations["tri"]) == math.comb(n, 3)
        assert r.iterations == [1, 0]
    rng = random.Random(0x7ea)
    for trial in range(60):
        n = rng.randint(5, 20)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 3 * n))]
        r = O.run(triangle_program(), bulk=tri_bulk(edges))
        assert len(r.relations["tri"]) == brute_triangles(edges, n), trial


def test_hub_bridge_closed_form():
    # crates/strata-gpu/examples/tri.rs:35-73 at a small size
    h, b = 6, 40
    edges = [(i, j) for i in range(h) for j in range(i + 1, h)]
    edges += [(hub, h + k) for k in range(b) for hub in range(h)]
    r = O.run(triangle_program(), bulk=tri_bulk(edges))
    assert len(r.relations["tri"]) == b * math.comb(h, 2) + math.comb(h, 3)


def test_routing_workload_lines():
    d = SymbolDict()
    edb = []
    with open(__file__.replace("test_oracle.py", "golden/routing_links.tsv")) as f:
        for line in f:
            a, b, w = line.split()
            edb.append(("link", (d.sym(a), d.sym(b)), int(w)))
    from strata_k.core_ir import CoreProgram, atom, pos, pred, rule, v
    prog = CoreProgram([pred("link", 2, TROP), pred("route", 2, TROP)],
                       [rule(atom("route", v(0), v(1)), [pos("link", v(0), v(1))], 2),
                        rule(atom("route", v(0), v(2)), [pos("link", v(0), v(1)), pos("route", v(1), v(2))], 3)], 1)
    r = O.run(prog, edb)
    got = {(t[0][1], t[1][1]): a for t, a in r.relations["route"].rows()}
    for k, w in G["routing_workload"]["lines"].items():
        x, y = k[len("route("):-1].split(", ")
        assert got[(d.intern(x), d.intern(y))] == w, k


# ------------------------------------------------- naive == semi == indexed

def _same(a, b):
    if a.status != b.status:
        return False
    if not a.ok:
        return True
    return all(a.relations[k].rows() == b.relations[k].rows() for k in a.relations)


def test_naive_equals_semi_naive_fuzz():
    for seed in range(400):
        prog, edb = random_program(seed)
        s = O.run(prog, edb, semi=True)
        n = O.run(prog, edb, semi=False)
        assert _same(s, n), seed


def test_indexed_equals_full_scan_fuzz():
    for seed in range(150):
        prog, edb = random_program(10_000 + seed, big=True)
        a = O.run(prog, edb, indexed=True)
        b = O.run(prog, edb, indexed=False)
        assert _same(a, b), seed
        assert a.iterations == b.iterations


def test_asp_filter_restatement_truth_table():
    """keep_literal (crates/strata-gpu/src/asp.rs:53-68): a positive literal
    dies on an impossible atom and is dropped on a certain one; a negative
    literal dies on a certain atom and is dropped on an impossible one;
    otherwise kept.  Atoms: 0 impossible, 1 possible, 2 certain."""
    poss, cert = [0, 1, 1], [0, 0, 1]
    cases = [([1], None), ([2], [2]), ([3], []), ([-1], []), ([-2], [-2]), ([-3], None),
             ([2, -2, 3, -1], [2, -2]), ([2, -3, 1], None), ([], [])]
    head = list(range(len(cases)))
    bs, bl = [0], []
    for lits, _ in cases:
        bl += lits
        bs.append(len(bl))
    oh, ob, ol = O.asp_simplify_filter(head, bs, bl, poss, cert)
    want = [(h, k) for h, (_, k) in zip(head, cases) if k is not None]
    assert oh == [h for h, _ in want]
    assert [ol[ob[i]:ob[i + 1]] for i in range(len(oh))] == [k for _, k in want]
    assert ob[-1] == len(ol)
