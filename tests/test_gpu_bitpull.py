"""GPU: word-parallel Bool pull rounds (a destination row ORs the delta rows
its fixed relation names, 32 lanes per word) against the oracle and against
push rounds, with the pull forced off (0), chosen by the cost model (1) and
forced on (2), with and without device-driven rounds (a hipGraph of
rounds relaunched until a device-side flag says the delta is empty).
Covers rows narrower than a word (the lane-level pull), one word, several
words and several word blocks, both linear forms and body orders, rows
spanning two cell columns, a second recursive rule that must push in the
same round, and batched programs (program id as a row column) whose round
counts differ."""
import random

import numpy as np
import pytest

import strata_k as sk
from oracle import oracle as O
from strata_k.core_ir import BOOL, Agg, CoreProgram, Sym, atom, neg, pos, pred, rule, tc_program, v

pytestmark = pytest.mark.gpu


def rows(rel):
    return [tuple(g.id if hasattr(g, "id") else ("int", g.value) for g in t) for t, _ in rel.rows()]


def oracle_rows(rel):
    return [tuple(x[1] if x[0] == "sym" else ("int", x[1]) for x in t) for t, _ in rel.rows()]


def random_edges(seed, n, m):
    rnd = random.Random(seed)
    es = set()
    while len(es) < m:
        a, b = rnd.randrange(n), rnd.randrange(n)
        if a != b:
            es.add((a, b))
    return sorted(es)


def check_modes(ctx, prog, edb, label):
    ref = O.run(prog, edb)
    assert ref.ok, ref.status
    for mode, graph in ((0, True), (1, True), (1, False), (2, True), (2, False)):
        db = ctx.run_semi_naive(prog, edb, pull=mode, graph=graph)
        lab = f"{label}[pull={mode},graph={graph}]"
        assert db.iterations == ref.iterations, (lab, db.iterations, ref.iterations)
        for p in prog.predicates:
            assert rows(db.relation(p.name)) == oracle_rows(ref.relations[p.name]), (lab, p.name)
    return ref


def right_linear():
    """path(X,Y) :- edge(X,Y).  path(X,Z) :- path(X,Y), edge(Y,Z)."""
    return CoreProgram([pred("edge", 2, BOOL), pred("path", 2, BOOL)],
                       [rule(atom("path", v(0), v(1)), [pos("edge", v(0), v(1))], 2),
                        rule(atom("path", v(0), v(2)), [pos("path", v(0), v(1)), pos("edge", v(1), v(2))], 3)], 1)


def delta_first_left_linear():
    """path(X,Y) :- edge(X,Y).  path(X,Z) :- path(Y,Z), edge(X,Y)."""
    return CoreProgram([pred("edge", 2, BOOL), pred("path", 2, BOOL)],
                       [rule(atom("path", v(0), v(1)), [pos("edge", v(0), v(1))], 2),
                        rule(atom("path", v(0), v(2)), [pos("path", v(1), v(2)), pos("edge", v(0), v(1))], 3)], 1)


@pytest.mark.parametrize("n,m", [(10, 25), (24, 60), (40, 90), (100, 260), (300, 700), (700, 1500)])
@pytest.mark.parametrize("shape", ["left", "right", "delta_first"])
def test_tc_shapes(ctx, n, m, shape):
    """Rows of 16 lanes (under a word: lane-level pull), 32, 64, 128, 512 and
    1024 lanes (up to 16 word blocks of 64 words)."""
    prog = {"left": tc_program(BOOL), "right": right_linear(), "delta_first": delta_first_left_linear()}[shape]
    edb = [("edge", (Sym(a), Sym(b)), None) for a, b in random_edges(1000 * n + m, n, m)]
    check_modes(ctx, prog, edb, f"{shape}{n}")


def test_chain_and_hub(ctx):
    """A 400-chain (400 rounds, one new lane per row per round) and a hub whose
    row has more than 256 in-edges (split over chunks)."""
    edb = [("edge", (Sym(i), Sym(i + 1)), None) for i in range(400)]
    ref = check_modes(ctx, tc_program(BOOL), edb, "chain400")
    assert len(ref.relations["path"]) == 400 * 401 // 2
    edb = [("edge", (Sym(0), Sym(i)), None) for i in range(1, 600)]
    edb += [("edge", (Sym(i), Sym(i % 37 + 600)), None) for i in range(1, 600)]
    edb += [("edge", (Sym(600 + i), Sym(600 + (i + 1) % 37)), None) for i in range(37)]
    check_modes(ctx, tc_program(BOOL), edb, "hub")
    check_modes(ctx, right_linear(), edb, "hub_right")


def test_three_column_rows(ctx):
    """r(A,X,Z) :- r(A,Y,Z), e(A,X,Y): rows are (A, X) pairs, lanes Z."""
    prog = CoreProgram([pred("e", 3, BOOL), pred("s", 3, BOOL), pred("r", 3, BOOL)],
                       [rule(atom("r", v(0), v(1), v(2)), [pos("s", v(0), v(1), v(2))], 3),
                        rule(atom("r", v(0), v(1), v(3)), [pos("r", v(0), v(2), v(3)), pos("e", v(0), v(1), v(2))],
                             4)], 1)
    rnd = random.Random(7)
    edb = []
    for a in range(3):
        for x, y in random_edges(50 + a, 70, 160):
            edb.append(("e", (Sym(a), Sym(x), Sym(y)), None))
        for _ in range(40):
            edb.append(("s", (Sym(a), Sym(rnd.randrange(70)), Sym(rnd.randrange(70))), None))
    check_modes(ctx, prog, edb, "rows3")


def test_second_rule_pushes(ctx):
    """Two recursive rules in one stratum: the left-linear one pulls, the
    right-linear one (lanes on the other column) pushes in the same rounds."""
    prog = CoreProgram([pred("edge", 2, BOOL), pred("link", 2, BOOL), pred("path", 2, BOOL)],
                       [rule(atom("path", v(0), v(1)), [pos("edge", v(0), v(1))], 2),
                        rule(atom("path", v(0), v(2)), [pos("edge", v(0), v(1)), pos("path", v(1), v(2))], 3),
                        rule(atom("path", v(0), v(2)), [pos("path", v(0), v(1)), pos("link", v(1), v(2))], 3)], 1)
    edb = [("edge", (Sym(a), Sym(b)), None) for a, b in random_edges(11, 120, 150)]
    edb += [("link", (Sym(a), Sym(b)), None) for a, b in random_edges(12, 120, 90)]
    check_modes(ctx, prog, edb, "two_rules")


def test_later_strata_read_the_closure(ctx):
    """The closure's stratum ends in device-driven rounds; later strata read
    its full state: a count aggregate per node and a negation."""
    prog = CoreProgram([pred("edge", 2, BOOL, 0), pred("node", 1, BOOL, 0), pred("path", 2, BOOL, 0),
                        pred("reach_cnt", 2, BOOL, 1), pred("unreached", 2, BOOL, 1)],
                       [rule(atom("path", v(0), v(1)), [pos("edge", v(0), v(1))], 2, 0),
                        rule(atom("path", v(0), v(2)), [pos("edge", v(0), v(1)), pos("path", v(1), v(2))], 3, 0),
                        rule(atom("reach_cnt", v(0), Agg("count", 1)), [pos("path", v(0), v(1))], 2, 1),
                        rule(atom("unreached", v(0), v(1)), [pos("node", v(0)), pos("node", v(1)),
                                                              neg("path", v(0), v(1))], 2, 1)], 2)
    edb = [("edge", (Sym(a), Sym(b)), None) for a, b in random_edges(77, 90, 170)]
    edb += [("node", (Sym(i),), None) for i in range(0, 90, 3)]
    check_modes(ctx, prog, edb, "strata")


def test_batched_tc_matches_plain(ctx):
    """Batched Bool closures (rows keyed by program id and node) equal their
    plain runs and the oracle, program by program."""
    progs = []
    for i in range(12):
        n = [20, 45, 130, 300][i % 4]
        progs.append(random_edges(500 + i, n, 2 * n + i))
    progs.append([(i, i + 1) for i in range(150)])  # 150 rounds
    progs.append([(i, i + 1) for i in range(60)] + [(60, 0)])  # a 61-cycle
    b = sk.Batch(ctx)
    for es in progs:
        s = np.array([a for a, _ in es], np.uint32)
        d = np.array([c for _, c in es], np.uint32)
        b.add(tc_program(BOOL), {"edge": ([s, d], None)})
    b.load()
    b.execute()
    for i, es in enumerate(progs):
        edb = [("edge", (Sym(a), Sym(c)), None) for a, c in es]
        ref = O.run(tc_program(BOOL), edb)
        cols, _ = b.relation_sym(i, "path")
        got = list(zip(cols[0].tolist(), cols[1].tolist()))
        assert got == oracle_rows(ref.relations["path"]), i
        assert b.iterations(i) == ref.iterations, (i, b.iterations(i), ref.iterations)
    b.free()
