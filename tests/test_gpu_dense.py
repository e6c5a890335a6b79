"""GPU: the dense cell path (relations ⊕-applied in place during their
stratum) against the oracle and against the sorted-list path, on the cases
its encoding singles out: the finite weight INT64_MAX (whose cell code is the
"absent" code), +inf, relations that start with facts, rules that read a
relation's full state (non-linear recursion), head constants, column domains
ranked from small EDB columns, 128-bit keys, and errors."""
import random

import numpy as np
import pytest

import strata_k as sk
from oracle import oracle as O
from strata_k import workloads as wl
from strata_k.core_ir import (BOOL, INF, TROP, Const, CoreProgram, Sym, atom, cost_program, pos, pred, rule,
                              tc_program, v)

pytestmark = pytest.mark.gpu
I64_MAX = 2**63 - 1


def rows(rel):
    return [(tuple(g.id if isinstance(g, Sym) else ("int", g.value) for g in t), "inf" if a is INF else a)
            for t, a in rel.rows()]


def oracle_rows(rel):
    return [(tuple(x[1] if x[0] == "sym" else ("int", x[1]) for x in t), a) for t, a in rel.rows()]


def both(ctx, prog, edb=(), bulk=None, label=""):
    ref = O.run(prog, edb, bulk=bulk)
    for dense in (True, False):
        lab = f"{label}[{'dense' if dense else 'sorted'}]"
        try:
            db = ctx.run_semi_naive(prog, edb, bulk=bulk, dense=dense)
        except sk.EngineError as e:
            assert not ref.ok, (lab, e.kind)
            continue
        assert ref.ok, (lab, ref.status)
        assert db.iterations == ref.iterations, (lab, db.iterations, ref.iterations)
        for p in prog.predicates:
            assert rows(db.relation(p.name)) == oracle_rows(ref.relations[p.name]), (lab, p.name)
    return ref


@pytest.mark.parametrize("case", range(6))
def test_weight_extremes(ctx, case):
    """INT64_MAX reached exactly (no overflow), INT64_MAX facts, +inf facts
    improved by finite paths, and +inf paths that stay +inf."""
    M = I64_MAX
    edges = {
        0: [(0, 1, M - 5), (1, 2, 5), (0, 3, INF), (3, 4, 1), (0, 4, M)],
        1: [(0, 1, M), (1, 2, 0), (2, 3, -1), (0, 3, INF)],
        2: [(0, 1, INF), (1, 2, INF), (0, 2, 7), (2, 0, 3)],
        3: [(0, 1, M - 1), (1, 0, 1)],  # the cycle overflows on its second lap
        4: [(0, 1, -M), (1, 2, -1), (2, 3, 0)],
        5: [(a, (a * 7 + 3) % 12, (M // 4) if a % 3 == 0 else a) for a in range(12)] + [(5, 6, INF)],
    }[case]
    edb = [("edge", (Sym(a), Sym(b)), w) for a, b, w in edges]
    both(ctx, tc_program(TROP), edb, label=f"extreme{case}")


@pytest.mark.parametrize("wt", [2**31 - 1, 2**32 - 2, 2**32 - 1, 2**32 + 5])
def test_narrow_cells_widen(ctx, wt):
    """Non-negative weights start on narrow (u32) cells when every fact fits;
    a chain whose costs pass 2^32 mid-fixpoint forces the cells wide and the
    round to be applied again; a fact at or past 2^32-1 starts wide."""
    edb = [("edge", (Sym(i), Sym(i + 1)), wt) for i in range(12)]
    edb += [("edge", (Sym(3), Sym(9)), 1), ("edge", (Sym(5), Sym(2)), INF), ("edge", (Sym(2), Sym(0)), 7)]
    both(ctx, tc_program(TROP), edb, label=f"narrow{wt}")
    src_edb = [e for e in edb] + [("src", (Sym(0),), 0), ("src", (Sym(4),), 0)]
    both(ctx, cost_program(), src_edb, label=f"narrowcost{wt}")


def test_idb_with_max_and_inf_facts(ctx):
    """A relation that starts with INT64_MAX and +inf facts: derivations below
    them are fresh, equal ones are not."""
    prog = CoreProgram([pred("a", 2, TROP), pred("h", 2, TROP)],
                       [rule(atom("h", v(0), v(1)), [pos("a", v(0), v(1))], 2),
                        rule(atom("h", v(0), v(2)), [pos("h", v(0), v(1)), pos("a", v(1), v(2))], 3)], 1)
    M = I64_MAX
    edb = [("a", (Sym(0), Sym(1)), 4), ("a", (Sym(1), Sym(2)), M - 10), ("a", (Sym(2), Sym(3)), INF),
           ("h", (Sym(0), Sym(1)), M), ("h", (Sym(0), Sym(2)), INF), ("h", (Sym(5), Sym(6)), INF),
           ("h", (Sym(2), Sym(3)), M), ("h", (Sym(1), Sym(2)), M - 10)]
    both(ctx, prog, edb, label="idbfacts")


@pytest.mark.parametrize("sem", [BOOL, TROP])
def test_nonlinear_closure(ctx, sem):
    """path(x,z) :- path(x,y), path(y,z): every round reads path's full state."""
    prog = CoreProgram([pred("edge", 2, sem), pred("path", 2, sem)],
                       [rule(atom("path", v(0), v(1)), [pos("edge", v(0), v(1))], 2),
                        rule(atom("path", v(0), v(2)), [pos("path", v(0), v(1)), pos("path", v(1), v(2))], 3)], 1)
    rng = random.Random(7)
    for trial in range(8):
        n = rng.randint(5, 60)
        es = {(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(n, 3 * n))}
        edb = [("edge", (Sym(a), Sym(b)), (rng.randint(0, 20) if sem == TROP else None)) for a, b in sorted(es)]
        both(ctx, prog, edb, label=f"nonlinear{trial}")


def test_ranked_domains_and_head_constants(ctx):
    """A 3-column relation whose first two columns come from small EDB columns
    (ranked domains), one rule writing a head constant, sparse Sym ids."""
    prog = CoreProgram(
        [pred("q", 2, TROP), pred("e", 2, TROP), pred("r", 3, TROP)],
        [rule(atom("r", v(0), v(1), v(2)), [pos("q", v(0), v(1)), pos("e", v(1), v(2))], 3),
         rule(atom("r", v(0), v(1), v(3)), [pos("r", v(0), v(1), v(2)), pos("e", v(2), v(3))], 4),
         rule(atom("r", v(0), Const(77777), v(2)), [pos("r", v(0), v(1), v(2)), pos("q", v(0), v(1))], 3)], 1)
    rng = np.random.default_rng(3)
    n = 3000
    s = (rng.integers(0, 400, n) * 97).astype(np.uint32)
    d = (rng.integers(0, 400, n) * 97).astype(np.uint32)
    w = rng.integers(0, 50, n).astype(np.int64)
    qa = np.array([5, 9, 9, 1000, 123456], np.uint32)
    qb = (np.array([3, 17, 40, 200, 399], np.uint32) * 97).astype(np.uint32)
    bulk = {"e": ([s, d], w, None), "q": ([qa, qb], np.array([0, 1, 2, 3, 4], np.int64), None)}
    both(ctx, prog, bulk=bulk, label="ranked")


def test_rows_without_outputs_outside_domain(ctx):
    """Sources with no edges: cost's first column is ranked over edge's first
    column (fewer rows than src), so those sources lie outside the domain; the
    rows that emit nothing must not be decoded."""
    edb = [("edge", (Sym(1), Sym(7)), 17), ("edge", (Sym(7), Sym(3)), 27), ("edge", (Sym(4), Sym(0)), 28)]
    edb += [("src", (Sym(s),), 0) for s in (1, 5, 6, 7)]
    both(ctx, cost_program(), edb, label="emptyrows")


def test_empty_and_single_cell(ctx):
    prog = CoreProgram([pred("e", 2, BOOL), pred("z", 0, BOOL), pred("p", 2, BOOL)],
                       [rule(atom("z"), [pos("e", v(0), v(1))], 2),
                        rule(atom("p", v(0), v(1)), [pos("e", v(0), v(1)), pos("z")], 2)], 1)
    both(ctx, prog, [], label="empty")
    both(ctx, prog, [("e", (Sym(3), Sym(4)), None)], label="one")


def test_errors_on_dense_path(ctx):
    # negative cycle: the cap universe+2 is reached
    edb = [("edge", (Sym(0), Sym(1)), 1), ("edge", (Sym(1), Sym(0)), -2), ("src", (Sym(0),), 0)]
    both(ctx, cost_program(), edb, label="negcycle")
    # overflow in ⊗ (the first round that adds past INT64_MAX)
    edb = [("edge", (Sym(0), Sym(1)), I64_MAX), ("edge", (Sym(1), Sym(2)), 1), ("src", (Sym(0),), 0)]
    ref = both(ctx, cost_program(), edb, label="overflow")
    assert ref.status == "overflow"


def test_config1_dense_equals_sorted(ctx):
    w = wl.config1()
    a = ctx.run_semi_naive(w.program, bulk=w.bulk, want=["path"], dense=True)
    b = ctx.run_semi_naive(w.program, bulk=w.bulk, want=["path"], dense=False)
    assert (a.relation("path").vals == b.relation("path").vals).all()
    assert a.iterations == b.iterations


def test_config2_full_dense_equals_sorted(ctx):
    """Full RMAT-20 multi-source cost: both paths give the same relation,
    weights and round count (a size-independent check of the dense path)."""
    w = wl.config2()
    out = []
    for dense in (True, False):
        run = ctx.prepare(w.program)
        run.set_dense(dense)
        for p, (cols, wt, inf) in w.bulk.items():
            run.add_sym(p, cols, wt, inf)
        run.load()
        run.execute()
        (s, z), cw, _ = run.relation_sym("cost")
        out.append((s, z, cw, run.iterations()))
        run.free()
    (s1, z1, w1, i1), (s2, z2, w2, i2) = out
    assert len(s1) == len(s2) and i1 == i2
    assert (s1 == s2).all() and (z1 == z2).all() and (w1 == w2).all()
    print("config2 full:", len(s1), "tuples,", i1[0], "rounds, dense == sorted")


def reach_program() -> CoreProgram:
    """Bool multi-source reachability, the shape of ``cost_program`` over Bool."""
    return CoreProgram(
        [pred("src", 1, BOOL), pred("edge", 2, BOOL), pred("cost", 2, BOOL)],
        [
            rule(atom("cost", v(0), v(1)), [pos("src", v(0)), pos("edge", v(0), v(1))], 2),
            rule(atom("cost", v(0), v(2)), [pos("cost", v(0), v(1)), pos("edge", v(1), v(2))], 3),
        ],
        1,
    )


@pytest.mark.parametrize("sem", [BOOL, TROP])
@pytest.mark.parametrize("nsrc", [1, 3, 17, 40, 64])
def test_pull_rounds_all_row_widths(ctx, sem, nsrc):
    """Destination-side (pull) rounds at every row width: the lane column
    holds nsrc sources, so a cell row has 1, 4, 32 or 64 lanes.  Tuples,
    weights and round counts equal the oracle's and the sorted path's, and,
    with pulls forced, the pull kernel must have run (Bool rows of whole
    words take the word-parallel pull)."""
    w = wl.config2(seed=5 + nsrc, scale=8, edge_factor=16, n_sources=nsrc)
    prog = cost_program() if sem == TROP else reach_program()
    bulk = w.bulk if sem == TROP else {p: (cols, None, None) for p, (cols, _, _) in w.bulk.items()}
    both(ctx, prog, bulk=bulk, label=f"pull-{sem}-{nsrc}")
    run = ctx.prepare(prog)
    run.set_pull(2)  # pull whenever the shape allows, whatever the round's size
    for p, (cols, wt, inf) in bulk.items():
        run.add_sym(p, cols, wt, inf)
    run.load()
    ctx.set_profiling(True)
    try:
        run.execute()
        names = {r["name"] for r in run.kernel_records() if r["launches"]}
    finally:
        ctx.set_profiling(False)
        run.free()
    want = "join_bitpull" if sem == BOOL and nsrc > 16 else "join_pull"
    assert want in names, sorted(names)
