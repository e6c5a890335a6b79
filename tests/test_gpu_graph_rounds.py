"""GPU: device-driven Trop pull rounds (a hipGraph of rounds with the round
cap, the overflow / widen status and the push-or-pull choice decided on the
device) against the oracle and against host-driven rounds: converging runs,
runs that hand back to push rounds, a widen of narrow cells mid-fixpoint, an
overflow and a negative cycle, each under pull by cost model, pull forced,
graph on and off."""
import random

import numpy as np
import pytest

import strata_k as sk
from oracle import oracle as O
from strata_k import workloads as wl
from strata_k.core_ir import INF, Sym, cost_program

pytestmark = pytest.mark.gpu

ERR = {"overflow": "Overflow", "negative_weight_cycle": "NegativeWeightCycle"}
# (pull mode, graph rounds, packed pull deltas)
MODES = [(1, True, True), (1, False, True), (2, True, True), (2, False, True), (0, True, True), (2, True, False),
         (1, False, False)]


def rows(rel):
    return [(tuple(g.id for g in t), "inf" if a is INF else a) for t, a in rel.rows()]


def oracle_rows(rel):
    return [(tuple(x[1] for x in t), a) for t, a in rel.rows()]


def check(ctx, edb=(), bulk=None, label="", prog=None, out="cost"):
    prog = prog or cost_program()
    ref = O.run(prog, edb, bulk=bulk)
    for pull, graph, pack in MODES:
        lab = f"{label}[pull={pull},graph={graph},pack={pack}]"
        try:
            db = ctx.run_semi_naive(prog, edb, bulk=bulk, pull=pull, graph=graph, pack=pack)
        except sk.EngineError as e:
            assert not ref.ok, (lab, e.kind)
            assert e.kind == ERR[ref.status], (lab, e.kind, ref.status)
            continue
        assert ref.ok, (lab, ref.status)
        assert db.iterations == ref.iterations, (lab, db.iterations, ref.iterations)
        assert rows(db.relation(out)) == oracle_rows(ref.relations[out]), lab
    return ref


@pytest.mark.parametrize("scale,nsrc", [(8, 4), (9, 17), (10, 64), (11, 33)])
def test_rmat_costs(ctx, scale, nsrc):
    w = wl.config2(seed=40 + scale, scale=scale, edge_factor=16, n_sources=nsrc)
    check(ctx, bulk=w.bulk, label=f"rmat{scale}-{nsrc}")


def test_long_chain_many_graph_launches(ctx):
    """A 300-node weighted chain plus shortcuts: 300 rounds, many launches of
    the four-round graph, the last one converging mid-graph."""
    edb = [("edge", (Sym(i), Sym(i + 1)), i + 1) for i in range(300)]
    edb += [("edge", (Sym(i), Sym(i + 7)), 1000 * (i + 1)) for i in range(0, 290, 13)]
    edb += [("src", (Sym(s),), 0) for s in (0, 5, 100)]
    ref = check(ctx, edb, label="chain300")
    assert ref.iterations[0] == 300


def test_widen_mid_fixpoint(ctx):
    """Narrow cells until a cost passes 2^32 - 1 in a device-driven round: the
    stratum restarts on wide cells with the same result and round count."""
    rnd = random.Random(3)
    edb = [("edge", (Sym(i), Sym(i + 1)), 2**30 + i) for i in range(40)]
    edb += [("edge", (Sym(rnd.randrange(41)), Sym(rnd.randrange(41))), rnd.randrange(1, 1000)) for _ in range(60)]
    edb += [("src", (Sym(s),), 0) for s in (0, 3, 9)]
    check(ctx, edb, label="widen")


def test_overflow_in_device_rounds(ctx):
    """Weights near INT64_MAX / 4: the fourth hop overflows in a device-driven
    round and the run fails with Overflow, as the oracle does."""
    big = (2**63 - 1) // 4
    edb = [("edge", (Sym(i), Sym(i + 1)), big + i) for i in range(8)]
    edb += [("src", (Sym(0),), 0)]
    ref = check(ctx, edb, label="overflow")
    assert ref.status == "overflow"


def test_negative_cycle_in_device_rounds(ctx):
    """A reachable negative cycle: rounds continue until the cap (universe + 2)
    is passed inside the device-driven rounds."""
    edb = [("edge", (Sym(i), Sym(i + 1)), 5) for i in range(6)]
    edb += [("edge", (Sym(6), Sym(2)), -30)]
    edb += [("src", (Sym(0),), 0), ("src", (Sym(4),), 0)]
    ref = check(ctx, edb, label="negcycle")
    assert ref.status == "negative_weight_cycle"


def test_packed_delta_widens_past_2_26(ctx):
    """Staged weights below 2^26 start on packed pull deltas; sums crossing
    2^26 after two hops widen the cells mid-fixpoint (the round is redone on
    wide cells), with results equal to the oracle's."""
    rng = random.Random(26)
    n = 400
    edb = [("edge", (Sym(i), Sym((i * 7 + 3) % n)), (1 << 25) + rng.randrange(1000)) for i in range(n)]
    edb += [("edge", (Sym(rng.randrange(n)), Sym(rng.randrange(n))), rng.randint(1, 50)) for _ in range(6 * n)]
    edb += [("src", (Sym(s),), 0) for s in range(0, n, 9)]
    check(ctx, edb, label="pack-widen")


def test_packed_pull_over_lower_stratum_weights(ctx):
    """The pull's fixed relation is derived by a lower stratum (two-hop sums
    of staged weights below 2^26), so its own weights reach 2^26 and beyond:
    the packed pull must not truncate them (they widen the cells)."""
    from strata_k.core_ir import TROP, CoreProgram, atom, pos, pred, rule, v
    prog = CoreProgram([pred("e", 2, TROP, 0), pred("big", 2, TROP, 0), pred("src", 1, TROP, 1),
                        pred("cost", 2, TROP, 1)],
                       [rule(atom("big", v(0), v(2)), [pos("e", v(0), v(1)), pos("e", v(1), v(2))], 3, 0),
                        rule(atom("cost", v(0), v(1)), [pos("src", v(0)), pos("big", v(0), v(1))], 2, 1),
                        rule(atom("cost", v(0), v(2)), [pos("cost", v(0), v(1)), pos("big", v(1), v(2))], 3, 1)], 2)
    rng = random.Random(27)
    n = 300
    for big in (False, True):
        hi = (1 << 25) + 5 if big else 1000
        edb = [("e", (Sym(rng.randrange(n)), Sym(rng.randrange(n))), rng.randint(1, hi)) for _ in range(5 * n)]
        edb += [("src", (Sym(s),), 0) for s in range(0, n, 11)]
        check(ctx, edb, label=f"lower-stratum-{big}", prog=prog)
