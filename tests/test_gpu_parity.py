"""GPU: the engine against the oracle, bit-exact, on the reference's vectors,
random graphs and random programs.  Everything goes through the C ABI."""
import math
import random

import numpy as np
import pytest

import strata_k as sk
from helpers import (SymbolDict, brute_4cliques, brute_triangles, golden, negation_program,
                     neg_same_stratum_program, orient_simple, outdeg_program, random_program, sum_program,
                     tri_bulk)
from oracle import oracle as O
from strata_k.core_ir import BOOL, INF, TROP, Agg, Const, IntV, Sym, cost_program, tc_program, triangle_program

pytestmark = pytest.mark.gpu
G = golden()

ERR = {"overflow": "Overflow", "negative_weight_cycle": "NegativeWeightCycle", "unknown_pred": "UnknownPred",
       "unsafe_head_var": "UnsafeHeadVar", "unsafe_neg_var": "UnsafeNegVar",
       "negation_not_stratified": "NegationNotStratified", "aggregate_in_recursion": "AggregateInRecursion",
       "malformed_aggregate_head": "MalformedAggregateHead", "non_integer_aggregand": "NonIntegerAggregand",
       "unsupported": "Unsupported", "panic": "SemiringMismatch"}


def engine_rows(rel):
    out = []
    for t, a in rel.rows():
        tt = tuple(("sym", g.id) if isinstance(g, Sym) else ("int", g.value) for g in t)
        out.append((tt, "inf" if a is INF else a))
    return out


def compare(ctx, prog, edb=(), bulk=None, label=""):
    """Run the oracle and the engine on both evaluation paths (dense cells and
    sorted lists); assert identical outcome, relations, weights and round counts."""
    ref = O.run(prog, edb, bulk=bulk)
    out = None
    for dense in (True, False):
        lab = f"{label}[{'dense' if dense else 'sorted'}]"
        try:
            db = ctx.run_semi_naive(prog, edb, bulk=bulk, dense=dense)
        except sk.EngineError as e:
            assert not ref.ok, f"{lab}: engine {e.kind} but oracle ok"
            assert e.kind == ERR[ref.status], f"{lab}: {e.kind} vs {ref.status}"
            continue
        assert ref.ok, f"{lab}: oracle {ref.status} but engine ok"
        assert db.iterations == ref.iterations, f"{lab}: iterations {db.iterations} vs {ref.iterations}"
        for p in prog.predicates:
            g, r = engine_rows(db.relation(p.name)), ref.relations[p.name].rows()
            assert g == r, f"{lab}: relation {p.name} differs ({len(g)} vs {len(r)})"
        if dense:
            out = db
    return out, ref


# ------------------------------------------------------ reference vectors

def test_interp_vectors(ctx):
    I = G["interp"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), None) for a, b in I["transitive_closure_bool"]["edges"]]
    db, _ = compare(ctx, tc_program(BOOL), edb, label="tc")
    assert len(db.relation("path")) == 6
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), w) for a, b, w in I["sssp_trop_min_plus"]["edges"]]
    db, _ = compare(ctx, tc_program(TROP, path="reach"), edb, label="sssp")
    got = {(t[0].id, t[1].id): a for t, a in db.relation("reach").rows()}
    assert got[(d.intern("a"), d.intern("c"))] == 5
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), w) for a, b, w in I["trop_overflow_is_error_not_wrap"]["edges"]]
    with pytest.raises(sk.EvalError) as e:
        ctx.run_semi_naive(tc_program(TROP, path="reach"), edb)
    assert e.value.kind == "Overflow"
    assert (e.value.info["lhs"], e.value.info["rhs"]) == (2**63 - 1, 1)
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), w) for a, b, w in I["negative_weight_cycle_is_detected"]["edges"]]
    with pytest.raises(sk.EvalError) as e:
        ctx.run_semi_naive(tc_program(TROP, path="reach"), edb)
    assert e.value.kind == "NegativeWeightCycle"
    g = I["stratified_negation_complement"]
    d = SymbolDict()
    edb = [("node", (d.sym(x),), None) for x in g["node"]]
    edb += [("edge", (d.sym(a), d.sym(b)), None) for a, b in g["edge"]]
    edb += [("reach", (d.sym(x),), None) for x in g["reach"]]
    db, _ = compare(ctx, negation_program(), edb, label="negation")
    assert [t[0].id for t, _ in db.relation("unreach").rows()] == [d.intern("c")]
    with pytest.raises(sk.EvalError) as e:
        ctx.run_semi_naive(neg_same_stratum_program(), [("q", (Sym(0),), None)])
    assert e.value.kind == "NegationNotStratified"
    g = I["count_aggregate_outdegree"]
    d = SymbolDict()
    edb = [("edge", (d.sym(a), d.sym(b)), None) for a, b in g["edges"]]
    db, _ = compare(ctx, outdeg_program(), edb, label="count")
    got = {t[0].id: t[1].value for t, _ in db.relation("outdeg").rows()}
    assert got == {d.intern(k): v for k, v in g["outdeg"].items()}
    g = I["sum_aggregate"]
    d = SymbolDict()
    edb = [("w", (d.sym(a), IntV(x)), None) for a, x in g["w"]]
    db, _ = compare(ctx, sum_program(), edb, label="sum")
    got = {t[0].id: t[1].value for t, _ in db.relation("total").rows()}
    assert got == {d.intern(k): v for k, v in g["total"].items()}


def test_tc_and_trop_fixed(ctx):
    for case in G["tc_diff"]["fixed_cases"]["cases"]:
        edb = [("edge", (Sym(a), Sym(b)), None) for a, b in case["edges"]]
        db, _ = compare(ctx, tc_program(BOOL), edb, label=str(case["edges"]))
        assert [(t[0].id, t[1].id) for t, _ in db.relation("path").rows()] == [tuple(p) for p in case["path"]]
    g = G["trop_diff"]["fixed_cases"]
    db, _ = compare(ctx, tc_program(TROP), [("edge", (Sym(a), Sym(b)), w) for a, b, w in g["edges"]])
    assert [(t[0].id, t[1].id, a) for t, a in db.relation("path").rows()] == [tuple(x) for x in g["reach"]]


def test_chain_400_and_weighted_chain_300(ctx):
    n = 400
    db = ctx.run_semi_naive(tc_program(BOOL), bulk={"edge": ([np.arange(n), np.arange(1, n + 1)], None, None)})
    rel = db.relation("path")
    want = np.array([(i, j) for i in range(n + 1) for j in range(i + 1, n + 1)], dtype=np.int64)
    assert len(rel) == 80200 and (rel.vals == want).all()
    assert db.iterations == [400]
    n = 300
    db = ctx.run_semi_naive(tc_program(TROP), bulk={"edge": ([np.arange(n), np.arange(1, n + 1)],
                                                             np.arange(1, n + 1), None)})
    rel = db.relation("path")
    want_w = [sum(range(i + 1, j + 1)) for i in range(n + 1) for j in range(i + 1, n + 1)]
    assert rel.w.tolist() == want_w
    assert db.iterations == [300]


def test_chain_union(ctx):
    K, Cn = 2000, 5
    src = np.array([k * (Cn + 1) + i for k in range(K) for i in range(Cn)])
    for sem in (BOOL, TROP):
        bulk = {"edge": ([src, src + 1], np.ones(len(src), np.int64) if sem == TROP else None, None)}
        db = ctx.run_semi_naive(tc_program(sem), bulk=bulk)
        rel = db.relation("path")
        assert len(rel) == K * Cn * (Cn + 1) // 2
        if sem == TROP:
            assert (rel.w == rel.vals[:, 1] - rel.vals[:, 0]).all()


def test_triangles_vectors(ctx):
    for case in G["wcoj_diff"]["fixed_cases"]["cases"]:
        db, _ = compare(ctx, triangle_program(), bulk=tri_bulk(case["edges"]), label="tri")
        assert len(db.relation("tri")) == case["triangles"]
    for n in G["wcoj_diff"]["complete_graphs"]["n"]:
        edges = [(a, b) for a in range(n) for b in range(a + 1, n)]
        db = ctx.run_semi_naive(triangle_program(), bulk=tri_bulk(edges))
        assert len(db.relation("tri")) == math.comb(n, 3)
        assert db.relation("ntri").rows() == [((IntV(math.comb(n, 3)),), None)]
        assert db.iterations == [1, 0]
    h, b = 8, 300
    edges = [(i, j) for i in range(h) for j in range(i + 1, h)] + [(hub, h + k) for k in range(b) for hub in range(h)]
    db = ctx.run_semi_naive(triangle_program(), bulk=tri_bulk(edges))
    assert len(db.relation("tri")) == b * math.comb(h, 2) + math.comb(h, 3)


@pytest.mark.parametrize("b", [1500, 2500])
def test_triangles_long_lists(ctx, b):
    """Hubs of b leaves each beside a random graph.  A hub's rows hold more
    than 64 list values together, so each takes a wave: (hub, hub) rows
    search a longer list of more than 1,024 values (past the LDS table) in
    global memory, and their 24 (b = 1500) or 40 (b = 2500) hit words per row
    span several of the heavy emit's 64-word chunks; (hub, leaf) rows have an
    empty shorter list (no hit words) and sit between them.  The random
    graph's rows take the one-lane merge.  The tri relation equals the
    oracle's and the count the closed form plus brute force."""
    h = 5
    edges = [(i, j) for i in range(h) for j in range(i + 1, h)] + [(hub, h + k) for k in range(b) for hub in range(h)]
    rnd = random.Random(17)
    base = h + b
    rand = set()
    while len(rand) < 2500:
        x, y = rnd.randrange(300), rnd.randrange(300)
        if x != y:
            rand.add((base + min(x, y), base + max(x, y)))
    edges += sorted(rand)
    db, _ = compare(ctx, triangle_program(), bulk=tri_bulk(edges), label="tri_long")
    want = b * math.comb(h, 2) + math.comb(h, 3) + brute_triangles([(x - base, y - base) for x, y in rand], 300)
    assert len(db.relation("tri")) == want


def test_routing_loaded_and_rendered(ctx):
    """The routing workload loaded from its TSV (first-seen interning, Trop
    weight column) and rendered canonically: the engine's full output text
    equals the oracle's, and the queried lines are the reference's pinned
    ones (crates/strata-cli/tests/workloads.rs:25-32)."""
    from strata_k.io import SymbolDict as Dict, load_input, render_db
    from test_io import _OracleDb, routing_program
    d = Dict()
    facts = load_input(__file__.replace("test_gpu_parity.py", "golden/routing_links.tsv"), "link", 2, TROP,
                       ["node", "node"], d)
    prog = routing_program()
    db = ctx.run_semi_naive(prog, facts)
    ref = O.run(prog, facts)
    assert render_db(db, d) == render_db(_OracleDb(ref), d)
    q = [("route", (Sym(d.get(a)), Sym(d.get(b)))) for a, b in (("n0_0", "n7_7"), ("n0_0", "n4_4"), ("n7_7", "n0_0"))]
    assert render_db(db, d, filter=q).splitlines() == ["route(n0_0, n4_4) = 2", "route(n0_0, n7_7) = 5",
                                                       "route(n7_7, n0_0) = 44"]


def test_routing_workload(ctx):
    d = SymbolDict()
    edb = []
    with open(__file__.replace("test_gpu_parity.py", "golden/routing_links.tsv")) as f:
        for line in f:
            a, b, w = line.split()
            edb.append(("link", (d.sym(a), d.sym(b)), int(w)))
    from strata_k.core_ir import CoreProgram, atom, pos, pred, rule, v
    prog = CoreProgram([pred("link", 2, TROP), pred("route", 2, TROP)],
                       [rule(atom("route", v(0), v(1)), [pos("link", v(0), v(1))], 2),
                        rule(atom("route", v(0), v(2)), [pos("link", v(0), v(1)), pos("route", v(1), v(2))], 3)], 1)
    db, _ = compare(ctx, prog, edb, label="routing")
    got = {(t[0].id, t[1].id): a for t, a in db.relation("route").rows()}
    for k, w in G["routing_workload"]["lines"].items():
        x, y = k[len("route("):-1].split(", ")
        assert got[(d.intern(x), d.intern(y))] == w


# ---------------------------------------------------------------- random

def test_random_graphs(ctx):
    rng = random.Random(0x5151)
    for trial in range(60):
        n = rng.randint(3, 40)
        m = rng.randint(0, 4 * n)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(m)]
        compare(ctx, tc_program(BOOL), [("edge", (Sym(a), Sym(b)), None) for a, b in edges], label=f"tc{trial}")
        wedges = [(a, b, rng.randint(-3 if trial % 7 == 0 else 0, 50)) for a, b in edges]
        compare(ctx, tc_program(TROP), [("edge", (Sym(a), Sym(b)), w) for a, b, w in wedges], label=f"trop{trial}")
        srcs = sorted(rng.sample(range(n), min(n, rng.randint(1, 4))))
        edb = [("edge", (Sym(a), Sym(b)), w) for a, b, w in wedges] + [("src", (Sym(s),), 0) for s in srcs]
        compare(ctx, cost_program(), edb, label=f"cost{trial}")
        compare(ctx, triangle_program(), bulk=tri_bulk(edges), label=f"tri{trial}")


def test_inf_weights(ctx):
    # +inf enters only through explicit facts (docs/ir-encoding.md rule 5); it absorbs ⊗ and loses ⊕
    edb = [("edge", (Sym(0), Sym(1)), INF), ("edge", (Sym(1), Sym(2)), 4), ("edge", (Sym(0), Sym(2)), INF),
           ("edge", (Sym(2), Sym(3)), 1), ("edge", (Sym(0), Sym(3)), 100)]
    db, ref = compare(ctx, tc_program(TROP), edb, label="inf")
    got = {(t[0].id, t[1].id): a for t, a in db.relation("path").rows()}
    assert got[(0, 1)] is INF and got[(0, 2)] is INF and got[(0, 3)] == 100 and got[(1, 3)] == 5


def test_json_program_and_high_ir_facts(ctx):
    """A serialised Core-IR program and High-IR facts with an "inf" weight
    (docs/ir-encoding.md rule 5) run on the device as on the oracle."""
    import json
    from strata_k import io as IO
    doc = json.dumps(IO.core_program_to_json(tc_program(TROP)))
    c = lambda x: {"kind": "const", "data": {"name": x}}  # noqa: E731
    items = [{"kind": "fact", "data": {"atom": {"pred": "edge", "args": [c(a), c(b)]}, "weight": w}}
             for a, b, w in (("s", "a", "inf"), ("a", "b", 4), ("s", "b", "inf"), ("b", "t", 1), ("s", "t", 100))]
    d = IO.SymbolDict()
    edb = IO.high_ir_facts({"ir_version": "0.1.0", "items": items}, d)
    db, ref = compare(ctx, IO.core_program_from_json(doc), edb, label="json")
    text = IO.render_db(db, d)
    assert "path(s, a) = inf\n" in text and "path(s, t) = 100\n" in text and "path(a, t) = 5\n" in text


def test_prepare_json(ctx):
    """sk_prepare_json: a serialised program runs as its struct form does;
    decoding errors and compound terms are refused with their reasons."""
    import json
    from strata_k import io as IO
    edb = [("edge", (Sym(a), Sym(b)), (a * 7 + b) % 13 + 1) for a in range(40) for b in range(40) if (a * b) % 11 == 3]
    edb += [("src", (Sym(0),), 0), ("src", (Sym(5),), 0)]
    prog = cost_program()
    text = json.dumps(IO.core_program_to_json(prog))
    ref = O.run(prog, edb)
    for r in (ctx.prepare_json(text), ctx.prepare(prog)):
        r.add_facts(edb)
        r.load()
        r.execute()
        assert engine_rows(r.relation("cost")) == ref.relations["cost"].rows()
        assert r.iterations() == ref.iterations
        r.free()
    with pytest.raises(sk.EngineError, match="missing field `num_strata`"):
        ctx.prepare_json('{"predicates": [], "rules": []}')
    d = IO.core_program_to_json(prog)
    d["rules"][0]["head"]["args"][0] = {"kind": "compound", "data": {"functor": 3, "args": []}}
    with pytest.raises(sk.EngineError):
        ctx.prepare_json(json.dumps(d))


def test_fuzz_programs(ctx):
    for seed in range(300):
        prog, edb = random_program(seed)
        compare(ctx, prog, edb, label=f"fuzz{seed}")


def test_fuzz_programs_big(ctx):
    for seed in range(150):
        prog, edb = random_program(50_000 + seed, big=True)
        compare(ctx, prog, edb, label=f"bigfuzz{seed}")


def test_error_paths(ctx):
    from strata_k.core_ir import CoreProgram, atom, neg, pos, pred, rule, v
    # unsafe head variable (naive.rs:486-490): slot 2 never bound
    prog = CoreProgram([pred("e", 2, BOOL), pred("h", 2, BOOL)],
                       [rule(atom("h", v(0), v(2)), [pos("e", v(0), v(1))], 3)], 1)
    compare(ctx, prog, [("e", (Sym(0), Sym(1)), None)], label="unsafe_head")
    compare(ctx, prog, [], label="unsafe_head_empty")
    # unsafe negated variable
    prog = CoreProgram([pred("e", 2, BOOL, 0), pred("q", 1, BOOL, 0), pred("h", 1, BOOL, 1)],
                       [rule(atom("h", v(0)), [neg("q", v(1)), pos("e", v(0), v(1))], 2, 1)], 2)
    compare(ctx, prog, [("e", (Sym(0), Sym(1)), None)], label="unsafe_neg")
    # unknown predicate in a body
    prog = CoreProgram([pred("e", 2, BOOL), pred("h", 2, BOOL)],
                       [rule(atom("h", v(0), v(1)), [pos("e", v(0), v(1)), pos("nope", v(1))], 2)], 1)
    compare(ctx, prog, [("e", (Sym(0), Sym(1)), None)], label="unknown")
    # aggregate in recursion
    prog = CoreProgram([pred("e", 2, BOOL, 0), pred("c", 2, BOOL, 0)],
                       [rule(atom("c", v(0), Agg("count", 1)), [pos("e", v(0), v(1))], 2, 0)], 1)
    compare(ctx, prog, [("e", (Sym(0), Sym(1)), None)], label="agg_rec")
    # non-integer aggregand
    prog = CoreProgram([pred("e", 2, BOOL, 0), pred("c", 2, BOOL, 1)],
                       [rule(atom("c", v(0), Agg("sum", 1)), [pos("e", v(0), v(1))], 2, 1)], 2)
    compare(ctx, prog, [("e", (Sym(0), Sym(1)), None)], label="nonint")
    # min / max / count with several groups and Int keys
    for op in ("min", "max", "count", "sum"):
        prog = CoreProgram([pred("w", 2, BOOL, 0), pred("r", 2, BOOL, 1)],
                           [rule(atom("r", v(0), Agg(op, 1)), [pos("w", v(0), v(1))], 2, 1)], 2)
        edb = [("w", (Sym(i % 5), IntV((i * 37) % 23 - 11)), None) for i in range(40)]
        compare(ctx, prog, edb, label=op)


def test_constants_and_repeated_vars(ctx):
    from strata_k.core_ir import CoreProgram, atom, pos, pred, rule, v, Int
    prog = CoreProgram(
        [pred("e", 2, BOOL, 0), pred("loop", 1, BOOL, 0), pred("from3", 1, BOOL, 0), pred("tag", 2, BOOL, 0)],
        [rule(atom("loop", v(0)), [pos("e", v(0), v(0))], 1),
         rule(atom("from3", v(0)), [pos("e", Const(3), v(0))], 1),
         rule(atom("tag", v(0), Const(99)), [pos("from3", v(0)), pos("e", v(0), v(1)), pos("e", v(1), v(0))], 2)],
        1)
    rng = random.Random(3)
    edb = [("e", (Sym(rng.randrange(12)), Sym(rng.randrange(12))), None) for _ in range(60)]
    compare(ctx, prog, edb, label="consts")


# --------------------------------------------------- crate-level functions

def test_crate_functions(ctx):
    rng = random.Random(0xa11ce)
    for trial in range(20):
        n = rng.randint(3, 30)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 3 * n))]
        got = sk.transitive_closure_bool(edges, ctx)
        ref = O.run(tc_program(BOOL), [("edge", (Sym(a), Sym(b)), None) for a, b in edges])
        assert got == [tuple(x) for x in ref.relations["path"].vals.tolist()]
        wedges = [(a, b, rng.randint(0, 9)) for a, b in edges]
        got = sk.shortest_paths_trop(wedges, ctx)
        ref = O.run(tc_program(TROP), [("edge", (Sym(a), Sym(b)), w) for a, b, w in wedges])
        rel = ref.relations["path"]
        assert got == list(zip(rel.vals[:, 0].tolist(), rel.vals[:, 1].tolist(), rel.w.tolist()))
        assert sk.count_triangles(edges, ctx) == brute_triangles(edges, n)
    assert sk.count_triangles([], ctx) == 0
    for case in G["wcoj_diff"]["fourcliques_fixed_and_complete"]["cases"]:
        assert sk.count_4cliques(case["edges"], ctx) == case["k4"]
    for n in G["wcoj_diff"]["fourcliques_fixed_and_complete"]["n"]:
        assert sk.count_4cliques([(a, b) for a in range(n) for b in range(a + 1, n)], ctx) == math.comb(n, 4)
    for trial in range(30):
        n = rng.randint(6, 20)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 4 * n))]
        assert sk.count_4cliques(edges, ctx) == brute_4cliques(edges, n), trial


def _k4_sets(edges):
    """Independent 4-clique count with Python sets: for each id-ordered edge
    (a, b), the common neighbours above b, and the edges among them."""
    nb = {}
    for x, y in edges:
        if x != y:
            nb.setdefault(x, set()).add(y)
            nb.setdefault(y, set()).add(x)
    total = 0
    for a in nb:
        for b in nb[a]:
            if b <= a:
                continue
            t = {c for c in nb[a] & nb[b] if c > b}
            total += sum(1 for c in t for d in nb[c] & t if d > c)
    return total


@pytest.mark.parametrize("mode", [(False, 0), (True, 0), (True, 3)], ids=["auto", "all_waves", "waves_lds_cap_3"])
def test_4cliques_every_path(ctx, mode):
    """count_4cliques (wcoj.rs:162-184) on the reference's fixtures and
    brute-force graphs (wcoj_diff.rs:109-171), complete graphs, hub graphs
    (closed form C(h,4) + b*C(h,3)) and a 6,000-edge random graph against a
    set-based count, with the lane path, the wave path, and the wave path
    with a 3-entry LDS table (common out-neighbour sets above 3 take the
    global-memory membership test)."""
    ctx.set_clique_test(*mode)
    try:
        for case in G["wcoj_diff"]["fourcliques_fixed_and_complete"]["cases"]:
            assert sk.count_4cliques(case["edges"], ctx) == case["k4"]
        for n in (4, 5, 10, 20, 40, 70):
            assert sk.count_4cliques([(a, b) for a in range(n) for b in range(a + 1, n)], ctx) == math.comb(n, 4), n
        rng = random.Random(0x4c1)
        for trial in range(60):
            n = rng.randint(6, 23)
            edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 4 * n))]
            assert sk.count_4cliques(edges, ctx) == brute_4cliques(edges, n), trial
        for h, b in ((5, 30), (9, 200)):
            edges = [(i, j) for i in range(h) for j in range(i + 1, h)]
            edges += [(hub, h + k) for k in range(b) for hub in range(h)]
            edges += [(h + k, hub) for k in range(0, b, 3) for hub in range(h)]  # both directions
            edges += [(h + k, h + k) for k in range(0, b, 7)]  # self-loops
            assert sk.count_4cliques(edges, ctx) == math.comb(h, 4) + b * math.comb(h, 3), (h, b)
        r = np.random.default_rng(5)
        edges = list(zip(r.integers(0, 300, 6000).tolist(), r.integers(0, 300, 6000).tolist()))
        edges += [(i, j) for i in range(280, 300) for j in range(i + 1, 300)]  # a planted K20
        assert sk.count_4cliques(edges, ctx) == _k4_sets(edges)
        assert sk.count_4cliques([], ctx) == 0
        assert sk.count_4cliques([(3, 3), (4, 4)], ctx) == 0
    finally:
        ctx.set_clique_test(False, 0)


def test_intern_terms_sort_cases(ctx):
    # crates/strata-gpu/src/sort.rs:623-664 sizes, full 64-bit keys and heavy duplicates
    rng = np.random.default_rng(0xa11ce)
    for n in (0, 1, 2, 3, 5, 7, 8, 100, 1000, 4097, 255, 256, 257, 5000, 300_000):
        for keys in (rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True),
                     rng.integers(0, 40, size=n, dtype=np.uint64)):
            got = sk.intern_terms(keys, ctx)
            _, inv = np.unique(keys, return_inverse=True)
            assert got == inv.tolist(), n


# ----------------------------------------------- pinned error programs

from helpers import ERROR_CASES  # noqa: E402


@pytest.mark.parametrize("case", ERROR_CASES, ids=[c.label for c in ERROR_CASES])
def test_error_cases(ctx, case):
    """The hand-written programs whose outcome is read off the reference's
    raise points (pinned in tests/test_oracle_pins.py): the engine gives the
    same status, slot or predicate, or the same rows, on both paths."""
    for dense in (True, False):
        try:
            db = ctx.run_semi_naive(case.prog, case.edb, dense=dense)
        except sk.EngineError as e:
            assert case.status != "ok", f"engine {e.kind}: {e.info}"
            assert e.kind == ERR[case.status], (e.kind, case.status)
            if case.slot is not None:
                assert e.info["slot"] == case.slot
            if case.pred is not None:
                assert e.info["pred"] == case.pred
            continue
        assert case.status == "ok", f"engine ok, reference {case.status}"
        for name, rows in (case.want or {}).items():
            got = [tuple(g.value if isinstance(g, IntV) else ("s", g.id) for g in t)
                   for t, _ in db.relation(name).rows()]
            assert got == rows, name
