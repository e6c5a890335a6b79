"""GPU: the benchmark configurations, bit-exact against the oracle at sizes
the oracle finishes in seconds, and through size-independent properties at
full size."""
import numpy as np
import pytest

import strata_k as sk
from oracle import oracle as O
from strata_k import workloads as wl
from strata_k.core_ir import INF, TROP, CoreProgram, pred

pytestmark = pytest.mark.gpu


def same_relation(g, r):
    assert len(g) == len(r)
    assert (g.vals == r.vals).all()
    if g.w is not None:
        assert (g.w == r.w).all()
        assert (g.inf == r.inf).all()


def test_min_by_key_long_runs(ctx):
    """⊕ on EDB load (Relation::combine, crates/strata-eval/src/store.rs:34-48):
    runs of equal keys spanning many device chunks, +inf mixed in
    (the min-by-key check of crates/strata-gpu/src/sort.rs:666-691, scaled up)."""
    rng = np.random.default_rng(7)
    n = 3_000_000
    keys = np.concatenate([rng.integers(0, 40, size=n // 2), np.full(n // 2, 17)]).astype(np.uint32)
    keys2 = rng.integers(0, 3, size=n).astype(np.uint32)
    w = rng.integers(-10**12, 10**12, size=n).astype(np.int64)
    inf = (rng.random(n) < 0.3).astype(np.uint8)
    inf[keys == 39] = 1  # a key whose every fact is +inf
    prog = CoreProgram([pred("p", 2, TROP)], [], 1)
    db = ctx.run_semi_naive(prog, bulk={"p": ([keys, keys2], w, inf)})
    rel = db.relation("p")
    ref = O.run(prog, bulk={"p": ([keys, keys2], w, inf)})
    same_relation(rel, ref.relations["p"])
    # independent: numpy min per key with +inf above every finite value
    k = keys.astype(np.int64) * 4 + keys2
    order = np.lexsort((w, inf, k))
    first = np.r_[True, k[order][1:] != k[order][:-1]]
    want_w = np.where(inf[order][first] == 1, 0, w[order][first])
    assert (rel.inf == inf[order][first]).all()
    assert (np.where(rel.inf == 1, 0, rel.w) == want_w).all()


def test_config1_full(ctx):
    """Plain Bool TC, 2,000 nodes / 10,000 edges: the whole relation, bit-exact."""
    w = wl.config1()
    ref = O.run(w.program, bulk=w.bulk, want=["path"])
    db = ctx.run_semi_naive(w.program, bulk=w.bulk, want=["path"])
    same_relation(db.relation("path"), ref.relations["path"])
    assert db.iterations == ref.iterations
    print("config1:", len(db.relation("path")), "tuples,", db.iterations[0], "rounds")


def test_config2_reduced(ctx):
    """Trop cheapest cost on RMAT scale 14, 16 sources: whole relation bit-exact."""
    w = wl.config2(seed=21, scale=14, n_sources=16)
    ref = O.run(w.program, bulk=w.bulk, want=["cost"])
    db = ctx.run_semi_naive(w.program, bulk=w.bulk, want=["cost"])
    same_relation(db.relation("cost"), ref.relations["cost"])
    assert db.iterations == ref.iterations


def test_config2_full_graph_partition(ctx):
    """Full RMAT-20 graph, all 64 sources on the GPU, checked whole:
      * every (S, Z, weight) row against the multi-source Dijkstra checker
        (orc_cost_dijkstra, pinned against hop-aware Bellman-Ford and the
        semi-naive restatement), and nothing else in the relation;
      * the round count equal to the checker's largest fewest-hops count over
        every source and target (seminaive.rs:94-104);
      * the first source's rows against the semi-naive restatement itself run
        with that source alone (cost(S,.) depends on S only)."""
    w = wl.config2()
    run = ctx.prepare(w.program)
    for p, (cols, wt, inf) in w.bulk.items():
        run.add_sym(p, cols, wt, inf)
    run.load()
    run.execute()
    (s, z), cw, _ = run.relation_sym("cost")
    iters = run.iterations()
    run.free()
    k = s.astype(np.uint64) << np.uint64(32) | z.astype(np.uint64)
    assert (k[1:] > k[:-1]).all()  # sorted, unique
    src = w.bulk["src"][0][0]
    assert set(np.unique(s).tolist()) <= set(src.tolist())
    w1 = wl.restrict_sources(w, 1)
    ref = O.run(w1.program, bulk=w1.bulk, want=["cost"])
    r = ref.relations["cost"]
    sel = s == src[0]
    assert sel.sum() == len(r)
    assert (z[sel] == r.vals[:, 1]).all() and (cw[sel] == r.w).all()
    assert iters[0] >= ref.iterations[0]
    (es, ed), ew, _ = w.bulk["edge"]
    W, _, max_hops = O.cost_dijkstra(w.params["n_nodes"], es, ed, ew, src)
    S, Z, WW = O.cost_rows(W, src)
    assert len(S) == len(s)
    assert (S == s).all() and (Z == z).all() and (WW == cw).all()
    assert iters == [max_hops]
    print("config2 full:", len(s), "tuples,", iters[0], "rounds, every source checked")


def test_config3_reduced(ctx):
    """WCOJ triangles on RMAT scale 12 (degree-oriented): tri and ntri bit-exact."""
    w = wl.config3(seed=31, scale=12)
    ref = O.run(w.program, bulk=w.bulk)
    db = ctx.run_semi_naive(w.program, bulk=w.bulk)
    same_relation(db.relation("tri"), ref.relations["tri"])
    assert db.relation("ntri").rows() == [((sk.IntV(len(ref.relations["tri"])),), None)]
    assert db.iterations == ref.iterations == [1, 0]


def test_config3_full_count(ctx):
    """RMAT scale 22: |tri| and ntri equal the oracle's per-edge merge count."""
    w = wl.config3()
    a, b = w.bulk["e"][0]
    run = ctx.prepare(w.program)
    run.add_sym("e", [a, b])
    run.load()
    run.execute()
    ntri = run.relation("ntri").rows()
    n = run.size("tri")
    run.free()
    want = O.tri_count_oriented(a, b, w.params["n_nodes"])
    assert n == want
    assert ntri == [((sk.IntV(want),), None)]
    print("config3 full:", n, "triangles")


def test_relation_beyond_2_26_rows(ctx):
    """Load, ⊕-normalise and read back a relation of ~70M distinct rows (more
    than 2^26, the size past which a clipped launch grid would drop rows), and
    join over it (CSR index over all rows)."""
    from strata_k.core_ir import BOOL, atom, pos, rule, v
    rng = np.random.default_rng(11)
    n = 72_000_000
    a = rng.integers(0, 1 << 20, n, dtype=np.uint32)
    b = rng.integers(0, 1 << 20, n, dtype=np.uint32)
    key = np.unique(a.astype(np.uint64) << np.uint64(32) | b)
    prog = CoreProgram([pred("e", 2, BOOL), pred("f", 1, BOOL), pred("h", 2, BOOL)],
                       [rule(atom("h", v(0), v(1)), [pos("f", v(0)), pos("e", v(0), v(1))], 2)], 1)
    fsel = np.array([7, 99_999, (1 << 20) - 1], dtype=np.uint32)
    run = ctx.prepare(prog)
    run.add_sym("e", [a, b])
    run.add_sym("f", [fsel])
    run.load()
    run.execute()
    (ga, gb), _, _ = run.relation_sym("e")
    (ha, hb), _, _ = run.relation_sym("h")
    run.free()
    assert len(ga) == len(key)
    assert (ga == (key >> np.uint64(32)).astype(np.uint32)).all() and (gb == (key & np.uint64(0xffffffff))).all()
    want = key[np.isin((key >> np.uint64(32)).astype(np.uint32), fsel)]
    assert (ha.astype(np.uint64) << np.uint64(32) | hb).tolist() == want.tolist()
