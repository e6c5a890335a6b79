"""GPU: batched evaluation and query serving, each answer bit-exact against
the plain single-program run and the oracle."""
import numpy as np
import pytest

import strata_k as sk
from oracle import oracle as O
from strata_k import workloads as wl
from strata_k.core_ir import INF, Sym, cost_program, tc_program, TROP

pytestmark = pytest.mark.gpu


def oracle_cols(prog, bulk, out):
    b = {k: (v[0], v[1], None) for k, v in bulk.items()}
    r = O.run(prog, bulk=b, want=[out])
    assert r.ok, r.status
    rel = r.relations[out]
    return [rel.vals[:, j] for j in range(rel.arity)], (rel.w if prog.predicates[-1].semiring == TROP else None), \
        r.iterations


def test_batch_matches_plain_and_oracle(ctx):
    progs = wl.batch_programs(seed=44, count=24, n_lo=30, n_hi=600, m_lo=60, m_hi=2500, cost_sources=2)
    b = sk.Batch(ctx)
    for p in progs:
        b.add(p.program, p.bulk)
    b.load()
    b.execute()
    for i, p in enumerate(progs):
        code, info = b.status(i)
        assert code == 0, (i, info)
        cols, w = b.relation_sym(i, p.output)
        ocols, ow, oit = oracle_cols(p.program, p.bulk, p.output)
        assert len(cols[0]) == len(ocols[0]), i
        for c, oc in zip(cols, ocols):
            assert (c == oc).all(), i
        if w is not None:
            assert (w == ow).all(), i
        assert b.iterations(i) == oit, (i, b.iterations(i), oit)
        # the same program alone through the plain engine
        db = ctx.run_semi_naive(p.program, bulk={k: (v[0], v[1], None) for k, v in p.bulk.items()}, want=[p.output])
        assert db.iterations == oit
    b.free()


def test_batch_per_program_source_domains(ctx):
    """Cost programs with 1 to 9 sources each, sources shared between some
    programs and a source without out-edges: the source column is ranked
    within its program (fewer cell bits than the union of all sources)."""
    progs = []
    for i in range(10):
        n = 80 + 37 * i
        s, d = wl.uniform_graph(900 + i, n, 3 * n)
        w = wl.weights(900 + i, len(s))
        k = 1 + (i * 4) % 9
        src = np.sort(wl.sample_distinct(900 + i, n, k)).astype(np.uint32)
        if i % 3 == 0:
            src = np.unique(np.concatenate([src, np.array([n + 5], np.uint32)]))  # no out-edges
        progs.append((s, d, w, src))
    b = sk.Batch(ctx)
    for s, d, w, src in progs:
        b.add(cost_program(), {"edge": ([s, d], w), "src": ([src], np.zeros(len(src), np.int64))})
    b.load()
    b.execute()
    for i, (s, d, w, src) in enumerate(progs):
        code, info = b.status(i)
        assert code == 0, (i, info)
        bulk = {"edge": ([s, d], w), "src": ([src], np.zeros(len(src), np.int64))}
        cols, bw = b.relation_sym(i, "cost")
        ocols, ow, oit = oracle_cols(cost_program(), bulk, "cost")
        assert len(cols[0]) == len(ocols[0]), i
        for c, oc in zip(cols, ocols):
            assert (c == oc).all(), i
        assert (bw == ow).all(), i
        assert b.iterations(i) == oit, (i, b.iterations(i), oit)
    b.free()


def test_batch_isolates_errors(ctx):
    """A program that fails (negative cycle, overflow) stops alone; the others
    are unaffected (each keeps its own cap universe+2, seminaive.rs:94-103)."""
    good = wl.batch_programs(seed=45, count=6, n_lo=20, n_hi=200, m_lo=40, m_hi=600)
    b = sk.Batch(ctx)
    idx = {}
    for p in good:
        b.add(p.program, p.bulk)
    # negative cycle: a->b 1, b->a -2 with a source
    neg = {"edge": ([np.array([0, 1], np.uint32), np.array([1, 0], np.uint32)], np.array([1, -2], np.int64)),
           "src": ([np.array([0], np.uint32)], np.array([0], np.int64))}
    idx["neg"] = b.add(cost_program(), neg)
    ovf = {"edge": ([np.array([0, 1], np.uint32), np.array([1, 2], np.uint32)], np.array([2**63 - 1, 1], np.int64)),
           "src": ([np.array([0], np.uint32)], np.array([0], np.int64))}
    idx["ovf"] = b.add(cost_program(), ovf)
    b.load()
    b.execute()
    assert b.status(idx["neg"])[0] == 6  # NegativeWeightCycle
    code, info = b.status(idx["ovf"])
    assert code == 2 and (info["lhs"], info["rhs"]) == (2**63 - 1, 1)
    for i, p in enumerate(good):
        assert b.status(i)[0] == 0
        cols, w = b.relation_sym(i, p.output)
        ocols, ow, oit = oracle_cols(p.program, p.bulk, p.output)
        assert all((c == oc).all() for c, oc in zip(cols, ocols))
        assert b.iterations(i) == oit
    b.free()


def _batch_run(ctx, progs, graph):
    b = sk.Batch(ctx)
    b.set_graph(graph)
    for prog, bulk in progs:
        b.add(prog, bulk)
    ctx.set_profiling(True)
    b.load()
    b.execute()
    recs = {r["name"] for r in b.kernel_records()}
    ctx.set_profiling(False)
    out = []
    for i, (prog, bulk) in enumerate(progs):
        code, info = b.status(i)
        rel = "cost" if any(p.name == "cost" for p in prog.predicates) else "path"
        if code:
            out.append((code, info["lhs"], info["rhs"]))
            continue
        cols, w = b.relation_sym(i, rel)
        out.append((0, [c.copy() for c in cols], None if w is None else w.copy(), b.iterations(i)))
    b.free()
    return out, recs


def _same_batch_results(a, b):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x[0] == y[0], i
        if x[0]:
            assert x == y, i
            continue
        assert all((c == d).all() for c, d in zip(x[1], y[1])), i
        assert (x[2] is None) == (y[2] is None) and (x[2] is None or (x[2] == y[2]).all()), i
        assert x[3] == y[3], i


def test_batch_trop_graph_rounds_equal_host_rounds(ctx):
    """The batched Trop group's device-driven rounds (a hipGraph of push
    rounds sized on the device, per-program round counts and caps) against
    the host-driven rounds: every status, tuple, weight and per-program round
    count is identical.  One batch runs narrow cells and holds a program whose
    sums pass 2^32 (the rounds leave, the cells widen, the round is finished
    on the host); the other holds a negative cycle and an overflow."""
    base = [(p.program, p.bulk) for p in wl.batch_programs(seed=46, count=24, n_lo=300, n_hi=4000, m_lo=1200,
                                                            m_hi=20000, cost_sources=2)]
    chain = np.arange(60, dtype=np.uint32)
    widen = {"edge": ([chain, chain + 1], np.full(60, 1 << 30, np.int64)),
             "src": ([np.array([0], np.uint32)], np.zeros(1, np.int64))}
    a = base + [(cost_program(), widen)]
    neg = {"edge": ([np.array([0, 1, 1], np.uint32), np.array([1, 0, 2], np.uint32)], np.array([1, -2, 5], np.int64)),
           "src": ([np.array([0], np.uint32)], np.array([0], np.int64))}
    ovf = {"edge": ([np.array([0, 1, 2], np.uint32), np.array([1, 2, 3], np.uint32)],
                    np.array([2**62, 2**62, 2**62], np.int64)),
           "src": ([np.array([0], np.uint32)], np.array([0], np.int64))}
    bb = base + [(cost_program(), neg), (cost_program(), ovf)]
    for progs in (a, bb):
        on, recs_on = _batch_run(ctx, progs, True)
        off, recs_off = _batch_run(ctx, progs, False)
        assert "graph_rounds_trop" in recs_on and "graph_rounds_trop" not in recs_off
        _same_batch_results(on, off)
    assert on[-2][0] == 6 and on[-1][0] == 2  # NegativeWeightCycle, Overflow
    # and the batch's answers against the oracle
    for (prog, bulk), r in zip(a, _batch_run(ctx, a, True)[0]):
        rel = "cost" if any(p.name == "cost" for p in prog.predicates) else "path"
        ocols, ow, oit = oracle_cols(prog, bulk, rel)
        assert r[0] == 0 and all((c == oc).all() for c, oc in zip(r[1], ocols)) and r[3] == oit
        if ow is not None:
            assert (r[2] == ow).all()


def test_batch_add_refuses_semiring_mismatch(ctx):
    """Weights go with Trop predicates only, as a plain run's staging
    requires: weights on a Bool predicate, or none on a Trop one, are
    SEMIRING_MISMATCH at add time, and the batch stays usable."""
    b = sk.Batch(ctx)
    e = [np.array([0, 1], np.uint32), np.array([1, 2], np.uint32)]
    with pytest.raises(sk.EngineError) as ei:
        b.add(tc_program(), {"edge": (e, np.array([1, 1], np.int64))})
    assert ei.value.code == 50
    with pytest.raises(sk.EngineError) as ei:
        b.add(cost_program(), {"edge": (e, None), "src": ([np.array([0], np.uint32)], np.zeros(1, np.int64))})
    assert ei.value.code == 50
    i = b.add(tc_program(), {"edge": (e, None)})
    b.load()
    b.execute()
    assert b.status(i)[0] == 0
    cols, _ = b.relation_sym(i, "path")
    assert list(zip(cols[0].tolist(), cols[1].tolist())) == [(0, 1), (0, 2), (1, 2)]
    b.free()


def _server_graph():
    s, d = wl.rmat_graph(51, 12, 60_000)
    w = wl.weights(51, len(s))
    return s, d, w, 1 << 12


def _oracle_answer(s, d, w, q):
    prog = wl.query_program(len(q["targets"]) > 0)
    bulk = {"edge": ([s, d], w, None), "src": ([q["sources"]], np.zeros(len(q["sources"]), np.int64), None)}
    out = "cost"
    if len(q["targets"]):
        bulk["tgt"] = ([q["targets"]], np.zeros(len(q["targets"]), np.int64), None)
        out = "ans"
    r = O.run(prog, bulk=bulk, want=[out, "cost"])
    rel = r.relations[out]
    return rel.vals[:, 0], rel.vals[:, 1], rel.w, r.iterations[0]


def test_wide_trop_keys(ctx):
    """Trop relations whose packed tuples exceed 64 bits (arity 3, ids >= 2^22):
    the fused cost(Q,S,Z) shape the server uses, through the general engine."""
    from strata_k.core_ir import CoreProgram, atom, pos, pred, rule, v
    prog = CoreProgram(
        [pred("e", 2, TROP), pred("s", 2, TROP), pred("c", 3, TROP)],
        [rule(atom("c", v(0), v(1), v(2)), [pos("s", v(0), v(1)), pos("e", v(1), v(2))], 3),
         rule(atom("c", v(0), v(1), v(3)), [pos("c", v(0), v(1), v(2)), pos("e", v(2), v(3))], 4)], 1)
    s, d = wl.rmat_graph(61, 10, 8000)
    scale = 4099  # spread ids up to ~2^22
    w = wl.weights(61, len(s))
    rng = np.random.default_rng(5)
    q = rng.integers(0, 5, size=12).astype(np.uint32)
    src = (rng.integers(0, 1024, size=12) * scale).astype(np.uint32)
    bulk = {"e": ([s * scale, d * scale], w, None), "s": ([q, src], np.zeros(12, np.int64), None)}
    ref = O.run(prog, bulk=bulk, want=["c"])
    db = ctx.run_semi_naive(prog, bulk=bulk, want=["c"])
    g, r = db.relation("c"), ref.relations["c"]
    assert len(g) == len(r)
    assert (g.vals == r.vals).all()
    assert (g.w == r.w).all()
    assert db.iterations == ref.iterations


def test_server_single_queries(ctx):
    s, d, w, n = _server_graph()
    qs = wl.config5_queries(seed=53, n_queries=6, n_nodes=n, max_sources=6, n_targets=16, filter_frac=0.3)
    for q in qs:
        srv = sk.Server(ctx, s, d, w, max_batch=4)
        t = srv.submit(q["sources"], q["targets"])
        while srv.step()[0]:
            pass
        res = srv.poll(t)
        a, b, ww = srv.take(t, res["rows"])
        oa, ob, ow, oit = _oracle_answer(s, d, w, q)
        assert len(a) == len(oa)
        assert (a == oa).all() and (b == ob).all()
        assert (ww == ow).all(), np.nonzero(ww != ow)[0][:10]
        assert res["iterations"] == oit
        srv.free()


def test_server_answers_exact_alone_and_in_mix(ctx):
    s, d, w, n = _server_graph()
    qs = wl.config5_queries(seed=52, n_queries=16, n_nodes=n, max_sources=12, n_targets=16, filter_frac=0.25)
    srv = sk.Server(ctx, s, d, w, max_batch=8)
    tickets = []
    # half arrive up front, the rest while the first ones are running
    for q in qs[:8]:
        tickets.append(srv.submit(q["sources"], q["targets"]))
    srv.step()
    srv.step()
    for q in qs[8:]:
        tickets.append(srv.submit(q["sources"], q["targets"]))
    for _ in range(10_000):
        a, wt = srv.step()
        if a == 0 and wt == 0:
            break
    mix = []
    for q, t in zip(qs, tickets):
        res = srv.poll(t)
        assert res is not None
        a, b, ww = srv.take(t, res["rows"])
        mix.append((a, b, ww, res["iterations"]))
        oa, ob, ow, oit = _oracle_answer(s, d, w, q)
        assert (a == oa).all() and (b == ob).all() and (ww == ow).all()
        assert res["iterations"] == oit
    srv.free()
    # each query alone on a fresh server gives the same answer and round count
    for q, m in zip(qs[:6], mix):
        srv = sk.Server(ctx, s, d, w, max_batch=8)
        t = srv.submit(q["sources"], q["targets"])
        while srv.step()[0]:
            pass
        res = srv.poll(t)
        a, b, ww = srv.take(t, res["rows"])
        assert (a == m[0]).all() and (b == m[1]).all() and (ww == m[2]).all() and res["iterations"] == m[3]
        srv.free()


def test_server_failures_stay_per_query(ctx):
    """A query reaching a negative cycle fails with NegativeWeightCycle and one
    reaching an overflowing path fails with Overflow, while queries running
    beside them still match the oracle (each query has its own cap and
    overflow record)."""
    # nodes 0..5: a normal chain, a negative cycle 10<->11, an overflow path 20->21->22
    src = np.array([0, 1, 2, 3, 10, 11, 20, 21], np.uint32)
    dst = np.array([1, 2, 3, 4, 11, 10, 21, 22], np.uint32)
    w = np.array([5, 7, 1, 2, 1, -3, 2**63 - 1, 1], np.int64)
    srv = sk.Server(ctx, src, dst, w, max_batch=4)
    good = srv.submit(np.array([0, 2], np.uint32))
    negc = srv.submit(np.array([10], np.uint32))
    ovf = srv.submit(np.array([20], np.uint32))
    for _ in range(1000):
        a, wt = srv.step()
        if a == 0 and wt == 0:
            break
    with pytest.raises(sk.EvalError) as e1:
        srv.poll(negc)
    assert e1.value.kind == "NegativeWeightCycle"
    with pytest.raises(sk.EvalError) as e2:
        srv.poll(ovf)
    assert e2.value.kind == "Overflow"
    res = srv.poll(good)
    a, b, ww = srv.take(good, res["rows"])
    q = dict(sources=np.array([0, 2], np.uint32), targets=np.empty(0, np.uint32))
    oa, ob, ow, oit = _oracle_answer(src, dst, w, q)
    assert (a == oa).all() and (b == ob).all() and (ww == ow).all() and res["iterations"] == oit
    # the failed instances' lanes are reset and handed out again
    q = dict(sources=np.array([0, 1, 2, 3, 21], np.uint32), targets=np.empty(0, np.uint32))
    again = srv.submit(q["sources"])
    for _ in range(100):
        a, wt = srv.step()
        if a == 0 and wt == 0:
            break
    res = srv.poll(again)
    a, b, ww = srv.take(again, res["rows"])
    oa, ob, ow, oit = _oracle_answer(src, dst, w, q)
    assert (a == oa).all() and (b == ob).all() and (ww == ow).all() and res["iterations"] == oit
    srv.free()


def _group_bytes(s, d):
    """Bytes the server charges per 64-lane group for this edge list: wide
    cells, touched and maxf bitmaps, row records and the packed-delta
    reserve (2^zbits rows of 64 cells, zbits = max(id bits, 8))."""
    n = int(max(s.max(), d.max())) + 1
    b = 1
    while (1 << b) < n:
        b += 1
    z = max(b, 8)
    gc = 1 << (z + 6)
    return gc * 8 + gc // 4 + (1 << z) * 8 + gc * 4


def _dijkstra_answer(s, d, w, q):
    n = int(max(s.max(), d.max())) + 1
    W, _, mh = O.cost_dijkstra(n, s, d, w, q["sources"])
    S, Z, WW = O.cost_rows(W, q["sources"])
    if len(q["targets"]):
        keep = np.isin(Z, q["targets"])
        S, Z, WW = S[keep], Z[keep], WW[keep]
    return S, Z, WW, mh


def _serve_all(srv, tickets, max_steps=100_000):
    done = {}
    for _ in range(max_steps):
        srv.step()
        for t in tickets:
            if t not in done:
                r = srv.poll(t)
                if r is not None:
                    done[t] = (r, srv.take(t, r["rows"]))
        if len(done) == len(tickets):
            return done
    raise AssertionError("queries did not finish")


def test_server_sources_beyond_lanes(ctx):
    """A server of one 64-lane group: queries of 150 and 300 sources (the
    second filtered to 64 targets) run in pieces as lanes free up, beside
    small queries arriving with them; every answer equals the Dijkstra
    checker's and every round count its largest fewest-hops count."""
    s, d, w, n = _server_graph()
    nn = int(max(s.max(), d.max())) + 1
    srv = sk.Server(ctx, s, d, w, budget_bytes=_group_bytes(s, d) + 4096, max_batch=4)
    assert srv.stats()["lanes"] == 64
    rng = np.random.default_rng(77)
    qs = [dict(sources=np.sort(rng.choice(nn, 150, replace=False)).astype(np.uint32),
               targets=np.empty(0, np.uint32)),
          dict(sources=np.sort(rng.choice(nn, 300, replace=False)).astype(np.uint32),
               targets=np.sort(rng.choice(nn, 64, replace=False)).astype(np.uint32))]
    qs += wl.config5_queries(seed=78, n_queries=5, n_nodes=nn, max_sources=9, n_targets=16, filter_frac=0.4)
    tickets = [srv.submit(q["sources"], q["targets"]) for q in qs]
    done = _serve_all(srv, tickets)
    for q, t in zip(qs, tickets):
        r, (gs, gd, gw) = done[t]
        S, Z, WW, mh = _dijkstra_answer(s, d, w, q)
        assert len(gs) == len(S)
        assert (gs == S).all() and (gd == Z).all() and (gw == WW).all()
        assert r["iterations"] == mh
    st = srv.stats()
    srv.free()
    print("sources beyond lanes:", st)


def test_server_streams_sources(ctx):
    """Answers taken source by source while the query runs (the server
    withholds lanes while an untaken answer is held): in source order, each
    equal to the checker's rows for that source."""
    s, d, w, n = _server_graph()
    nn = int(max(s.max(), d.max())) + 1
    srv = sk.Server(ctx, s, d, w, budget_bytes=_group_bytes(s, d) + 4096, max_batch=2)
    srv.set_hold(1)
    rng = np.random.default_rng(79)
    q = dict(sources=np.sort(rng.choice(nn, 200, replace=False)).astype(np.uint32), targets=np.empty(0, np.uint32))
    t = srv.submit(q["sources"])
    S, Z, WW, mh = _dijkstra_answer(s, d, w, q)
    got = 0
    for _ in range(100_000):
        srv.step()
        fin = False
        while True:
            p = srv.poll_source(t)
            if p is None:
                break
            if p["index"] == len(q["sources"]):
                fin = True
                break
            assert p["index"] == got and p["source"] == q["sources"][got]
            z, ww = srv.take_source(t, p["rows"])
            keep = S == p["source"]
            assert (z == Z[keep]).all() and (ww == WW[keep]).all(), got
            got += 1
        if fin:
            break
    assert got == len(q["sources"])
    assert srv.poll(t)["iterations"] == mh
    srv.free()


def test_server_pull_rounds(ctx):
    """64 sources fill one group; rounds whose push join would emit at least
    |E| rows pull all 64 lanes in one pass over the in-edges.  Answers equal
    the checker's and pull rounds ran."""
    s, d, w, n = _server_graph()
    nn = int(max(s.max(), d.max())) + 1
    srv = sk.Server(ctx, s, d, w, max_batch=4)
    rng = np.random.default_rng(80)
    qs = [dict(sources=np.sort(rng.choice(nn, 64, replace=False)).astype(np.uint32), targets=np.empty(0, np.uint32)),
          dict(sources=np.sort(rng.choice(nn, 40, replace=False)).astype(np.uint32),
               targets=np.sort(rng.choice(nn, 64, replace=False)).astype(np.uint32))]
    tickets = [srv.submit(q["sources"], q["targets"]) for q in qs]
    done = _serve_all(srv, tickets)
    for q, t in zip(qs, tickets):
        r, (gs, gd, gw) = done[t]
        S, Z, WW, mh = _dijkstra_answer(s, d, w, q)
        assert (gs == S).all() and (gd == Z).all() and (gw == WW).all()
        assert r["iterations"] == mh
    st = srv.stats()
    srv.free()
    assert st["narrow"] == 1 and st["pull_rounds"] > 0, st


def test_server_widens_past_2_26(ctx):
    """Edge weights below 2^26 start narrow cells with pull rounds; path sums
    pass 2^26 mid-run, the cells widen in place and the round is applied
    again: answers (and those of a query after the widening) equal the
    checker's."""
    rng = np.random.default_rng(81)
    n = 3000
    cs = np.arange(n - 1, dtype=np.uint32)
    cw = (1 << 24) + rng.integers(0, 1000, n - 1)  # a chain whose sums pass 2^26 after 4 hops
    # sparse shortcuts of weight >= 2^23: far nodes stay many hops away, so
    # the sums of the chain's later rounds reach 2^26
    xs = rng.integers(0, n, 3_000).astype(np.uint32)
    xd = rng.integers(0, n, 3_000).astype(np.uint32)
    xw = rng.integers(1 << 23, 1 << 24, 3_000)
    s = np.concatenate([cs, xs]).astype(np.uint32)
    d = np.concatenate([cs + 1, xd]).astype(np.uint32)
    w = np.concatenate([cw, xw]).astype(np.int64)
    srv = sk.Server(ctx, s, d, w, max_batch=4)
    assert srv.stats()["narrow"] == 1
    qs = [dict(sources=np.sort(rng.choice(n, 64, replace=False)).astype(np.uint32), targets=np.empty(0, np.uint32)),
          dict(sources=np.array([0, 7], np.uint32), targets=np.empty(0, np.uint32))]
    t0 = srv.submit(qs[0]["sources"])
    d0 = _serve_all(srv, [t0])
    assert srv.stats()["narrow"] == 0
    t1 = srv.submit(qs[1]["sources"])
    d1 = _serve_all(srv, [t1])
    for q, (r, (gs, gd, gw)) in zip(qs, [d0[t0], d1[t1]]):
        S, Z, WW, mh = _dijkstra_answer(s, d, w, q)
        assert (gs == S).all() and (gd == Z).all() and (gw == WW).all()
        assert r["iterations"] == mh
    srv.free()


# ------------------------------------------------ full-shape configurations

NONE64 = np.iinfo(np.int64).max


def test_config4_full_shape(ctx):
    """The batched configuration at its recorded shape: 512 programs (256
    Bool TC with n <= 2000, m <= 10,000; 256 Trop cost, n in [1e3, 2e5], m in
    [4e3, 2e6], one source), in one batch.  Every program equals its own plain
    run (rows, weights, round counts); every cost program equals the
    multi-source Dijkstra checker (weights and round count = largest fewest
    hops); every TC program's closure and round count equal unit-weight
    Dijkstra from every node (reachability, longest shortest path)."""
    progs = wl.batch_programs(seed=4, count=512, cost_sources=1, tc_node_cap=2000, tc_edge_cap=10_000)
    b = sk.Batch(ctx)
    for p in progs:
        b.add(p.program, p.bulk)
    b.load()
    b.execute()
    n_rows = 0
    for i, p in enumerate(progs):
        assert b.status(i)[0] == 0, i
        cols, w = b.relation_sym(i, p.output)
        its = b.iterations(i)
        n_rows += len(cols[0])
        run = ctx.prepare(p.program)
        for pred, (pc, pw) in p.bulk.items():
            run.add_sym(pred, pc, pw, None)
        run.load()
        run.execute()
        (pcols, pw, _), pits = run.relation_sym(p.output), run.iterations()
        run.free()
        assert its == pits, i
        assert all((x == y).all() for x, y in zip(cols, pcols)), i
        (es, ed), ew = p.bulk["edge"][0], p.bulk["edge"][1]
        n = int(max(es.max(), ed.max())) + 1 if len(es) else 1
        if p.output == "cost":
            assert (w == pw).all(), i
            srcs = p.bulk["src"][0][0]
            W, _, mh = O.cost_dijkstra(n, es, ed, ew, srcs)
            S, Z, WW = O.cost_rows(W, srcs)
            assert (cols[0] == S).all() and (cols[1] == Z).all() and (w == WW).all(), i
            assert its == [mh], i
        else:
            W, _, mh = O.cost_dijkstra(n, es, ed, np.ones(len(es), np.int64), np.arange(n))
            S, Z, _ = O.cost_rows(W, np.arange(n))
            assert (cols[0] == S).all() and (cols[1] == Z).all(), i
            assert its == [mh], i
    b.free()
    print("config4 full shape:", len(progs), "programs,", n_rows, "derived tuples, all exact")


def test_config5_rmat22_answers(ctx):
    """The served configuration on an RMAT scale-22 graph (4,194,304 nodes,
    65M edges) with queries of up to 6 sources, a third filtered to 64
    targets, all submitted at once and run together: every answer's rows and
    weights equal the multi-source Dijkstra checker's (restricted to the
    targets when filtered), and its round count the largest fewest-hops count
    over its sources."""
    s, d, w, n = wl.config5_graph(seed=5, scale=22)
    qs = wl.config5_queries(seed=9, n_queries=10, n_nodes=n, max_sources=6, n_targets=64, filter_frac=0.3,
                            rate_per_s=1e9)
    srv = sk.Server(ctx, s, d, w, max_batch=16)
    tickets = [srv.submit(q["sources"], q["targets"]) for q in qs]
    done = {}
    for _ in range(100_000):
        srv.step()
        for i, tk in enumerate(tickets):
            if i not in done:
                r = srv.poll(tk)
                if r is not None:
                    done[i] = (r, srv.take(tk, r["rows"]))
        if len(done) == len(qs):
            break
    srv.free()
    assert len(done) == len(qs)
    for i, q in enumerate(qs):
        r, (gs, gd, gw) = done[i]
        W, _, mh = O.cost_dijkstra(n, s, d, w, q["sources"])
        S, Z, WW = O.cost_rows(W, q["sources"])
        if len(q["targets"]):
            keep = np.isin(Z, q["targets"])
            S, Z, WW = S[keep], Z[keep], WW[keep]
        assert len(gs) == len(S), i
        assert (gs == S).all() and (gd == Z).all() and (gw == WW).all(), i
        assert r["iterations"] == mh, i
    print("config5 rmat22:", sum(r["rows"] for r, _ in done.values()), "answer rows, all exact")
