"""GPU: relations whose packed tuples need 128-bit keys (arity * bits > 64),
piece by piece: EDB ⊕ load (sort + min), Bool closure (sort + merge),
non-recursive Trop derivations (sort + min into an empty relation), and Trop
derivations merged into a relation that already holds facts."""
import numpy as np
import pytest

from oracle import oracle as O
from strata_k.core_ir import BOOL, TROP, CoreProgram, atom, pos, pred, rule, v

pytestmark = pytest.mark.gpu

SPREAD = 4099  # ids up to ~2^22: 3 columns need more than 64 bits
BIG = 1023 * SPREAD  # one id >= 2^21 in every EDB forces 22-bit columns


def wide(x):
    x = np.asarray(x, dtype=np.uint32).copy()
    x[0] = BIG
    return x


def check(ctx, prog, bulk, outs):
    ref = O.run(prog, bulk=bulk, want=outs)
    assert ref.ok, ref.status
    for dense in (True, False):
        db = ctx.run_semi_naive(prog, bulk=bulk, want=outs, dense=dense)
        for o in outs:
            g, r = db.relation(o), ref.relations[o]
            assert len(g) == len(r), (dense, o, len(g), len(r))
            assert (g.vals == r.vals).all(), (dense, o)
            if g.w is not None:
                bad = np.nonzero(g.w != r.w)[0]
                assert len(bad) == 0, (dense, o, bad[:5], g.w[bad[:5]], r.w[bad[:5]])
        assert db.iterations == ref.iterations, dense


def test_wide_edb_min_load(ctx):
    rng = np.random.default_rng(1)
    n = 400_000
    a = (rng.integers(0, 30, n) * SPREAD).astype(np.uint32)
    b = (rng.integers(0, 20, n) * SPREAD).astype(np.uint32)
    c = (rng.integers(0, 1024, n) * SPREAD).astype(np.uint32)
    w = rng.integers(-1000, 1000, n).astype(np.int64)
    prog = CoreProgram([pred("p", 3, TROP)], [], 1)
    check(ctx, prog, {"p": ([a, b, c], w, None)}, ["p"])


def test_wide_bool_closure(ctx):
    rng = np.random.default_rng(2)
    m = 3000
    q = rng.integers(0, 4, m).astype(np.uint32)
    x = wide(rng.integers(0, 300, m) * SPREAD)
    y = (rng.integers(0, 300, m) * SPREAD).astype(np.uint32)
    prog = CoreProgram([pred("e", 3, BOOL), pred("p", 3, BOOL)],
                       [rule(atom("p", v(0), v(1), v(2)), [pos("e", v(0), v(1), v(2))], 3),
                        rule(atom("p", v(0), v(1), v(3)), [pos("p", v(0), v(1), v(2)), pos("e", v(0), v(2), v(3))],
                             4)], 1)
    check(ctx, prog, {"e": ([q, x, y], None, None)}, ["p"])


def test_wide_trop_nonrecursive(ctx):
    rng = np.random.default_rng(3)
    n = 20_000
    cols = [wide(rng.integers(0, 50, n) * SPREAD) for _ in range(3)]
    w1 = rng.integers(0, 100, n).astype(np.int64)
    w2 = rng.integers(0, 100, n).astype(np.int64)
    prog = CoreProgram([pred("a", 3, TROP), pred("b", 3, TROP), pred("h", 3, TROP)],
                       [rule(atom("h", v(0), v(1), v(2)), [pos("a", v(0), v(1), v(2))], 3),
                        rule(atom("h", v(0), v(1), v(2)), [pos("b", v(0), v(1), v(2))], 3)], 1)
    perm = rng.permutation(n)
    check(ctx, prog, {"a": (cols, w1, None), "b": ([c[perm] for c in cols], w2, None)}, ["h"])


def test_wide_trop_merge_into_facts(ctx):
    rng = np.random.default_rng(4)
    n = 20_000
    cols = [wide(rng.integers(0, 40, n) * SPREAD) for _ in range(3)]
    w1 = rng.integers(0, 100, n).astype(np.int64)
    idx = rng.choice(n, n // 2, replace=False)
    w2 = rng.integers(0, 100, n // 2).astype(np.int64)
    prog = CoreProgram([pred("a", 3, TROP), pred("h", 3, TROP)],
                       [rule(atom("h", v(0), v(1), v(2)), [pos("a", v(0), v(1), v(2))], 3)], 1)
    check(ctx, prog, {"a": (cols, w1, None), "h": ([c[idx] for c in cols], w2, None)}, ["h"])


def test_wide_trop_fresh_propagates(ctx):
    """g copies h; g only sees h's improvements through h's delta, so g == h
    at the fixpoint iff the merge's fresh rows carry every strict improvement."""
    rng = np.random.default_rng(5)
    n = 20_000
    cols = [wide(rng.integers(0, 40, n) * SPREAD) for _ in range(3)]
    w1 = rng.integers(0, 100, n).astype(np.int64)
    idx = rng.choice(n, n // 2, replace=False)
    w2 = rng.integers(0, 100, n // 2).astype(np.int64)
    prog = CoreProgram([pred("a", 3, TROP), pred("h", 3, TROP), pred("g", 3, TROP)],
                       [rule(atom("h", v(0), v(1), v(2)), [pos("a", v(0), v(1), v(2))], 3),
                        rule(atom("g", v(0), v(1), v(2)), [pos("h", v(0), v(1), v(2))], 3)], 1)
    check(ctx, prog, {"a": (cols, w1, None), "h": ([c[idx] for c in cols], w2, None)}, ["h", "g"])


def test_wide_trop_join_emit(ctx):
    """One join (64-bit probe relation) emitting 128-bit Trop head keys."""
    rng = np.random.default_rng(6)
    s = wide(rng.integers(0, 500, 6000) * SPREAD)
    d = (rng.integers(0, 500, 6000) * SPREAD).astype(np.uint32)
    w = rng.integers(1, 1000, 6000).astype(np.int64)
    q = rng.integers(0, 6, 40).astype(np.uint32)
    src = (rng.integers(0, 500, 40) * SPREAD).astype(np.uint32)
    prog = CoreProgram([pred("e", 2, TROP), pred("s", 2, TROP), pred("c", 3, TROP)],
                       [rule(atom("c", v(0), v(1), v(2)), [pos("s", v(0), v(1)), pos("e", v(1), v(2))], 3)], 1)
    check(ctx, prog, {"e": ([s, d], w, None), "s": ([q, src], np.zeros(40, np.int64), None)}, ["c"])
