"""Repeated executes of one prepared run (reset between them), as a caller
that re-runs a loaded program does: every execute's relations and round
counts equal the oracle's.  With an aggregate in the program every relation
is snapshotted; relations no rule writes keep their device state across a
reset unless the dictionary grew under them (new aggregate Ints sorting
before existing ones re-encode every relation), when they are restored."""
import random

import pytest

from oracle import oracle as O
from strata_k.core_ir import BOOL, Agg, CoreProgram, IntV, Sym, atom, pos, pred, rule, v

pytestmark = pytest.mark.gpu


def _program():
    # outdeg(X, count<Y>) :- edge(X,Y).   path(X,Y) :- edge(X,Y).
    # path(X,Z) :- path(X,Y), edge(Y,Z).  big(X, V) is an EDB relation of Ints.
    return CoreProgram([pred("edge", 2, BOOL, 0), pred("big", 2, BOOL, 0), pred("path", 2, BOOL, 0),
                        pred("outdeg", 2, BOOL, 1)],
                       [rule(atom("path", v(0), v(1)), [pos("edge", v(0), v(1))], 2, 0),
                        rule(atom("path", v(0), v(2)), [pos("path", v(0), v(1)), pos("edge", v(1), v(2))], 3, 0),
                        rule(atom("outdeg", v(0), Agg("count", 1)), [pos("edge", v(0), v(1))], 2, 1)], 2)


def _rows(rel):
    """The engine's rows as the oracle lists them (Bool relations: no weights)."""
    return [tuple(("sym", g.id) if isinstance(g, Sym) else ("int", g.value) for g in t) for t, _ in rel.rows()]


@pytest.mark.parametrize("big_ints", [False, True])
def test_reset_reexecute_equals_oracle(ctx, big_ints):
    rnd = random.Random(7 if big_ints else 5)
    n = 300
    edges = sorted({(rnd.randrange(n), rnd.randrange(n)) for _ in range(1200)})
    facts = [("edge", (Sym(a), Sym(b)), None) for a, b in edges]
    # Ints above every count: the counts sort before them, so the dictionary
    # shifts and every relation is re-encoded (and must be restored)
    base = 10 ** 9 if big_ints else 0
    facts += [("big", (Sym(i), IntV(base + i)), None) for i in range(0, n, 3)]
    prog = _program()
    ref = O.run(prog, facts)
    assert ref.ok
    run = ctx.prepare(prog)
    try:
        run.add_facts(facts)
        run.load()
        for rep in range(3):
            if rep:
                run.reset()
            run.execute()
            assert run.iterations() == ref.iterations, f"execute {rep}"
            for p in ("edge", "big", "path", "outdeg"):
                got = _rows(run.relation(p))
                want = [t for t, _ in ref.relations[p].rows()]
                assert got == want, f"execute {rep}: relation {p} differs ({len(got)} vs {len(want)})"
    finally:
        run.free()
