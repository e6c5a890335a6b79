"""Core-IR data model: the program that goes into the engine.

Mirrors the reference's Core-IR ADTs (crates/strata-ir/src/core/mod.rs:19-94),
ground values (crates/strata-ir/src/value.rs:19-24), the tropical weight with a
distinct +inf (crates/strata-ir/src/trop.rs:35-62) and the interpreter's
annotation (crates/strata-eval/src/value.rs:16-20).  Names, fields and meaning
follow the reference so a caller of ``strata_eval::run_semi_naive`` finds the
same vocabulary.

Annotations on the Python side:
  * ``None``  -> ``Ann::Unit`` (Bool fact),
  * ``int``   -> ``Ann::W(Weight::Finite(i))``,
  * ``INF``   -> ``Ann::W(Weight::PosInf)``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Union

BOOL = "bool"
TROP = "trop"
SEMIRINGS = (BOOL, TROP)

# AggOp (crates/strata-ir/src/high/program.rs:153-159), serialised order.
AGG_OPS = ("count", "sum", "min", "max", "prob_or")


class _PosInf:
    """``Weight::PosInf``: the ⊕ identity and ⊗ annihilator; above every i64."""

    _inst = None

    def __new__(cls):
        if cls._inst is None:
            cls._inst = super().__new__(cls)
        return cls._inst

    def __repr__(self) -> str:
        return "INF"

    def __reduce__(self):
        return (_PosInf, ())


INF = _PosInf()


# ----------------------------------------------------------------- ground values

@dataclass(frozen=True, order=True)
class Sym:
    """``GroundVal::Sym(SymbolId)`` — an interned constant (u32 id)."""

    id: int


@dataclass(frozen=True, order=True)
class IntV:
    """``GroundVal::Int(i64)``."""

    value: int


GroundVal = Union[Sym, IntV]


# ------------------------------------------------------------------------- terms

@dataclass(frozen=True)
class Var:
    """``CoreTerm::Var { slot }`` — canonical variable slot in its rule."""

    slot: int


@dataclass(frozen=True)
class Const:
    """``CoreTerm::Const { sym }`` — interned constant."""

    sym: int


@dataclass(frozen=True)
class Int:
    """``CoreTerm::Int { value }``."""

    value: int


@dataclass(frozen=True)
class Agg:
    """``CoreTerm::Agg { op, slot }`` — head-only aggregate over a slot."""

    op: str
    slot: int


CoreTerm = Union[Var, Const, Int, Agg]


@dataclass
class CoreAtom:
    pred: str
    args: List[CoreTerm]


@dataclass
class Pos:
    """``CoreLiteral::Pos``."""

    atom: CoreAtom


@dataclass
class Neg:
    """``CoreLiteral::Neg`` — resolved against a strictly lower stratum."""

    atom: CoreAtom


CoreLiteral = Union[Pos, Neg]


@dataclass
class CoreRule:
    head: CoreAtom
    body: List[CoreLiteral]
    stratum: int
    var_count: int
    neg_weight_cycle_check: bool = False


@dataclass
class CorePred:
    name: str
    arity: int
    semiring: str
    stratum: int


@dataclass
class CoreProgram:
    predicates: List[CorePred]
    rules: List[CoreRule]
    num_strata: int

    def rules_in_stratum(self, k: int):
        """Rules of stratum ``k`` in declaration order (core/mod.rs:36-38)."""
        return [r for r in self.rules if r.stratum == k]


# -------------------------------------------------------------- small builders

def v(slot: int) -> Var:
    return Var(slot)


def atom(pred: str, *args: CoreTerm) -> CoreAtom:
    return CoreAtom(pred, list(args))


def pos(pred: str, *args: CoreTerm) -> Pos:
    return Pos(CoreAtom(pred, list(args)))


def neg(pred: str, *args: CoreTerm) -> Neg:
    return Neg(CoreAtom(pred, list(args)))


def rule(head: CoreAtom, body: Sequence[CoreLiteral], var_count: int, stratum: int = 0) -> CoreRule:
    return CoreRule(head, list(body), stratum, var_count)


def pred(name: str, arity: int, semiring: str, stratum: int = 0) -> CorePred:
    assert semiring in SEMIRINGS
    return CorePred(name, arity, semiring, stratum)


# ------------------------------------------------------------ standard programs

def tc_program(sem: str = BOOL, edge: str = "edge", path: str = "path") -> CoreProgram:
    """``path(X,Y) :- edge(X,Y).  path(X,Z) :- edge(X,Y), path(Y,Z).``

    The left-linear closure every reference GPU test and example runs
    (crates/strata-gpu/examples/race.rs:57-99, strata-eval tests/interp.rs:60-79).
    """
    return CoreProgram(
        [pred(edge, 2, sem), pred(path, 2, sem)],
        [
            rule(atom(path, v(0), v(1)), [pos(edge, v(0), v(1))], 2),
            rule(atom(path, v(0), v(2)), [pos(edge, v(0), v(1)), pos(path, v(1), v(2))], 3),
        ],
        1,
    )


def cost_program() -> CoreProgram:
    """Multi-source cheapest cost, right-linear, all Trop.

    ``cost(S,Y) :- src(S), edge(S,Y).  cost(S,Z) :- cost(S,Y), edge(Y,Z).``
    ``src`` is declared Trop with weight-0 facts: a Bool literal under a Trop
    head panics in the reference interpreter (crates/strata-eval/src/value.rs:56-62),
    and weight 0 is how a Bool ``true`` enters Trop (crates/strata-ir/src/trop.rs:44-45).
    """
    return CoreProgram(
        [pred("src", 1, TROP), pred("edge", 2, TROP), pred("cost", 2, TROP)],
        [
            rule(atom("cost", v(0), v(1)), [pos("src", v(0)), pos("edge", v(0), v(1))], 2),
            rule(atom("cost", v(0), v(2)), [pos("cost", v(0), v(1)), pos("edge", v(1), v(2))], 3),
        ],
        1,
    )


def triangle_program() -> CoreProgram:
    """``tri(X,Y,Z) :- e(X,Y), e(Y,Z), e(X,Z).  ntri(count<X>) :- tri(X,_Y,_Z).``"""
    return CoreProgram(
        [pred("e", 2, BOOL, 0), pred("tri", 3, BOOL, 0), pred("ntri", 1, BOOL, 1)],
        [
            rule(atom("tri", v(0), v(1), v(2)),
                 [pos("e", v(0), v(1)), pos("e", v(1), v(2)), pos("e", v(0), v(2))], 3, 0),
            rule(atom("ntri", Agg("count", 0)), [pos("tri", v(0), v(1), v(2))], 3, 1),
        ],
        2,
    )
