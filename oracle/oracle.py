"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (built from sk_oracle.c by oracle/Makefile) and
runs a Core-IR program through the C restatement of the reference interpreter
(crates/strata-eval/src/seminaive.rs, naive.rs).  Only tests/, the smoke check
and bench.py's cpu_baseline leg use this module; the engine never does.

Programs are duck-typed: any object with the attributes of the Core-IR model
(predicates / rules / num_strata, CoreAtom.pred / args, Var.slot, ...) works.

Also here: ``asp_simplify_filter``, a plain-Python restatement of the
reference's CPU grounding-simplification filter (crates/strata-gpu/src/asp.rs:53-99),
the checker of the device filter.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORC_SYM, ORC_INT = 0, 1
T_VAR, T_CONST, T_INT, T_AGG = 0, 1, 2, 3
AGG_OPS = ("count", "sum", "min", "max", "prob_or")

STATUS = {
    0: "ok",
    1: "unknown_pred",
    2: "overflow",
    3: "unsafe_head_var",
    4: "unsafe_neg_var",
    5: "negation_not_stratified",
    6: "negative_weight_cycle",
    7: "aggregate_in_recursion",
    8: "malformed_aggregate_head",
    9: "non_integer_aggregand",
    10: "unsupported",
    50: "panic",
    60: "bad_input",
}


class Val(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("v", C.c_int64)]


class Ann(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("inf", C.c_uint8), ("pad", C.c_uint8 * 6), ("w", C.c_int64)]


class Term(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("a", C.c_uint32), ("b", C.c_int64)]


class Atom(C.Structure):
    _fields_ = [("pred", C.c_char_p), ("negated", C.c_uint32), ("first_term", C.c_uint32),
                ("n_terms", C.c_uint32), ("pad", C.c_uint32)]


class Rule(C.Structure):
    _fields_ = [("head", C.c_uint32), ("first_lit", C.c_uint32), ("n_lits", C.c_uint32),
                ("stratum", C.c_uint32), ("var_count", C.c_uint32), ("pad", C.c_uint32)]


class Pred(C.Structure):
    _fields_ = [("name", C.c_char_p), ("arity", C.c_uint32), ("semiring", C.c_uint32),
                ("stratum", C.c_uint32), ("pad", C.c_uint32)]


class Program(C.Structure):
    _fields_ = [("preds", C.POINTER(Pred)), ("n_preds", C.c_uint32),
                ("rules", C.POINTER(Rule)), ("n_rules", C.c_uint32),
                ("atoms", C.POINTER(Atom)), ("n_atoms", C.c_uint32),
                ("terms", C.POINTER(Term)), ("n_terms", C.c_uint32),
                ("num_strata", C.c_uint32)]


class Error(C.Structure):
    _fields_ = [("code", C.c_int32), ("slot", C.c_uint32), ("lhs", C.c_int64), ("rhs", C.c_int64),
                ("pred", C.c_char * 128), ("msg", C.c_char * 256)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_db_new.restype = C.c_void_p
        L.orc_db_new.argtypes = [C.POINTER(Program)]
        L.orc_db_free.argtypes = [C.c_void_p]
        L.orc_insert.restype = C.c_int32
        L.orc_insert.argtypes = [C.c_void_p, C.c_char_p, C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(Error)]
        L.orc_insert_sym.restype = C.c_int32
        L.orc_insert_sym.argtypes = [C.c_void_p, C.c_char_p, C.c_uint64, C.c_uint32, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(Error)]
        L.orc_run.restype = C.c_int32
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(Error)]
        L.orc_iterations.restype = C.c_uint64
        L.orc_iterations.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_universe.restype = C.c_uint64
        L.orc_universe.argtypes = [C.c_void_p]
        L.orc_unify_steps.restype = C.c_uint64
        L.orc_unify_steps.argtypes = [C.c_void_p]
        L.orc_rel_size.restype = C.c_int64
        L.orc_rel_size.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_rel_arity.restype = C.c_int64
        L.orc_rel_arity.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_rel_copy.restype = C.c_int32
        L.orc_rel_copy.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p]
        L.orc_cost_dijkstra.restype = C.c_int32
        L.orc_cost_dijkstra.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
        L.orc_tri_count_oriented.restype = C.c_uint64
        L.orc_tri_count_oriented.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64]
        _lib = L
    return _lib


def _sem_code(s) -> int:
    return 1 if str(s).lower() == "trop" else 0


def _term(t) -> Term:
    name = type(t).__name__
    if name == "Var":
        return Term(T_VAR, t.slot, 0)
    if name == "Const":
        return Term(T_CONST, t.sym, 0)
    if name == "Int":
        return Term(T_INT, 0, t.value)
    if name == "Agg":
        return Term(T_AGG, AGG_OPS.index(t.op), t.slot)
    raise TypeError(f"unsupported term {t!r}")


class FlatProgram:
    """Core-IR flattened into the C layout; keeps the buffers alive."""

    def __init__(self, prog):
        self._names = []
        preds = []
        for p in prog.predicates:
            nm = p.name.encode()
            self._names.append(nm)
            preds.append(Pred(nm, p.arity, _sem_code(p.semiring), p.stratum, 0))
        atoms, terms, rules = [], [], []

        def add_atom(a, negated):
            nm = a.pred.encode()
            self._names.append(nm)
            first = len(terms)
            for t in a.args:
                terms.append(_term(t))
            atoms.append(Atom(nm, 1 if negated else 0, first, len(a.args), 0))
            return len(atoms) - 1

        for r in prog.rules:
            h = add_atom(r.head, False)
            first = len(atoms)
            for lit in r.body:
                add_atom(lit.atom, type(lit).__name__ == "Neg")
            rules.append(Rule(h, first, len(r.body), r.stratum, r.var_count, 0))
        self.preds = (Pred * max(1, len(preds)))(*preds)
        self.rules = (Rule * max(1, len(rules)))(*rules)
        self.atoms = (Atom * max(1, len(atoms)))(*atoms)
        self.terms = (Term * max(1, len(terms)))(*terms)
        self.prog = Program(self.preds, len(preds), self.rules, len(rules), self.atoms, len(atoms),
                            self.terms, len(terms), prog.num_strata)
        self.arity = {p.name: p.arity for p in prog.predicates}


@dataclass
class OracleRelation:
    arity: int
    kinds: np.ndarray   # (n, arity) uint32: 0 Sym, 1 Int
    vals: np.ndarray    # (n, arity) int64
    ann_kind: np.ndarray  # (n,) uint8: 0 Unit, 1 W
    inf: np.ndarray     # (n,) uint8
    w: np.ndarray       # (n,) int64

    def __len__(self):
        return int(self.vals.shape[0])

    def rows(self):
        """Sorted list of (tuple of GroundVal-like, ann) in BTreeMap order."""
        from_kind = lambda k, x: ("sym", int(x)) if k == ORC_SYM else ("int", int(x))
        out = []
        for i in range(len(self)):
            t = tuple(from_kind(self.kinds[i, j], self.vals[i, j]) for j in range(self.arity))
            if self.ann_kind[i] == 0:
                a = None
            elif self.inf[i]:
                a = "inf"
            else:
                a = int(self.w[i])
            out.append((t, a))
        return out

    def sym_columns(self) -> List[np.ndarray]:
        assert (self.kinds == ORC_SYM).all()
        return [self.vals[:, j].astype(np.uint32) for j in range(self.arity)]


@dataclass
class OracleResult:
    status: str
    error: dict
    iterations: List[int]
    universe: int
    steps: int
    seconds: float
    relations: Dict[str, OracleRelation] = field(default_factory=dict)

    @property
    def ok(self) -> bool:
        return self.status == "ok"


def _ann_struct(a) -> Ann:
    x = Ann()
    if a is None:
        x.kind = 0
    elif isinstance(a, str) and a == "inf" or repr(a) == "INF":
        x.kind, x.inf, x.w = 1, 1, 0
    else:
        x.kind, x.w = 1, int(a)
    return x


def _val_struct(g) -> Val:
    name = type(g).__name__
    if name == "Sym":
        return Val(ORC_SYM, 0, g.id)
    if name == "IntV":
        return Val(ORC_INT, 0, g.value)
    if isinstance(g, tuple) and g[0] == "sym":
        return Val(ORC_SYM, 0, g[1])
    if isinstance(g, tuple) and g[0] == "int":
        return Val(ORC_INT, 0, g[1])
    raise TypeError(f"unsupported ground value {g!r}")


def run(prog, edb=(), bulk: Optional[dict] = None, semi: bool = True, indexed: bool = True,
        want: Optional[List[str]] = None) -> OracleResult:
    """Run ``prog`` over the EDB.

    ``edb`` is the reference's ``&[(&str, Tuple, Ann)]`` list; ``bulk`` maps a
    predicate name to ``(columns, weights, inf)`` numpy arrays of Sym ids for
    large inputs.  Facts are inserted in the order given (edb, then bulk).
    """
    L = lib()
    fp = FlatProgram(prog)
    db = L.orc_db_new(C.byref(fp.prog))
    err = Error()
    try:
        for p, t, a in edb:
            row = (Val * max(1, len(t)))(*[_val_struct(g) for g in t])
            ann = _ann_struct(a)
            L.orc_insert(db, p.encode(), 1, C.cast(row, C.c_void_p), C.cast(C.pointer(ann), C.c_void_p),
                         C.byref(err))
        for p, (cols, w, inf) in (bulk or {}).items():
            cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in cols]
            n = len(cols[0]) if cols else 0
            ptrs = (C.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
            wp = None
            if w is not None:
                w = np.ascontiguousarray(w, dtype=np.int64)
                wp = w.ctypes.data
            ip = None
            if inf is not None:
                inf = np.ascontiguousarray(inf, dtype=np.uint8)
                ip = inf.ctypes.data
            L.orc_insert_sym(db, p.encode(), n, len(cols), C.cast(ptrs, C.c_void_p), wp, ip, C.byref(err))
        t0 = time.perf_counter()
        code = L.orc_run(db, 1 if semi else 0, 1 if indexed else 0, C.byref(err))
        dt = time.perf_counter() - t0
        res = OracleResult(
            status=STATUS.get(code, f"code{code}"),
            error=dict(code=code, slot=err.slot, lhs=err.lhs, rhs=err.rhs,
                       pred=err.pred.decode(errors="replace"), msg=err.msg.decode(errors="replace")),
            iterations=[int(L.orc_iterations(db, k)) for k in range(prog.num_strata)],
            universe=int(L.orc_universe(db)),
            steps=int(L.orc_unify_steps(db)),
            seconds=dt,
        )
        if code == 0:
            names = want if want is not None else [p.name for p in prog.predicates]
            for nm in dict.fromkeys(names):
                n = L.orc_rel_size(db, nm.encode())
                ar = L.orc_rel_arity(db, nm.encode())
                vals = (Val * max(1, n * ar))()
                anns = (Ann * max(1, n))()
                L.orc_rel_copy(db, nm.encode(), vals, anns)
                va = np.frombuffer(vals, dtype=np.dtype([("kind", "<u4"), ("pad", "<u4"), ("v", "<i8")]),
                                   count=n * ar)
                an = np.frombuffer(anns, dtype=np.dtype([("kind", "u1"), ("inf", "u1"), ("pad", "u1", 6),
                                                         ("w", "<i8")]), count=n)
                res.relations[nm] = OracleRelation(
                    ar, va["kind"].reshape(n, ar).copy(), va["v"].reshape(n, ar).copy(),
                    an["kind"].copy(), an["inf"].copy(), an["w"].copy())
        return res
    finally:
        L.orc_db_free(db)


def cost_dijkstra(nnodes: int, src, dst, w, sources):
    """Multi-source cheapest cost by Dijkstra over (weight, hops) labels
    (orc_cost_dijkstra): returns (W, H, max_hops) with W[s, z] the least
    weight of a walk of one or more edges from sources[s] to z (INT64_MAX when
    none), H[s, z] the fewest edges among such walks, and max_hops the
    stratum's semi-naive round count."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.int64)
    sources = np.ascontiguousarray(sources, dtype=np.uint32)
    ns = len(sources)
    W = np.empty((ns, nnodes), dtype=np.int64)
    H = np.empty((ns, nnodes), dtype=np.uint32)
    mh = C.c_uint64()
    rc = lib().orc_cost_dijkstra(nnodes, len(src), src.ctypes.data, dst.ctypes.data, w.ctypes.data, ns,
                                 sources.ctypes.data, W.ctypes.data, H.ctypes.data, C.byref(mh))
    if rc:
        raise ValueError(f"orc_cost_dijkstra: {STATUS.get(rc, rc)}")
    return W, H, int(mh.value)


def cost_rows(W, sources):
    """The (S, Z, weight) rows of cost in key order from cost_dijkstra's W."""
    sources = np.asarray(sources, dtype=np.int64)
    order = np.argsort(sources, kind="stable")
    ss, zz, ww = [], [], []
    for i in order:
        z = np.nonzero(W[i] != np.iinfo(np.int64).max)[0]
        ss.append(np.full(len(z), sources[i], np.int64))
        zz.append(z.astype(np.int64))
        ww.append(W[i, z])
    if not ss:
        e = np.zeros(0, np.int64)
        return e, e, e
    return np.concatenate(ss), np.concatenate(zz), np.concatenate(ww)


def tri_count_oriented(a, b, nnodes: int) -> int:
    """Triangles of an oriented simple graph given as sorted unique pairs."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return int(lib().orc_tri_count_oriented(a.ctypes.data, b.ctypes.data, len(a), nnodes))


# ------------------------------------------------------ ASP simplification

def asp_simplify_filter(head, body_start, body_lit, poss, cert):
    """CPU restatement of simplify_filter_cpu (crates/strata-gpu/src/asp.rs:53-99):
    per rule in order, a positive literal over an impossible atom or a negative
    one over a certain atom kills the rule; certain positive and impossible
    negative literals are dropped; kept literals stay in order."""
    oh, ob, ol = [], [0], []
    for i in range(len(head)):
        kept, alive = [], True
        for lit in body_lit[body_start[i]:body_start[i + 1]]:
            lit = int(lit)
            a, positive = abs(lit) - 1, lit > 0
            if positive:
                if not poss[a]:
                    alive = False
                    break
                if not cert[a]:
                    kept.append(lit)
            else:
                if cert[a]:
                    alive = False
                    break
                if poss[a]:
                    kept.append(lit)
        if not alive:
            continue
        oh.append(int(head[i]))
        ol.extend(kept)
        ob.append(len(ol))
    return oh, ob, ol
