"""ctypes binding of libstrata_k.so (the gfx950 semi-naive engine).

Mirrors the reference's host API for this path:

  * ``run_semi_naive(prog, edb)`` -> ``Db``          strata_eval::run_semi_naive
    (crates/strata-eval/src/seminaive.rs:29), raising ``EvalError`` with the
    reference's variant names (crates/strata-eval/src/naive.rs:26-54);
  * ``transitive_closure_bool`` / ``shortest_paths_trop`` / ``count_triangles``
    / ``count_4cliques`` / ``intern_terms`` / ``binary_plan_intermediate``
    strata_gpu (crates/strata-gpu/src/lib.rs:60-149), raising ``GpuError``.

There is no CPU fallback: without the built library or a device every entry
point raises ``GpuError('NotBuilt')``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .core_ir import INF, IntV, Sym

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libstrata_k.so")

SK_OK = 0
STATUS_NAMES = {
    0: "Ok",
    1: "UnknownPred",
    2: "Overflow",
    3: "UnsafeHeadVar",
    4: "UnsafeNegVar",
    5: "NegationNotStratified",
    6: "NegativeWeightCycle",
    7: "AggregateInRecursion",
    8: "MalformedAggregateHead",
    9: "NonIntegerAggregand",
    10: "Unsupported",
    50: "SemiringMismatch",
    100: "NotBuilt",
    101: "Cuda",
    102: "OutOfBudget",
    103: "UnsupportedShape",
    104: "InvalidArgument",
    105: "NotReady",
}
EVAL_ERRORS = set(range(1, 11))
T_VAR, T_CONST, T_INT, T_AGG = 0, 1, 2, 3
AGG_OPS = ("count", "sum", "min", "max", "prob_or")
KERNEL_CLASSES = ("radix_sort", "dedup_min", "merge", "join", "leapfrog", "misc")


class EngineError(Exception):
    """Base of the engine's errors; ``kind`` is the reference variant name."""

    def __init__(self, code: int, info: Optional[dict] = None):
        self.code = code
        self.kind = STATUS_NAMES.get(code, f"code{code}")
        self.info = info or {}
        super().__init__(f"{self.kind}: {self.info.get('message', '')} {self.info.get('pred', '')}".strip())


class EvalError(EngineError):
    """strata_eval::EvalError (crates/strata-eval/src/naive.rs:26-54)."""


class GpuError(EngineError):
    """strata_gpu::GpuError (crates/strata-gpu/src/lib.rs:14-19) and engine limits."""


def _raise(code: int, info: Optional[dict] = None):
    if code in EVAL_ERRORS:
        raise EvalError(code, info)
    raise GpuError(code, info)


# ------------------------------------------------------------------ C layout

class SkTerm(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("a", C.c_uint32), ("b", C.c_int64)]


class SkAtom(C.Structure):
    _fields_ = [("pred", C.c_char_p), ("negated", C.c_uint32), ("first_term", C.c_uint32),
                ("n_terms", C.c_uint32), ("pad", C.c_uint32)]


class SkRule(C.Structure):
    _fields_ = [("head", C.c_uint32), ("first_lit", C.c_uint32), ("n_lits", C.c_uint32),
                ("stratum", C.c_uint32), ("var_count", C.c_uint32), ("pad", C.c_uint32)]


class SkPred(C.Structure):
    _fields_ = [("name", C.c_char_p), ("arity", C.c_uint32), ("semiring", C.c_uint32),
                ("stratum", C.c_uint32), ("pad", C.c_uint32)]


class SkProgram(C.Structure):
    _fields_ = [("preds", C.POINTER(SkPred)), ("n_preds", C.c_uint32),
                ("rules", C.POINTER(SkRule)), ("n_rules", C.c_uint32),
                ("atoms", C.POINTER(SkAtom)), ("n_atoms", C.c_uint32),
                ("terms", C.POINTER(SkTerm)), ("n_terms", C.c_uint32),
                ("num_strata", C.c_uint32)]


class SkVal(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("v", C.c_int64)]


class SkErrorInfo(C.Structure):
    _fields_ = [("code", C.c_int32), ("slot", C.c_uint32), ("lhs", C.c_int64), ("rhs", C.c_int64),
                ("pred", C.c_char * 128), ("message", C.c_char * 256)]


class SkKernelStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("ms", C.c_double), ("bytes", C.c_double), ("items", C.c_uint64)]


_VP = C.c_void_p
_U64 = C.c_uint64
_I64 = C.c_int64
_U32 = C.c_uint32

# (name, restype, argtypes) of every exported function declared in include/strata_k.h
SIGNATURES = [
    ("sk_ctx_create", C.c_int, [C.c_int, _U64, C.POINTER(_VP)]),
    ("sk_ctx_destroy", None, [_VP]),
    ("sk_ctx_synchronize", C.c_int, [_VP]),
    ("sk_ctx_error", None, [_VP, C.POINTER(SkErrorInfo)]),
    ("sk_ctx_device_bytes", _U64, [_VP]),
    ("sk_ctx_device_peak", _U64, [_VP]),
    ("sk_device_available", C.c_int, []),
    ("sk_build_info", C.c_char_p, []),
    ("sk_prepare", C.c_int, [_VP, C.POINTER(SkProgram), C.POINTER(_VP)]),
    ("sk_run_free", None, [_VP]),
    ("sk_run_add_facts", C.c_int, [_VP, C.c_char_p, _U64, _VP, _VP, _VP]),
    ("sk_run_add_facts_sym", C.c_int, [_VP, C.c_char_p, _U64, _U32, _VP, _VP, _VP]),
    ("sk_run_load", C.c_int, [_VP]),
    ("sk_run_execute", C.c_int, [_VP]),
    ("sk_run_reset", C.c_int, [_VP]),
    ("sk_run_set_dense", C.c_int, [_VP, C.c_int]),
    ("sk_run_set_pull", C.c_int, [_VP, C.c_int]),
    ("sk_run_set_graph", C.c_int, [_VP, C.c_int]),
    ("sk_run_set_pack", C.c_int, [_VP, C.c_int]),
    ("sk_run_error", None, [_VP, C.POINTER(SkErrorInfo)]),
    ("sk_result_size", _I64, [_VP, C.c_char_p]),
    ("sk_result_arity", _I64, [_VP, C.c_char_p]),
    ("sk_result_read", C.c_int, [_VP, C.c_char_p, _VP, _VP, _VP, _VP]),
    ("sk_result_read_sym", C.c_int, [_VP, C.c_char_p, _VP, _VP, _VP]),
    ("sk_result_iterations", _U64, [_VP, _U32]),
    ("sk_result_universe", _U64, [_VP]),
    ("sk_run_stats", C.c_int, [_VP, _U32, C.POINTER(SkKernelStats)]),
    ("sk_run_kernel_count", _U32, [_VP]),
    ("sk_run_kernel_record", C.c_int, [_VP, _U32, C.c_char_p, C.c_size_t, C.POINTER(_U32),
                                       C.POINTER(SkKernelStats)]),
    ("sk_set_profiling", None, [_VP, C.c_int]),
    ("sk_ctx_kernel_reset", C.c_int, [_VP]),
    ("sk_ctx_kernel_count", _U32, [_VP]),
    ("sk_ctx_kernel_record", C.c_int, [_VP, _U32, C.c_char_p, C.c_size_t, C.POINTER(_U32),
                                       C.POINTER(SkKernelStats)]),
    ("sk_run_execute_ms", C.c_double, [_VP]),
    ("sk_run_semi_naive", C.c_int, [_VP, C.POINTER(SkProgram), _U64, _VP, _VP, _VP, _VP, _VP, _VP,
                                    C.POINTER(_VP)]),
    ("sk_free", None, [_VP]),
    ("sk_transitive_closure_bool", C.c_int, [_VP, _VP, _VP, _U64, C.POINTER(_VP), C.POINTER(_VP),
                                             C.POINTER(_U64)]),
    ("sk_shortest_paths_trop", C.c_int, [_VP, _VP, _VP, _VP, _U64, C.POINTER(_VP), C.POINTER(_VP),
                                         C.POINTER(_VP), C.POINTER(_U64)]),
    ("sk_count_triangles", C.c_int, [_VP, _VP, _VP, _U64, C.POINTER(_U64)]),
    ("sk_count_4cliques", C.c_int, [_VP, _VP, _VP, _U64, C.POINTER(_U64)]),
    ("sk_ctx_set_clique_test", C.c_int, [_VP, C.c_int, _U32]),
    ("sk_intern_terms", C.c_int, [_VP, _VP, _U64, _VP]),
    ("sk_binary_plan_intermediate", _U64, [_VP, _VP, _U64]),
    ("sk_batch_create", C.c_int, [_VP, C.POINTER(_VP)]),
    ("sk_batch_free", None, [_VP]),
    ("sk_batch_add", C.c_int, [_VP, C.POINTER(SkProgram), _U32, _VP, _VP, _VP, _VP, C.POINTER(_U32)]),
    ("sk_batch_set_graph", C.c_int, [_VP, C.c_int]),
    ("sk_batch_load", C.c_int, [_VP]),
    ("sk_batch_execute", C.c_int, [_VP]),
    ("sk_batch_status", C.c_int, [_VP, _U32, C.POINTER(SkErrorInfo)]),
    ("sk_batch_iterations", _U64, [_VP, _U32, _U32]),
    ("sk_batch_result_size", _I64, [_VP, _U32, C.c_char_p]),
    ("sk_batch_result_read_sym", C.c_int, [_VP, _U32, C.c_char_p, _VP, _VP, _VP]),
    ("sk_batch_execute_ms", C.c_double, [_VP]),
    ("sk_asp_simplify_filter", C.c_int, [_VP, _VP, _VP, _VP, _U64, _VP, _VP, _U64, _VP, _VP, _VP,
                                         C.POINTER(_U64), C.POINTER(_U64)]),
    ("sk_program_from_json", C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(_VP), C.c_char_p, C.c_size_t]),
    ("sk_program_json_get", C.POINTER(SkProgram), [_VP]),
    ("sk_program_json_free", None, [_VP]),
    ("sk_prepare_json", C.c_int, [_VP, C.c_char_p, C.c_size_t, C.POINTER(_VP)]),
    ("sk_asp_simplify_filter_device", C.c_int, [_VP, _VP, _VP, _VP, _U64, _U64, _VP, _VP, _U64, _VP, _VP, _VP,
                                                C.POINTER(_U64), C.POINTER(_U64)]),
    ("sk_asp_simplify_filter_device_noalloc", C.c_int, [_VP, _VP, _VP, _VP, _U64, _U64, _VP, _VP, _U64, _VP, _VP,
                                                        _VP, C.POINTER(_U64), C.POINTER(_U64)]),
    ("sk_sort_u64", C.c_int, [_VP, _VP, _U64, _VP]),
    ("sk_sort_unique_u64", C.c_int, [_VP, _VP, _U64, _VP, C.POINTER(_U64)]),
    ("sk_sort_unique_pairs", C.c_int, [_VP, _VP, _VP, _U64, _VP, _VP, C.POINTER(_U64)]),
    ("sk_sort_min_pairs", C.c_int, [_VP, _VP, _VP, _VP, _U64, _VP, _VP, _VP, C.POINTER(_U64)]),
    ("sk_radix_sort_device", C.c_int, [_VP, _VP, _VP, _VP, _VP, _U64, C.c_int, C.POINTER(C.c_int)]),
    ("sk_ctx_set_autotune", C.c_int, [_VP, C.c_int]),
    ("sk_ctx_force_radix", C.c_int, [_VP, C.c_int]),
    ("sk_ctx_radix_pick", C.c_int, [_VP, C.c_int, C.c_int, _U64, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]),
    ("sk_radix_candidates", C.c_int, []),
    ("sk_ctx_radix_timing", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int]),
    ("sk_radix_candidate", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("sk_ctx_radix_choice", C.c_int, [_VP, C.c_int, C.c_int, _U64, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("sk_batch_kernel_count", _U32, [_VP]),
    ("sk_batch_kernel_record", C.c_int, [_VP, _U32, C.c_char_p, C.c_size_t, C.POINTER(_U32),
                                        C.POINTER(SkKernelStats)]),
    ("sk_server_create", C.c_int, [_VP, _VP, _VP, _VP, _U64, _U64, _U32, C.POINTER(_VP)]),
    ("sk_server_free", None, [_VP]),
    ("sk_server_submit", C.c_int, [_VP, _VP, _U32, _VP, _U32, C.POINTER(_U64)]),
    ("sk_server_step", C.c_int, [_VP, C.POINTER(_U32), C.POINTER(_U32)]),
    ("sk_server_poll", C.c_int, [_VP, _U64, C.POINTER(_U64), C.POINTER(_U64), C.POINTER(C.c_double)]),
    ("sk_server_take", C.c_int, [_VP, _U64, _VP, _VP, _VP]),
    ("sk_server_release", C.c_int, [_VP, _U64]),
    ("sk_server_poll_source", C.c_int, [_VP, _U64, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U64)]),
    ("sk_server_take_source", C.c_int, [_VP, _U64, _VP, _VP]),
    ("sk_server_set_hold", C.c_int, [_VP, _U64]),
    ("sk_server_stats", C.c_int, [_VP, _VP, _U32]),
    # workload generators (host code in the same library)
    ("sk_wl_uniform_graph", _U64, [_U64, _U32, _U64, _VP, _VP]),
    ("sk_wl_tri_hubs", _U64, [_U32, _U32, _U32, _U32, _VP, _VP]),
    ("sk_wl_rmat", _U64, [_U64, _U32, _U64, C.c_double, C.c_double, C.c_double, _U32, _VP, _VP]),
    ("sk_wl_normalize", _U64, [_VP, _VP, _U64, C.c_int, C.c_int]),
    ("sk_wl_weights", None, [_U64, _U64, _I64, _I64, _VP]),
    ("sk_wl_sample_distinct", _U32, [_U64, _U32, _U32, _VP]),
    ("sk_wl_orient_by_degree", None, [_VP, _VP, _U64, _U32]),
    ("sk_wl_log_uniform", _U64, [_U64, _U64, _U64, _U64]),
    ("sk_wl_uniform01", C.c_double, [_U64, _U64]),
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libstrata_k.so and bind every declared function (raises when absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GpuError(100, {"message": f"{path} is not built"})
    L = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _err_info(struct: SkErrorInfo) -> dict:
    return dict(code=struct.code, slot=struct.slot, lhs=struct.lhs, rhs=struct.rhs,
                pred=struct.pred.decode(errors="replace"), message=struct.message.decode(errors="replace"))


# --------------------------------------------------------------- flattening

def _sem(s) -> int:
    return 1 if str(s).lower() == "trop" else 0


def _term(t) -> SkTerm:
    name = type(t).__name__
    if name == "Var":
        return SkTerm(T_VAR, t.slot, 0)
    if name == "Const":
        return SkTerm(T_CONST, t.sym, 0)
    if name == "Int":
        return SkTerm(T_INT, 0, t.value)
    if name == "Agg":
        return SkTerm(T_AGG, AGG_OPS.index(t.op), t.slot)
    raise TypeError(f"unsupported term {t!r}")


class FlatProgram:
    """A Core-IR program flattened into the C ABI's ``sk_program`` layout."""

    def __init__(self, prog):
        self._keep = []
        preds = []
        for p in prog.predicates:
            nm = p.name.encode()
            self._keep.append(nm)
            preds.append(SkPred(nm, p.arity, _sem(p.semiring), p.stratum, 0))
        atoms, terms, rules = [], [], []

        def add_atom(a, negated):
            nm = a.pred.encode()
            self._keep.append(nm)
            first = len(terms)
            terms.extend(_term(t) for t in a.args)
            atoms.append(SkAtom(nm, 1 if negated else 0, first, len(a.args), 0))
            return len(atoms) - 1

        for r in prog.rules:
            h = add_atom(r.head, False)
            first = len(atoms)
            for lit in r.body:
                add_atom(lit.atom, type(lit).__name__ == "Neg")
            rules.append(SkRule(h, first, len(r.body), r.stratum, r.var_count, 0))
        self.preds = (SkPred * max(1, len(preds)))(*preds)
        self.rules = (SkRule * max(1, len(rules)))(*rules)
        self.atoms = (SkAtom * max(1, len(atoms)))(*atoms)
        self.terms = (SkTerm * max(1, len(terms)))(*terms)
        self.c = SkProgram(self.preds, len(preds), self.rules, len(rules), self.atoms, len(atoms),
                           self.terms, len(terms), prog.num_strata)


# ------------------------------------------------------------------ results

@dataclass
class Relation:
    """Relation contents in BTreeMap order (store.rs:17-21): sorted, unique."""

    name: str
    arity: int
    semiring: str
    kinds: Optional[np.ndarray]  # (n, arity) uint32 0 Sym / 1 Int, None when all Sym
    vals: np.ndarray             # (n, arity) int64
    w: Optional[np.ndarray]      # (n,) int64 Trop weights
    inf: Optional[np.ndarray]    # (n,) uint8 +inf flags

    def __len__(self):
        return int(self.vals.shape[0])

    def rows(self):
        """List of (tuple of Sym/IntV, ann) with ann None (Unit), int, or INF."""
        out = []
        for i in range(len(self)):
            t = tuple((Sym(int(self.vals[i, j])) if self.kinds is None or self.kinds[i, j] == 0
                       else IntV(int(self.vals[i, j]))) for j in range(self.arity))
            if self.w is None:
                a = None
            elif self.inf is not None and self.inf[i]:
                a = INF
            else:
                a = int(self.w[i])
            out.append((t, a))
        return out


class Db:
    """Result of a run: one ``Relation`` per declared predicate (store.rs:61-94)."""

    def __init__(self, relations: Dict[str, Relation], iterations: List[int], universe: int,
                 device_ms: float, stats: dict):
        self._rels = relations
        self.iterations = iterations
        self.universe = universe
        self.device_ms = device_ms
        self.stats = stats

    def relation(self, name: str) -> Optional[Relation]:
        return self._rels.get(name)

    def predicates(self):
        return sorted(self._rels)


# ------------------------------------------------------------------ context

def radix_candidates() -> int:
    """Number of radix sort candidate configurations."""
    return int(load_library().sk_radix_candidates())


def radix_candidate(c: int, wide: bool = False) -> dict:
    """Candidate c: digit bits, items per thread (128-bit keys when wide),
    threads per block, and whether its passes are single-sweep."""
    d, it, b, sw = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    code = load_library().sk_radix_candidate(int(c), 1 if wide else 0, C.byref(d), C.byref(it), C.byref(b),
                                             C.byref(sw))
    if code:
        _raise(code)
    return {"digit_bits": d.value, "items_per_thread": it.value, "block": b.value, "sweep": bool(sw.value)}


class Context:
    """One device, one stream, one memory pool (sk_ctx)."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        self.L = load_library()
        if not self.L.sk_device_available():
            raise GpuError(100, {"message": "no HIP device is available"})
        h = C.c_void_p()
        code = self.L.sk_ctx_create(device, budget_bytes, C.byref(h))
        if code:
            _raise(code, {"message": "sk_ctx_create failed"})
        self.h = h
        import weakref
        self._children = weakref.WeakSet()  # runs, batches and servers on this context

    def _adopt(self, child):
        self._children.add(child)

    def close(self):
        if self.h:
            # every object living on the context is released before it
            for ch in list(self._children):
                ch.free()
            self.L.sk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_profiling(self, on: bool):
        self.L.sk_set_profiling(self.h, 1 if on else 0)

    def synchronize(self):
        self.L.sk_ctx_synchronize(self.h)

    @property
    def device_peak_bytes(self) -> int:
        return int(self.L.sk_ctx_device_peak(self.h))

    # ----------------------------------------------------- radix autotuner
    def set_autotune(self, mode: int):
        """0: fixed 8-bit digits; 1 (default): time a size class on its first
        sort; 2: time every class now."""
        code = self.L.sk_ctx_set_autotune(self.h, int(mode))
        if code:
            _raise(code)

    def set_clique_test(self, force_heavy: bool = False, lds_cap: int = 0):
        """4-clique counter testing knob: every edge to the wave path, and
        that path's LDS table capped at lds_cap entries (0: 1024)."""
        code = self.L.sk_ctx_set_clique_test(self.h, 1 if force_heavy else 0, int(lds_cap))
        if code:
            _raise(code)

    def force_radix(self, candidate: int):
        """Pin every sort to one candidate configuration (-1: the tuner)."""
        code = self.L.sk_ctx_force_radix(self.h, int(candidate))
        if code:
            _raise(code)

    def radix_choice(self, n: int, key_bits: int, wide: bool = False, weighted: bool = False) -> dict:
        d, it, est = C.c_int(), C.c_int(), C.c_double()
        code = self.L.sk_ctx_radix_choice(self.h, 1 if wide else 0, 1 if weighted else 0, int(n), int(key_bits),
                                          C.byref(d), C.byref(it), C.byref(est))
        if code:
            _raise(code)
        pick = C.c_int()
        code = self.L.sk_ctx_radix_pick(self.h, 1 if wide else 0, 1 if weighted else 0, int(n), int(key_bits),
                                        C.byref(pick), None)
        if code:
            _raise(code)
        desc = radix_candidate(pick.value, wide)
        return {"candidate": pick.value, "digit_bits": d.value, "items_per_thread": it.value,
                "block": desc["block"], "sweep": desc["sweep"], "est_ms": est.value}

    def radix_timing(self, cls: int, wide: bool = False, weighted: bool = False) -> list:
        """The tuner's per-pass time (ms) of every candidate for size class
        cls (records 2^(16 + 2 cls))."""
        nc = radix_candidates()
        buf = (C.c_float * nc)()
        code = self.L.sk_ctx_radix_timing(self.h, 1 if wide else 0, 1 if weighted else 0, int(cls), buf, nc)
        if code:
            _raise(code)
        return [round(float(v), 4) for v in buf]

    def radix_sort_device(self, keys_ptr: int, alt_ptr: int, n: int, key_bits: int, w_ptr: int = 0,
                          walt_ptr: int = 0) -> bool:
        """Sort device-resident u64 keys (and i64 weights) on the context's
        stream; True when the result is in the alternate buffers."""
        flag = C.c_int()
        code = self.L.sk_radix_sort_device(self.h, keys_ptr, alt_ptr, w_ptr or None, walt_ptr or None, int(n),
                                           int(key_bits), C.byref(flag))
        if code:
            _raise(code)
        return bool(flag.value)

    def asp_simplify_filter_device(self, head_ptr: int, bs_ptr: int, lits_ptr: int, n_rules: int, n_lits: int,
                                   poss_ptr: int, cert_ptr: int, n_atoms: int, out_head_ptr: int, out_bs_ptr: int,
                                   out_lits_ptr: int, noalloc: bool = False):
        """The ASP simplify filter over device-resident buffers; returns the
        kept rule and literal counts once the filter has finished.  noalloc
        takes the entry point that allocates no device memory."""
        nr, nl = C.c_uint64(), C.c_uint64()
        fn = self.L.sk_asp_simplify_filter_device_noalloc if noalloc else self.L.sk_asp_simplify_filter_device
        code = fn(self.h, head_ptr or None, bs_ptr, lits_ptr or None, int(n_rules), int(n_lits), poss_ptr or None,
                  cert_ptr or None, int(n_atoms), out_head_ptr or None, out_bs_ptr, out_lits_ptr or None,
                  C.byref(nr), C.byref(nl))
        if code:
            _raise(code)
        return nr.value, nl.value

    def kernel_reset(self):
        """Clear the context-level kernel records (crate functions, sorts, the
        ASP filter) collected while profiling is on."""
        code = self.L.sk_ctx_kernel_reset(self.h)
        if code:
            _raise(code)

    def kernel_records(self) -> List[dict]:
        return _kernel_records(self.L, self.h, self.L.sk_ctx_kernel_count, self.L.sk_ctx_kernel_record)

    # ------------------------------------------------------------ staged run
    def prepare(self, prog) -> "Run":
        return Run(self, prog)

    def prepare_json(self, json_text) -> "Run":
        """Prepare a serialised CoreProgram (sk_prepare_json)."""
        return Run(self, program_from_json(json_text), json_text=json_text)

    def run_semi_naive(self, prog, edb: Iterable = (), bulk: Optional[dict] = None,
                       want: Optional[Sequence[str]] = None, dense: bool = True, pull: int = 1,
                       graph: bool = True, pack: bool = True) -> Db:
        """``run_semi_naive(prog, edb)`` (crates/strata-eval/src/seminaive.rs:29).

        ``edb``: iterable of ``(pred, tuple-of-Sym/IntV, ann)``; ``bulk``: pred ->
        ``(columns, weights, inf)`` numpy arrays of Sym ids for large inputs.
        ``dense=False`` keeps every relation on the sorted-list path; ``pull``
        selects destination-side rounds (see ``Run.set_pull``), ``graph``
        device-driven rounds (see ``Run.set_graph``), ``pack`` packed pull
        deltas (see ``Run.set_pack``).
        """
        run = self.prepare(prog)
        try:
            run.set_dense(dense)
            run.set_pull(pull)
            run.set_graph(graph)
            run.set_pack(pack)
            run.add_facts(edb)
            for p, (cols, w, inf) in (bulk or {}).items():
                run.add_sym(p, cols, w, inf)
            run.load()
            run.execute()
            return run.db(want)
        finally:
            run.free()


class Run:
    """A prepared program (sk_run): add facts, load, execute, read results.
    ``json_text``: prepare from a serialised CoreProgram instead
    (sk_prepare_json); ``prog`` is then its decoded form."""

    def __init__(self, ctx: Context, prog, json_text: Optional[str] = None):
        self.ctx = ctx
        self.L = ctx.L
        self.prog = prog
        h = C.c_void_p()
        if json_text is not None:
            self.flat = None
            raw = json_text.encode() if isinstance(json_text, str) else bytes(json_text)
            code = self.L.sk_prepare_json(ctx.h, raw, len(raw), C.byref(h))
        else:
            self.flat = FlatProgram(prog)
            code = self.L.sk_prepare(ctx.h, C.byref(self.flat.c), C.byref(h))
        if code:
            info = SkErrorInfo()
            self.L.sk_ctx_error(ctx.h, C.byref(info))
            _raise(code, _err_info(info))
        self.h = h
        ctx._adopt(self)
        self._keep = []

    def free(self):
        if self.h and self.ctx.h:
            self.L.sk_run_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _check(self, code):
        if code:
            info = SkErrorInfo()
            self.L.sk_run_error(self.h, C.byref(info))
            d = _err_info(info)
            _raise(code if d["code"] == 0 else d["code"], d)

    def add_facts(self, edb: Iterable):
        """Stage ``(pred, tuple, ann)`` facts in order (seminaive.rs:30-36)."""
        groups: Dict[Tuple[str, bool], list] = {}
        order = []
        for p, t, a in edb:
            key = (p, a is not None)
            if key not in groups:
                groups[key] = []
                order.append(key)
            groups[key].append((t, a))
        # facts of one predicate keep their relative order; ⊕ makes the
        # result independent of the insertion order anyway
        for (p, weighted) in order:
            facts = groups[(p, weighted)]
            arity = len(facts[0][0]) if facts else 0
            vals = (SkVal * max(1, len(facts) * max(1, arity)))()
            k = 0
            for t, _ in facts:
                for g in t:
                    if type(g).__name__ == "Sym":
                        vals[k] = SkVal(0, 0, g.id)
                    elif type(g).__name__ == "IntV":
                        vals[k] = SkVal(1, 0, g.value)
                    else:
                        raise TypeError(f"unsupported ground value {g!r}")
                    k += 1
            w = inf = None
            if weighted:
                w = np.array([0 if a is INF else int(a) for _, a in facts], dtype=np.int64)
                inf = np.array([1 if a is INF else 0 for _, a in facts], dtype=np.uint8)
            self._check(self.L.sk_run_add_facts(self.h, p.encode(), len(facts), C.cast(vals, C.c_void_p),
                                                 w.ctypes.data if w is not None else None,
                                                 inf.ctypes.data if inf is not None else None))
            self._keep.append((vals, w, inf))

    def add_sym(self, pred: str, cols, w=None, inf=None):
        """Stage all-Sym facts given as column arrays of u32 ids."""
        cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in cols]
        n = len(cols[0]) if cols else 0
        ptrs = (C.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        wp = ip = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.int64)
            wp = w.ctypes.data
        if inf is not None:
            inf = np.ascontiguousarray(inf, dtype=np.uint8)
            ip = inf.ctypes.data
        self._check(self.L.sk_run_add_facts_sym(self.h, pred.encode(), n, len(cols), C.cast(ptrs, C.c_void_p),
                                                wp, ip))

    def load(self):
        self._check(self.L.sk_run_load(self.h))
        self._keep = []

    def execute(self):
        self._check(self.L.sk_run_execute(self.h))

    def reset(self):
        """Restore the post-load relation state (re-run over the resident EDB)."""
        self._check(self.L.sk_run_reset(self.h))

    def set_dense(self, on: bool):
        """Dense cell arrays for eligible relations (True, the default) or the
        sorted-list path for every relation (False)."""
        self._check(self.L.sk_run_set_dense(self.h, 1 if on else 0))

    def set_pull(self, mode: int):
        """Pull rounds: 0 never, 1 by the cost model (default), 2 whenever
        the rule's shape allows.  Results are identical in every mode."""
        self._check(self.L.sk_run_set_pull(self.h, int(mode)))

    def set_graph(self, on: bool):
        """Device-driven rounds captured in a hipGraph once every delta
        variant takes the word-parallel pull (True, the default)."""
        self._check(self.L.sk_run_set_graph(self.h, 1 if on else 0))

    def set_pack(self, on: bool):
        """Packed pull deltas when every Trop weight stays below 2^26 (True,
        the default); results are identical either way."""
        self._check(self.L.sk_run_set_pack(self.h, 1 if on else 0))

    @property
    def execute_ms(self) -> float:
        return float(self.L.sk_run_execute_ms(self.h))

    def iterations(self) -> List[int]:
        return [int(self.L.sk_result_iterations(self.h, k)) for k in range(self.prog.num_strata)]

    def universe(self) -> int:
        return int(self.L.sk_result_universe(self.h))

    def size(self, pred: str) -> int:
        return int(self.L.sk_result_size(self.h, pred.encode()))

    def relation(self, pred: str) -> Relation:
        p = next(x for x in self.prog.predicates if x.name == pred)
        n = self.size(pred)
        a = int(self.L.sk_result_arity(self.h, pred.encode()))
        rows = (SkVal * max(1, n * max(1, a)))()
        trop = str(p.semiring).lower() == "trop"
        w = np.zeros(max(1, n), dtype=np.int64)
        inf = np.zeros(max(1, n), dtype=np.uint8)
        code = self.L.sk_result_read(self.h, pred.encode(), C.cast(rows, C.c_void_p), w.ctypes.data,
                                     inf.ctypes.data, None)
        if code:
            _raise(code)
        arr = np.frombuffer(rows, dtype=np.dtype([("kind", "<u4"), ("pad", "<u4"), ("v", "<i8")]),
                            count=n * a)
        kinds = arr["kind"].reshape(n, a).copy()
        vals = arr["v"].reshape(n, a).copy()
        return Relation(pred, a, "trop" if trop else "bool", None if (kinds == 0).all() else kinds, vals,
                        w[:n].copy() if trop else None, inf[:n].copy() if trop else None)

    def relation_sym(self, pred: str):
        """All-Sym relation as (list of u32 column arrays, weights or None, inf or None)."""
        p = next(x for x in self.prog.predicates if x.name == pred)
        n = self.size(pred)
        cols = [np.empty(max(1, n), dtype=np.uint32) for _ in range(p.arity)]
        trop = str(p.semiring).lower() == "trop"
        w = np.empty(max(1, n), dtype=np.int64) if trop else None
        inf = np.empty(max(1, n), dtype=np.uint8) if trop else None
        ptrs = (C.c_void_p * max(1, p.arity))(*[c.ctypes.data for c in cols])
        code = self.L.sk_result_read_sym(self.h, pred.encode(), C.cast(ptrs, C.c_void_p),
                                         w.ctypes.data if trop else None, inf.ctypes.data if trop else None)
        if code:
            _raise(code)
        return [c[:n] for c in cols], (w[:n] if trop else None), (inf[:n] if trop else None)

    def stats(self) -> dict:
        out = {}
        for i, nm in enumerate(KERNEL_CLASSES + ("all",)):
            s = SkKernelStats()
            self.L.sk_run_stats(self.h, i, C.byref(s))
            out[nm] = dict(launches=int(s.launches), ms=float(s.ms), bytes=float(s.bytes), items=int(s.items))
        return out

    def kernel_records(self) -> List[dict]:
        return _kernel_records(self.L, self.h, self.L.sk_run_kernel_count, self.L.sk_run_kernel_record)

    def db(self, want: Optional[Sequence[str]] = None) -> Db:
        names = want if want is not None else [p.name for p in self.prog.predicates]
        rels = {nm: self.relation(nm) for nm in dict.fromkeys(names)}
        return Db(rels, self.iterations(), self.universe(), self.execute_ms, self.stats())


# ------------------------------------------------- reference crate functions

_default_ctx: Optional[Context] = None


def _kernel_records(L, h, count_fn, record_fn) -> List[dict]:
    recs = []
    for i in range(int(count_fn(h))):
        nm = C.create_string_buffer(64)
        cls = C.c_uint32()
        s = SkKernelStats()
        record_fn(h, i, nm, 64, C.byref(cls), C.byref(s))
        recs.append(dict(name=nm.value.decode(), cls=KERNEL_CLASSES[cls.value], launches=int(s.launches),
                         ms=float(s.ms), bytes=float(s.bytes), items=int(s.items)))
    return recs


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _take(L, ptr, n, dtype):
    if n == 0:
        L.sk_free(ptr)
        return np.empty(0, dtype=dtype)
    arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), shape=(n,)).copy()
    L.sk_free(ptr)
    return arr


def _pairs(edges) -> Tuple[np.ndarray, np.ndarray]:
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2) if len(edges) else np.zeros((0, 2), np.int64)
    return np.ascontiguousarray(e[:, 0], dtype=np.uint32), np.ascontiguousarray(e[:, 1], dtype=np.uint32)


def transitive_closure_bool(edges, ctx: Optional[Context] = None) -> List[Tuple[int, int]]:
    """strata_gpu::transitive_closure_bool (crates/strata-gpu/src/lib.rs:60-69)."""
    ctx = ctx or default_context()
    a, b = _pairs(edges)
    pa, pb, n = C.c_void_p(), C.c_void_p(), C.c_uint64()
    code = ctx.L.sk_transitive_closure_bool(ctx.h, a.ctypes.data, b.ctypes.data, len(a), C.byref(pa), C.byref(pb),
                                            C.byref(n))
    if code:
        info = SkErrorInfo()
        ctx.L.sk_ctx_error(ctx.h, C.byref(info))
        _raise(code, _err_info(info))
    s = _take(ctx.L, pa, n.value, np.uint32)
    d = _take(ctx.L, pb, n.value, np.uint32)
    return list(zip(s.tolist(), d.tolist()))


def shortest_paths_trop(edges, ctx: Optional[Context] = None) -> List[Tuple[int, int, int]]:
    """strata_gpu::shortest_paths_trop (lib.rs:77-86), interpreter semantics."""
    ctx = ctx or default_context()
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 3) if len(edges) else np.zeros((0, 3), np.int64)
    a = np.ascontiguousarray(e[:, 0], dtype=np.uint32)
    b = np.ascontiguousarray(e[:, 1], dtype=np.uint32)
    w = np.ascontiguousarray(e[:, 2], dtype=np.int64)
    pa, pb, pw, n = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_uint64()
    code = ctx.L.sk_shortest_paths_trop(ctx.h, a.ctypes.data, b.ctypes.data, w.ctypes.data, len(a), C.byref(pa),
                                        C.byref(pb), C.byref(pw), C.byref(n))
    if code:
        info = SkErrorInfo()
        ctx.L.sk_ctx_error(ctx.h, C.byref(info))
        _raise(code, _err_info(info))
    s = _take(ctx.L, pa, n.value, np.uint32)
    d = _take(ctx.L, pb, n.value, np.uint32)
    ww = _take(ctx.L, pw, n.value, np.int64)
    return list(zip(s.tolist(), d.tolist(), ww.tolist()))


def count_triangles(edges, ctx: Optional[Context] = None) -> int:
    """strata_gpu::count_triangles (lib.rs:94-101)."""
    ctx = ctx or default_context()
    a, b = _pairs(edges)
    out = C.c_uint64()
    code = ctx.L.sk_count_triangles(ctx.h, a.ctypes.data, b.ctypes.data, len(a), C.byref(out))
    if code:
        _raise(code)
    return int(out.value)


def count_4cliques(edges, ctx: Optional[Context] = None) -> int:
    """strata_gpu::count_4cliques (lib.rs:109-114)."""
    ctx = ctx or default_context()
    a, b = _pairs(edges)
    out = C.c_uint64()
    code = ctx.L.sk_count_4cliques(ctx.h, a.ctypes.data, b.ctypes.data, len(a), C.byref(out))
    if code:
        _raise(code)
    return int(out.value)


def intern_terms(keys, ctx: Optional[Context] = None) -> List[int]:
    """strata_gpu::intern_terms (lib.rs:122-130): index in the sorted distinct table."""
    ctx = ctx or default_context()
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
    out = np.empty(max(1, len(k)), dtype=np.uint32)
    code = ctx.L.sk_intern_terms(ctx.h, k.ctypes.data, len(k), out.ctypes.data)
    if code:
        _raise(code)
    return out[:len(k)].tolist()


def sort_u64(keys, ctx: Optional[Context] = None) -> np.ndarray:
    """gpu_sort_u64 (crates/strata-gpu/src/sort.rs:396): the keys in ascending order."""
    ctx = ctx or default_context()
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
    out = np.empty(max(1, len(k)), dtype=np.uint64)
    code = ctx.L.sk_sort_u64(ctx.h, k.ctypes.data, len(k), out.ctypes.data)
    if code:
        _raise(code)
    return out[:len(k)]


def sort_unique_u64(keys, ctx: Optional[Context] = None) -> np.ndarray:
    """gpu_sort_unique_u64 (sort.rs:409): sorted distinct keys."""
    ctx = ctx or default_context()
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
    out = np.empty(max(1, len(k)), dtype=np.uint64)
    m = C.c_uint64()
    code = ctx.L.sk_sort_unique_u64(ctx.h, k.ctypes.data, len(k), out.ctypes.data, C.byref(m))
    if code:
        _raise(code)
    return out[:m.value]


def sort_unique_pairs(pairs, ctx: Optional[Context] = None) -> List[Tuple[int, int]]:
    """sort_unique_pairs (sort.rs:518)."""
    ctx = ctx or default_context()
    a, b = _pairs(pairs)
    oa = np.empty(max(1, len(a)), np.uint32)
    ob = np.empty(max(1, len(a)), np.uint32)
    m = C.c_uint64()
    code = ctx.L.sk_sort_unique_pairs(ctx.h, a.ctypes.data, b.ctypes.data, len(a), oa.ctypes.data, ob.ctypes.data,
                                      C.byref(m))
    if code:
        _raise(code)
    return list(zip(oa[:m.value].tolist(), ob[:m.value].tolist()))


def sort_min_pairs(triples, ctx: Optional[Context] = None) -> List[Tuple[int, int, int]]:
    """sort_min_pairs (sort.rs:565): per distinct (a, b) the minimum weight."""
    ctx = ctx or default_context()
    e = np.asarray(triples, dtype=np.int64).reshape(-1, 3) if len(triples) else np.zeros((0, 3), np.int64)
    a = np.ascontiguousarray(e[:, 0], dtype=np.uint32)
    b = np.ascontiguousarray(e[:, 1], dtype=np.uint32)
    w = np.ascontiguousarray(e[:, 2], dtype=np.int64)
    n = len(a)
    oa, ob = np.empty(max(1, n), np.uint32), np.empty(max(1, n), np.uint32)
    ow = np.empty(max(1, n), np.int64)
    m = C.c_uint64()
    code = ctx.L.sk_sort_min_pairs(ctx.h, a.ctypes.data, b.ctypes.data, w.ctypes.data, n, oa.ctypes.data,
                                   ob.ctypes.data, ow.ctypes.data, C.byref(m))
    if code:
        _raise(code)
    k = m.value
    return list(zip(oa[:k].tolist(), ob[:k].tolist(), ow[:k].tolist()))


def asp_simplify_filter(head, body_start, body_lit, poss, cert, ctx: Optional[Context] = None):
    """simplify_filter (crates/strata-gpu/src/asp.rs:121-135) over a ground
    program in CSR form; returns (head, body_start, body_lit) as numpy arrays."""
    ctx = ctx or default_context()
    h = np.ascontiguousarray(head, dtype=np.int32)
    bs = np.ascontiguousarray(body_start, dtype=np.uint32)
    bl = np.ascontiguousarray(body_lit, dtype=np.int32)
    p = np.ascontiguousarray(poss, dtype=np.uint8)
    c = np.ascontiguousarray(cert, dtype=np.uint8)
    if len(bs) != len(h) + 1 or len(p) != len(c):
        raise ValueError("body_start needs n_rules + 1 offsets; poss and cert one entry per atom")
    oh = np.empty(max(1, len(h)), np.int32)
    ob = np.empty(len(h) + 1, np.uint32)
    ol = np.empty(max(1, len(bl)), np.int32)
    nr, nl = C.c_uint64(), C.c_uint64()
    code = ctx.L.sk_asp_simplify_filter(ctx.h, h.ctypes.data, bs.ctypes.data, bl.ctypes.data, len(h), p.ctypes.data,
                                        c.ctypes.data, len(p), oh.ctypes.data, ob.ctypes.data, ol.ctypes.data,
                                        C.byref(nr), C.byref(nl))
    if code:
        _raise(code)
    return oh[:nr.value].copy(), ob[:nr.value + 1].copy(), ol[:nl.value].copy()


def program_from_json(json_text):
    """A serialised CoreProgram decoded by the library's decoder
    (sk_program_from_json; no device), returned as a core_ir.CoreProgram."""
    from . import core_ir as ir
    L = load_library()
    raw = json_text.encode() if isinstance(json_text, str) else bytes(json_text)
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    code = L.sk_program_from_json(raw, len(raw), C.byref(h), err, 256)
    if code:
        _raise(code, {"message": err.value.decode(errors="replace")})
    try:
        p = L.sk_program_json_get(h).contents
        ops = ir.AGG_OPS

        def term(t):
            if t.kind == 0:
                return ir.Var(t.a)
            if t.kind == 1:
                return ir.Const(t.a)
            if t.kind == 2:
                return ir.Int(t.b)
            if t.kind == 3:
                return ir.Agg(ops[t.a], t.b)
            _raise(103, {"message": "compound terms (@terms) are outside this engine"})

        def atom(i):
            a = p.atoms[i]
            return ir.CoreAtom(a.pred.decode(), [term(p.terms[a.first_term + k]) for k in range(a.n_terms)])

        preds = [ir.CorePred(p.preds[i].name.decode(), p.preds[i].arity, ir.SEMIRINGS[p.preds[i].semiring],
                             p.preds[i].stratum) for i in range(p.n_preds)]
        rules = []
        for i in range(p.n_rules):
            r = p.rules[i]
            body = [(ir.Neg if p.atoms[r.first_lit + k].negated else ir.Pos)(atom(r.first_lit + k))
                    for k in range(r.n_lits)]
            rules.append(ir.CoreRule(atom(r.head), body, r.stratum, r.var_count))
        return ir.CoreProgram(preds, rules, p.num_strata)
    finally:
        L.sk_program_json_free(h)


def binary_plan_intermediate(edges) -> int:
    """strata_gpu::binary_plan_intermediate (lib.rs:138-143); host-only."""
    L = load_library()
    a, b = _pairs(edges)
    return int(L.sk_binary_plan_intermediate(a.ctypes.data, b.ctypes.data, len(a)))


# ------------------------------------------------------------------ batched

class Batch:
    """Many independent programs evaluated together (sk_batch_*)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.L = ctx.L
        h = C.c_void_p()
        code = self.L.sk_batch_create(ctx.h, C.byref(h))
        if code:
            _raise(code)
        self.h = h
        ctx._adopt(self)
        self.progs = []
        self._keep = []

    def add(self, prog, bulk: dict) -> int:
        """Add one program with all-Sym EDB: pred -> (columns, weights or None)."""
        flat = FlatProgram(prog)
        names = list(bulk)
        n_edb = len(names)
        pn = (C.c_char_p * max(1, n_edb))(*[n.encode() for n in names])
        ns = (C.c_uint64 * max(1, n_edb))()
        colptrs = (C.c_void_p * max(1, n_edb))()
        wptrs = (C.c_void_p * max(1, n_edb))()
        keep = [flat, pn]
        for i, nm in enumerate(names):
            cols, w = bulk[nm][0], bulk[nm][1]
            cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in cols]
            ns[i] = len(cols[0]) if cols else 0
            arr = (C.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
            colptrs[i] = C.cast(arr, C.c_void_p)
            keep += [cols, arr]
            if w is not None:
                w = np.ascontiguousarray(w, dtype=np.int64)
                wptrs[i] = w.ctypes.data
                keep.append(w)
        idx = C.c_uint32()
        code = self.L.sk_batch_add(self.h, C.byref(flat.c), n_edb, C.cast(pn, C.c_void_p), C.cast(ns, C.c_void_p),
                                   C.cast(colptrs, C.c_void_p), C.cast(wptrs, C.c_void_p), C.byref(idx))
        if code:
            _raise(code)
        self.progs.append(prog)
        self._keep.append(keep)  # the library reads the facts at sk_batch_load
        return idx.value

    def set_graph(self, enabled: bool):
        """Device-driven rounds where a stratum allows them (default on)."""
        code = self.L.sk_batch_set_graph(self.h, 1 if enabled else 0)
        if code:
            _raise(code)

    def load(self):
        code = self.L.sk_batch_load(self.h)
        self._keep = []
        if code:
            _raise(code)

    def execute(self):
        code = self.L.sk_batch_execute(self.h)
        if code:
            _raise(code)

    @property
    def execute_ms(self) -> float:
        return float(self.L.sk_batch_execute_ms(self.h))

    def kernel_records(self) -> List[dict]:
        """Per-kernel records of the last execute, summed over the groups
        (profiling on in the context)."""
        return _kernel_records(self.L, self.h, self.L.sk_batch_kernel_count, self.L.sk_batch_kernel_record)

    def status(self, i: int) -> Tuple[int, dict]:
        info = SkErrorInfo()
        code = self.L.sk_batch_status(self.h, i, C.byref(info))
        return int(code), _err_info(info)

    def iterations(self, i: int) -> List[int]:
        return [int(self.L.sk_batch_iterations(self.h, i, k)) for k in range(self.progs[i].num_strata)]

    def relation_sym(self, i: int, pred: str):
        p = next(x for x in self.progs[i].predicates if x.name == pred)
        n = int(self.L.sk_batch_result_size(self.h, i, pred.encode()))
        if n < 0:
            code, info = self.status(i)
            _raise(code or 104, info)
        cols = [np.empty(max(1, n), dtype=np.uint32) for _ in range(p.arity)]
        trop = str(p.semiring).lower() == "trop"
        w = np.empty(max(1, n), dtype=np.int64) if trop else None
        ptrs = (C.c_void_p * max(1, p.arity))(*[c.ctypes.data for c in cols])
        code = self.L.sk_batch_result_read_sym(self.h, i, pred.encode(), C.cast(ptrs, C.c_void_p),
                                               w.ctypes.data if trop else None, None)
        if code:
            _raise(code)
        return [c[:n] for c in cols], (w[:n] if trop else None)

    def free(self):
        if self.h and self.ctx.h:
            self.L.sk_batch_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ serving

class Server:
    """Resident-EDB cheapest-cost query serving (sk_server_*)."""

    def __init__(self, ctx: Context, src, dst, w, budget_bytes: int = 0, max_batch: int = 64):
        self.ctx = ctx
        self.L = ctx.L
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.int64)
        h = C.c_void_p()
        code = self.L.sk_server_create(ctx.h, src.ctypes.data, dst.ctypes.data, w.ctypes.data, len(src),
                                       budget_bytes, max_batch, C.byref(h))
        if code:
            _raise(code)
        self.h = h
        ctx._adopt(self)

    def submit(self, sources, targets=()) -> int:
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32))
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.uint32))
        tk = C.c_uint64()
        code = self.L.sk_server_submit(self.h, s.ctypes.data if len(s) else None, len(s),
                                       t.ctypes.data if len(t) else None, len(t), C.byref(tk))
        if code:
            _raise(code)
        return tk.value

    def step(self) -> Tuple[int, int]:
        a, wt = C.c_uint32(), C.c_uint32()
        code = self.L.sk_server_step(self.h, C.byref(a), C.byref(wt))
        if code:
            _raise(code)
        return a.value, wt.value

    def poll(self, ticket: int):
        """None while running; else dict(rows, iterations, latency_ms)."""
        n, it, lat = C.c_uint64(), C.c_uint64(), C.c_double()
        code = self.L.sk_server_poll(self.h, ticket, C.byref(n), C.byref(it), C.byref(lat))
        if code == 105:
            return None
        if code:
            _raise(code)
        return dict(rows=n.value, iterations=it.value, latency_ms=lat.value)

    def take(self, ticket: int, n: int):
        s = np.empty(max(1, n), dtype=np.uint32)
        d = np.empty(max(1, n), dtype=np.uint32)
        w = np.empty(max(1, n), dtype=np.int64)
        code = self.L.sk_server_take(self.h, ticket, s.ctypes.data, d.ctypes.data, w.ctypes.data)
        if code:
            _raise(code)
        return s[:n], d[:n], w[:n]

    def release(self, ticket: int):
        self.L.sk_server_release(self.h, ticket)

    def poll_source(self, ticket: int):
        """Streaming: None while the next source's answer runs; else
        dict(index, source, rows) (index == number of sources once every
        source was taken)."""
        i, src, n = C.c_uint32(), C.c_uint32(), C.c_uint64()
        code = self.L.sk_server_poll_source(self.h, ticket, C.byref(i), C.byref(src), C.byref(n))
        if code == 105:
            return None
        if code:
            _raise(code)
        return dict(index=i.value, source=src.value, rows=n.value)

    def take_source(self, ticket: int, n: int, discard: bool = False):
        """The next source's rows (Z, w), sorted by Z, then released on the
        device (with discard, released without a copy)."""
        if discard:
            code = self.L.sk_server_take_source(self.h, ticket, None, None)
            if code:
                _raise(code)
            return None
        d = np.empty(max(1, n), dtype=np.uint32)
        w = np.empty(max(1, n), dtype=np.int64)
        code = self.L.sk_server_take_source(self.h, ticket, d.ctypes.data, w.ctypes.data)
        if code:
            _raise(code)
        return d[:n], w[:n]

    def set_hold(self, nbytes: int):
        code = self.L.sk_server_set_hold(self.h, nbytes)
        if code:
            _raise(code)

    def stats(self) -> dict:
        out = np.zeros(9, dtype=np.uint64)
        code = self.L.sk_server_stats(self.h, out.ctypes.data, 9)
        if code:
            _raise(code)
        keys = ("steps", "pull_rounds", "push_joins", "lanes", "narrow", "groups", "edges", "pull_pairs",
                "push_pairs")
        return {k: int(v) for k, v in zip(keys, out)}

    def free(self):
        if self.h and self.ctx.h:
            self.L.sk_server_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
