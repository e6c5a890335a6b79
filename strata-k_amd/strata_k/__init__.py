"""strata_k — MI355X (gfx950) semi-naive evaluation engine for Strata/K Core-IR.

The product path is ``libstrata_k.so`` (hand-written HIP kernels + C++ host
engine behind a C ABI, include/strata_k.h); this package is its Python front
end: the Core-IR model (``core_ir``), the ctypes binding (``engine``) and the
seeded synthetic workloads (``workloads``).
"""
from .core_ir import (AGG_OPS, BOOL, INF, TROP, Agg, Const, CoreAtom, CorePred, CoreProgram, CoreRule, Int,
                      IntV, Neg, Pos, Sym, Var, atom, cost_program, neg, pos, pred, rule, tc_program,
                      triangle_program, v)
from .engine import (Batch, Context, Db, EngineError, EvalError, GpuError, Relation, Run, Server,
                     asp_simplify_filter, binary_plan_intermediate, count_4cliques, count_triangles, default_context,
                     intern_terms, load_library, program_from_json, radix_candidate, radix_candidates, shortest_paths_trop, sort_min_pairs, sort_u64,
                     sort_unique_pairs, sort_unique_u64, transitive_closure_bool)
