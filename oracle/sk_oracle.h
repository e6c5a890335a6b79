/*
 * sk_oracle.h — CPU oracle: a plain-C restatement of the reference
 * interpreter's naive and semi-naive evaluation (strata-eval).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU engine is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Followed reference files (paths relative to the reference repository):
 *   crates/strata-eval/src/seminaive.rs   run_semi_naive / saturate_stratum /
 *                                          apply_round / solve
 *   crates/strata-eval/src/naive.rs       run / saturate_stratum / solve /
 *                                          eval_aggregate_rule / fold_aggregate /
 *                                          unify / ground_atom / negation_holds
 *   crates/strata-eval/src/store.rs       Relation::combine, Db
 *   crates/strata-eval/src/value.rs       Ann::oplus / otimes / otimes_id
 *   crates/strata-ir/src/trop.rs          Weight (i64 + distinct +inf), checked add
 *   crates/strata-ir/src/core/mod.rs      CoreProgram / CoreRule / CoreAtom / CoreTerm
 *
 * The flattened Core-IR layout below is the same one the engine's C ABI
 * (include/strata_k.h) uses, defined independently here so the checker shares
 * no code with the thing it checks.
 */
#ifndef SK_ORACLE_H
#define SK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GroundVal (crates/strata-ir/src/value.rs:19-24).  Ordering: kind first
 * (Sym < Int < Term, the derive(Ord) variant order), then the payload. */
enum { ORC_SYM = 0, ORC_INT = 1 };
typedef struct {
    uint32_t kind;
    uint32_t pad;
    int64_t v;
} orc_val;

/* Ann (crates/strata-eval/src/value.rs:16-20): Unit, or W(Weight) where a
 * Weight is Finite(i64) or PosInf (crates/strata-ir/src/trop.rs:35-39). */
enum { ORC_ANN_UNIT = 0, ORC_ANN_W = 1 };
typedef struct {
    uint8_t kind;
    uint8_t inf;
    uint8_t pad[6];
    int64_t w;
} orc_ann;

/* Semiring (core/mod.rs:19-22). */
enum { ORC_BOOL = 0, ORC_TROP = 1 };

/* CoreTerm (core/mod.rs:79-94).  VAR: a = slot.  CONST: a = symbol id.
 * INT: b = value.  AGG: a = op (AggOp order count,sum,min,max,prob_or), b = slot.
 * COMPOUND (@terms) is not evaluated by this oracle. */
enum { ORC_T_VAR = 0, ORC_T_CONST = 1, ORC_T_INT = 2, ORC_T_AGG = 3, ORC_T_COMPOUND = 4 };
enum { ORC_AGG_COUNT = 0, ORC_AGG_SUM = 1, ORC_AGG_MIN = 2, ORC_AGG_MAX = 3, ORC_AGG_PROB_OR = 4 };
typedef struct {
    uint32_t kind;
    uint32_t a;
    int64_t b;
} orc_term;

/* CoreAtom / CoreLiteral: negated != 0 is CoreLiteral::Neg. */
typedef struct {
    const char* pred;
    uint32_t negated;
    uint32_t first_term;
    uint32_t n_terms;
    uint32_t pad;
} orc_atom;

/* CoreRule (core/mod.rs:50-61): head is an atom index, the body is
 * atoms[first_lit .. first_lit + n_lits) in declaration order. */
typedef struct {
    uint32_t head;
    uint32_t first_lit;
    uint32_t n_lits;
    uint32_t stratum;
    uint32_t var_count;
    uint32_t pad;
} orc_rule;

/* CorePred (core/mod.rs:42-47). */
typedef struct {
    const char* name;
    uint32_t arity;
    uint32_t semiring;
    uint32_t stratum;
    uint32_t pad;
} orc_pred;

typedef struct {
    const orc_pred* preds;
    uint32_t n_preds;
    const orc_rule* rules;
    uint32_t n_rules;
    const orc_atom* atoms;
    uint32_t n_atoms;
    const orc_term* terms;
    uint32_t n_terms;
    uint32_t num_strata;
} orc_program;

/* EvalError (crates/strata-eval/src/naive.rs:26-54), plus PANIC for the
 * reference's panics (annotation mismatch in oplus/otimes, value.rs:56-62). */
enum {
    ORC_OK = 0,
    ORC_ERR_UNKNOWN_PRED = 1,
    ORC_ERR_OVERFLOW = 2,
    ORC_ERR_UNSAFE_HEAD_VAR = 3,
    ORC_ERR_UNSAFE_NEG_VAR = 4,
    ORC_ERR_NEGATION_NOT_STRATIFIED = 5,
    ORC_ERR_NEGATIVE_WEIGHT_CYCLE = 6,
    ORC_ERR_AGGREGATE_IN_RECURSION = 7,
    ORC_ERR_MALFORMED_AGGREGATE_HEAD = 8,
    ORC_ERR_NON_INTEGER_AGGREGAND = 9,
    ORC_ERR_UNSUPPORTED = 10,
    ORC_PANIC = 50,
    ORC_ERR_BAD_INPUT = 60
};

typedef struct {
    int32_t code;
    uint32_t slot;
    int64_t lhs;
    int64_t rhs;
    char pred[128];
    char msg[256];
} orc_error;

typedef struct orc_db orc_db;

/* Db::from_program (store.rs:67-73). */
orc_db* orc_db_new(const orc_program* prog);
void orc_db_free(orc_db* db);

/* Seed EDB facts, ⊕-combined on insert (seminaive.rs:30-36).  Row-major
 * values, n rows of the predicate's arity.  anns may be NULL (all Unit). */
int32_t orc_insert(orc_db* db, const char* pred, uint64_t n, const orc_val* rows,
                   const orc_ann* anns, orc_error* err);
/* Same for all-Sym rows given as `arity` column arrays (cols[j][i]).  w == NULL
 * gives Bool (Unit) facts, otherwise Trop facts with optional +inf flags. */
int32_t orc_insert_sym(orc_db* db, const char* pred, uint64_t n, uint32_t arity,
                       const uint32_t* const* cols, const int64_t* w,
                       const uint8_t* inf, orc_error* err);

/* Evaluate to the least fixpoint.  semi != 0 runs run_semi_naive
 * (seminaive.rs:29), else naive run (naive.rs:96).  indexed != 0 narrows each
 * body-literal scan to rows agreeing with the already-bound columns; the rows
 * visited that unify, and their order, are exactly those of the full scan. */
int32_t orc_run(orc_db* db, int semi, int indexed, orc_error* err);

/* Per-stratum round count: semi-naive `iters` (delta rounds after the
 * bootstrap, seminaive.rs:96-104) or naive `iters` (naive.rs:186-231). */
uint64_t orc_iterations(const orc_db* db, uint32_t stratum);
/* Distinct Sym values after the EDB load (naive.rs:143-155). */
uint64_t orc_universe(const orc_db* db);
/* Total unify attempts performed (work counter for the CPU baseline). */
uint64_t orc_unify_steps(const orc_db* db);

/* Relation contents in BTreeMap order (sorted). */
int64_t orc_rel_size(orc_db* db, const char* pred);
int64_t orc_rel_arity(orc_db* db, const char* pred);
int32_t orc_rel_copy(orc_db* db, const char* pred, orc_val* rows, orc_ann* anns);

/* Triangle count of an acyclically oriented simple graph (sorted unique
 * pairs a[i] -> b[i], ids < nnodes): crates/strata-gpu/src/wcoj.rs:27-43. */
uint64_t orc_tri_count_oriented(const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t nnodes);

/* Multi-source cheapest cost, an independent check of
 *   cost(S,Y) :- src(S), edge(S,Y).  cost(S,Z) :- cost(S,Y), edge(Y,Z).
 * by Dijkstra over (weight, hops) labels (weights must be non-negative).
 * Edges are (src[i], dst[i], w[i]), ids < nnodes; duplicates allowed (the
 * least weight wins).  For source s, out_w[s*nnodes + z] is the least weight
 * of a walk of >= 1 edges to z (INT64_MAX when none) and out_h the fewest
 * edges among such walks (UINT32_MAX when none).  *out_max_hops is the
 * largest hop count over every source: the stratum's semi-naive round count
 * (seminaive.rs:94-104).  OpenMP over sources. */
int32_t orc_cost_dijkstra(uint64_t nnodes, uint64_t nedges, const uint32_t* src, const uint32_t* dst,
                          const int64_t* w, uint32_t nsources, const uint32_t* sources, int64_t* out_w,
                          uint32_t* out_h, uint64_t* out_max_hops);

#ifdef __cplusplus
}
#endif
#endif
