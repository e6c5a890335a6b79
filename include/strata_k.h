/*
 * strata_k.h — C ABI of the MI355X (gfx950) semi-naive evaluation engine.
 *
 * The engine takes the same checked Core-IR program and ground EDB facts as
 * the reference interpreter and returns the same per-relation tuple sets,
 * Trop weights (with a distinct +inf) and per-stratum semi-naive round
 * counts.  Plain C types only: pointers, sizes, fixed-layout structs.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   sk_run_semi_naive        strata_eval::run_semi_naive
 *                            (crates/strata-eval/src/seminaive.rs:29)
 *   sk_prepare / sk_run_*    the same call split into stages so EDB upload,
 *                            device fixpoint and result download are timed
 *                            separately (no reference counterpart for the split)
 *   sk_status                strata_eval::EvalError (crates/strata-eval/src/naive.rs:26-54)
 *                            + strata_gpu::GpuError (crates/strata-gpu/src/lib.rs:14-19)
 *   sk_transitive_closure_bool  strata_gpu::transitive_closure_bool (lib.rs:67)
 *   sk_shortest_paths_trop      strata_gpu::shortest_paths_trop (lib.rs:84)
 *   sk_count_triangles          strata_gpu::count_triangles (lib.rs:99)
 *   sk_count_4cliques           strata_gpu::count_4cliques (lib.rs:112)
 *   sk_intern_terms             strata_gpu::intern_terms (lib.rs:128)
 *   sk_binary_plan_intermediate strata_gpu::binary_plan_intermediate (lib.rs:141)
 *   sk_batch_*                  no reference counterpart: many independent
 *                               programs evaluated together
 *   sk_server_*                 no reference counterpart: resident-EDB query
 *                               serving with continuous batching
 *
 * Threading: a context owns one device, one HIP stream and its memory pool;
 * calls on one context must not overlap.  Use one context per host thread.
 */
#ifndef STRATA_K_H
#define STRATA_K_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- status codes */

/* EvalError variants keep their reference names; SK_ERR_NOT_BUILT and
 * SK_ERR_DEVICE are GpuError::{NotBuilt, Cuda}. */
typedef enum sk_status {
    SK_OK = 0,
    SK_ERR_UNKNOWN_PRED = 1,             /* EvalError::UnknownPred(String)       */
    SK_ERR_OVERFLOW = 2,                 /* EvalError::Overflow(TropOverflow)    */
    SK_ERR_UNSAFE_HEAD_VAR = 3,          /* EvalError::UnsafeHeadVar(u32)        */
    SK_ERR_UNSAFE_NEG_VAR = 4,           /* EvalError::UnsafeNegVar(u32)         */
    SK_ERR_NEGATION_NOT_STRATIFIED = 5,  /* EvalError::NegationNotStratified     */
    SK_ERR_NEGATIVE_WEIGHT_CYCLE = 6,    /* EvalError::NegativeWeightCycle       */
    SK_ERR_AGGREGATE_IN_RECURSION = 7,   /* EvalError::AggregateInRecursion      */
    SK_ERR_MALFORMED_AGGREGATE_HEAD = 8, /* EvalError::MalformedAggregateHead    */
    SK_ERR_NON_INTEGER_AGGREGAND = 9,    /* EvalError::NonIntegerAggregand       */
    SK_ERR_UNSUPPORTED = 10,             /* EvalError::Unsupported(&str)         */
    SK_ERR_SEMIRING_MISMATCH = 50,       /* the reference panics here (value.rs:56-62) */
    SK_ERR_NOT_BUILT = 100,              /* GpuError::NotBuilt                   */
    SK_ERR_DEVICE = 101,                 /* GpuError::Cuda(String)               */
    SK_ERR_OUT_OF_BUDGET = 102,          /* the HBM budget cannot hold the round */
    SK_ERR_UNSUPPORTED_SHAPE = 103,      /* a rule shape the device planner does not lower */
    SK_ERR_INVALID_ARGUMENT = 104,
    SK_ERR_NOT_READY = 105               /* sk_server_poll: the answer is still running */
} sk_status;

typedef struct sk_error_info {
    int32_t code;       /* sk_status */
    uint32_t slot;      /* UnsafeHeadVar / UnsafeNegVar slot */
    int64_t lhs, rhs;   /* TropOverflow { lhs, rhs } */
    char pred[128];     /* UnknownPred / NegationNotStratified / AggregateInRecursion */
    char message[256];
} sk_error_info;

/* ------------------------------------------------------------------ Core-IR */

/* Semiring (crates/strata-ir/src/core/mod.rs:19-22). */
enum { SK_BOOL = 0, SK_TROP = 1 };

/* CoreTerm (core/mod.rs:79-94).  VAR: a = slot.  CONST: a = SymbolId.
 * INT: b = value.  AGG: a = AggOp (count, sum, min, max, prob_or), b = slot.
 * COMPOUND (@terms) is rejected with SK_ERR_UNSUPPORTED. */
enum { SK_T_VAR = 0, SK_T_CONST = 1, SK_T_INT = 2, SK_T_AGG = 3, SK_T_COMPOUND = 4 };
enum { SK_AGG_COUNT = 0, SK_AGG_SUM = 1, SK_AGG_MIN = 2, SK_AGG_MAX = 3, SK_AGG_PROB_OR = 4 };

typedef struct sk_term {
    uint32_t kind;
    uint32_t a;
    int64_t b;
} sk_term;

/* CoreAtom / CoreLiteral (core/mod.rs:63-75): negated != 0 is Neg. */
typedef struct sk_atom {
    const char* pred;
    uint32_t negated;
    uint32_t first_term; /* index into sk_program.terms */
    uint32_t n_terms;
    uint32_t pad;
} sk_atom;

/* CoreRule (core/mod.rs:50-61): head is an atom index; the body is
 * atoms[first_lit .. first_lit + n_lits) in declaration order. */
typedef struct sk_rule {
    uint32_t head;
    uint32_t first_lit;
    uint32_t n_lits;
    uint32_t stratum;
    uint32_t var_count;
    uint32_t pad;
} sk_rule;

/* CorePred (core/mod.rs:42-47). */
typedef struct sk_pred {
    const char* name;
    uint32_t arity;
    uint32_t semiring;
    uint32_t stratum;
    uint32_t pad;
} sk_pred;

/* CoreProgram (core/mod.rs:27-32), flattened. */
typedef struct sk_program {
    const sk_pred* preds;
    uint32_t n_preds;
    const sk_rule* rules;
    uint32_t n_rules;
    const sk_atom* atoms;
    uint32_t n_atoms;
    const sk_term* terms;
    uint32_t n_terms;
    uint32_t num_strata;
} sk_program;

/* GroundVal (crates/strata-ir/src/value.rs:19-24): kind 0 = Sym(id), 1 = Int. */
enum { SK_VAL_SYM = 0, SK_VAL_INT = 1 };
typedef struct sk_val {
    uint32_t kind;
    uint32_t pad;
    int64_t v;
} sk_val;

/* ---------------------------------------------------------------- context */

typedef struct sk_ctx sk_ctx;
typedef struct sk_run sk_run;

/* Open device `device` (HIP ordinal).  hbm_budget_bytes bounds the engine's
 * device memory (0 = 90% of free memory). */
sk_status sk_ctx_create(int device, uint64_t hbm_budget_bytes, sk_ctx** out);
void sk_ctx_destroy(sk_ctx* ctx);
/* Wait for all work queued on the context's stream. */
sk_status sk_ctx_synchronize(sk_ctx* ctx);
/* Last error of the context (or of the run, see sk_run_error). */
void sk_ctx_error(const sk_ctx* ctx, sk_error_info* out);
/* Device bytes currently held by the context's pool / high-water mark. */
uint64_t sk_ctx_device_bytes(const sk_ctx* ctx);
uint64_t sk_ctx_device_peak(const sk_ctx* ctx);
/* 1 if this library carries gfx950 device code and a device is present. */
int sk_device_available(void);
/* Library build identifier (architecture and compile flags). */
const char* sk_build_info(void);

/* --------------------------------------------------- staged run (drop-in) */

/* Lower the program onto device operators.  Returns SK_ERR_UNSUPPORTED_SHAPE
 * (with a message) for a rule the planner does not lower; there is no CPU
 * fallback. */
sk_status sk_prepare(sk_ctx* ctx, const sk_program* prog, sk_run** out);

/* A serialised CoreProgram: the JSON strata-ir's serde derives write for it
 * (crates/strata-ir/src/core/mod.rs:26-94; docs/ir-encoding.md: snake_case
 * fields, unknown fields ignored, {"kind","data"} tags for literals and
 * terms), decoded into an sk_program owned by *out (sk_program_json_get,
 * valid until sk_program_json_free).  Needs no device.  A malformed document
 * is SK_ERR_INVALID_ARGUMENT with the reason, naming the offending field's
 * path, in err (err_len bytes). */
typedef struct sk_program_json sk_program_json;
sk_status sk_program_from_json(const char* json, size_t len, sk_program_json** out, char* err, size_t err_len);
const sk_program* sk_program_json_get(const sk_program_json* p);
void sk_program_json_free(sk_program_json* p);
/* sk_prepare from a serialised CoreProgram (the form a Rust caller gets from
 * serde_json::to_string(&CoreProgram)); a decoding error is
 * SK_ERR_INVALID_ARGUMENT with its reason in sk_ctx_error's message. */
sk_status sk_prepare_json(sk_ctx* ctx, const char* json, size_t len, sk_run** out);
void sk_run_free(sk_run* run);

/* Seed EDB facts, ⊕-combined on insert (seminaive.rs:30-36).  Rows are
 * row-major GroundVals of the predicate's arity.  weights == NULL gives Unit
 * (Bool) facts; otherwise Trop facts W(weights[i]), or W(+inf) where
 * inf != NULL && inf[i] != 0.  Copied; the caller keeps ownership. */
sk_status sk_run_add_facts(sk_run* run, const char* pred, uint64_t n, const sk_val* rows,
                           const int64_t* weights, const uint8_t* inf);
/* Same for all-Sym facts given as `arity` column arrays cols[j][i]. */
sk_status sk_run_add_facts_sym(sk_run* run, const char* pred, uint64_t n, uint32_t arity,
                               const uint32_t* const* cols, const int64_t* weights,
                               const uint8_t* inf);
/* Move the staged EDB to the device and ⊕-normalise it there.  After this
 * call the EDB is resident in HBM; sk_run_execute does no host transfer of
 * relation data. */
sk_status sk_run_load(sk_run* run);
/* The device fixpoint: every stratum in order (seminaive.rs:52-65). */
sk_status sk_run_execute(sk_run* run);
/* Restore the relations to their state right after sk_run_load, so the
 * fixpoint can be executed again over the same resident EDB. */
sk_status sk_run_reset(sk_run* run);
/* Evaluation path switch (no counterpart in the reference; a tuning and
 * testing knob).  enabled != 0 (the default): a relation of a recursive
 * stratum whose column domains multiply out to a bounded number of cells is
 * ⊕-applied in place in a dense cell array while its stratum runs;
 * enabled == 0: every relation uses the sorted-list path (sort, ⊕-unique,
 * merge).  Results, round counts and errors are identical either way. */
sk_status sk_run_set_dense(sk_run* run, int enabled);
/* Destination-side (pull) rounds for rules of the linear recursive shape over
 * dense cells: 0 never, 1 when the cost model prefers them (default), 2 on
 * every delta round the shape allows.  Results are identical in every mode.
 * No reference counterpart (an evaluation strategy switch, like
 * sk_run_set_dense). */
sk_status sk_run_set_pull(sk_run* run, int mode);
/* Device-driven rounds (enabled != 0, the default): once every delta variant
 * of a stratum takes the word-parallel Bool pull, its remaining rounds run as
 * a hipGraph of rounds relaunched until a device-side convergence flag is
 * set, with no host read inside a launch.  Results and round counts are
 * identical either way.  No reference counterpart. */
sk_status sk_run_set_graph(sk_run* run, int enabled);
/* Packed pull deltas (default on): when every staged Trop weight is below
 * 2^26 and the cells are narrow, a pull round reads its delta as one 32-bit
 * word per entry ((weight << 6) | lane) and 8-byte row records; a derived
 * weight at or above 2^26 widens the cells.  Results and round counts are
 * identical either way.  No reference counterpart. */
sk_status sk_run_set_pack(sk_run* run, int enabled);
void sk_run_error(const sk_run* run, sk_error_info* out);

/* Results: rows in BTreeMap order (sorted by GroundVal tuple). */
int64_t sk_result_size(sk_run* run, const char* pred);
int64_t sk_result_arity(sk_run* run, const char* pred);
/* rows: n*arity sk_val (may be NULL); weights / inf: n each (NULL to skip);
 * ann_is_w: n flags, 1 where the row's annotation is W (may be NULL). */
sk_status sk_result_read(sk_run* run, const char* pred, sk_val* rows, int64_t* weights,
                         uint8_t* inf, uint8_t* ann_is_w);
/* All-Sym relations: cols[j] receives n u32 ids per column. */
sk_status sk_result_read_sym(sk_run* run, const char* pred, uint32_t* const* cols,
                             int64_t* weights, uint8_t* inf);
/* Semi-naive round count of a stratum, in the reference convention: delta
 * rounds after the bootstrap step, including the final empty one
 * (seminaive.rs:96-104). */
uint64_t sk_result_iterations(const sk_run* run, uint32_t stratum);
/* distinct Sym values after the EDB load (the Bellman-Ford universe). */
uint64_t sk_result_universe(const sk_run* run);

/* Timing and traffic accounting of the last sk_run_execute, measured with HIP
 * events on the context's stream.  kernel_class: 0 radix sort, 1 dedup/min,
 * 2 merge (set-difference / min-merge), 3 join (count + emit), 4 leapfrog,
 * 5 scan and small helpers, 6 all. */
typedef struct sk_kernel_stats {
    uint64_t launches;
    double ms;                /* summed kernel time */
    double bytes;             /* algorithmic bytes */
    uint64_t items;           /* elements processed */
} sk_kernel_stats;
sk_status sk_run_stats(const sk_run* run, uint32_t kernel_class, sk_kernel_stats* out);
/* Per-kernel records of the last execute (profiling on): count, then record
 * i's name, kernel class and totals. */
uint32_t sk_run_kernel_count(const sk_run* run);
sk_status sk_run_kernel_record(const sk_run* run, uint32_t i, char* name, size_t name_len, uint32_t* kernel_class,
                               sk_kernel_stats* out);
void sk_set_profiling(sk_ctx* ctx, int enabled);
/* Per-kernel records of context-level calls (the crate functions, sorts and
 * the ASP filter) with profiling on, accumulated since the last
 * sk_ctx_kernel_reset or sk_run_execute (no reference counterpart). */
sk_status sk_ctx_kernel_reset(sk_ctx* ctx);
uint32_t sk_ctx_kernel_count(sk_ctx* ctx);
sk_status sk_ctx_kernel_record(const sk_ctx* ctx, uint32_t i, char* name, size_t name_len, uint32_t* kernel_class,
                               sk_kernel_stats* out);
/* Wall-clock milliseconds of the last sk_run_execute (host timer). */
double sk_run_execute_ms(const sk_run* run);

/* One call: prepare, add facts, load, execute — run_semi_naive
 * (seminaive.rs:29).  Results stay readable through *out until
 * sk_run_free. */
sk_status sk_run_semi_naive(sk_ctx* ctx, const sk_program* prog, uint64_t n_facts,
                            const char* const* fact_preds, const uint32_t* fact_arity,
                            const sk_val* fact_rows, const int64_t* fact_weights,
                            const uint8_t* fact_inf, const uint8_t* fact_is_w, sk_run** out);

/* ------------------------------------------- reference GPU crate functions */

/* Buffers returned by these functions are allocated by the library; free
 * them with sk_free. */
void sk_free(void* p);

/* path(X,Y) :- edge(X,Y).  path(X,Z) :- edge(X,Y), path(Y,Z).  Sorted,
 * deduplicated result (crates/strata-gpu/src/lib.rs:60-69). */
sk_status sk_transitive_closure_bool(sk_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                                     uint64_t n, uint32_t** out_src, uint32_t** out_dst,
                                     uint64_t* out_n);
/* Trop closure of the same rules (lib.rs:77-86).  Semantics follow the
 * interpreter: ⊗ is checked addition (SK_ERR_OVERFLOW), the negative-cycle
 * cap is universe + 2 (SK_ERR_NEGATIVE_WEIGHT_CYCLE). */
sk_status sk_shortest_paths_trop(sk_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                                 const int64_t* w, uint64_t n, uint32_t** out_src,
                                 uint32_t** out_dst, int64_t** out_w, uint64_t* out_n);
/* Triangles of the undirected graph (duplicates, both directions and
 * self-loops allowed in the input; each triangle counted once) (lib.rs:94-101). */
sk_status sk_count_triangles(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n,
                             uint64_t* out);
/* 4-cliques of the undirected graph (lib.rs:109-114, count_4cliques at
 * crates/strata-gpu/src/wcoj.rs:162-184; duplicates, both directions and
 * self-loops allowed in the input; each 4-clique counted once).  The edge
 * list is simplified, oriented by (degree, id) and indexed on the device. */
sk_status sk_count_4cliques(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n,
                            uint64_t* out);
/* Testing knob of the 4-clique counter (no reference counterpart): send
 * every edge to the wave-per-edge path, and cap that path's LDS table of
 * common out-neighbours at lds_cap entries (1 .. 1024; 0 = 1024) so larger
 * sets take the global-memory membership test. */
sk_status sk_ctx_set_clique_test(sk_ctx* ctx, int force_heavy, uint32_t lds_cap);
/* Each key's index in the sorted distinct key table (lib.rs:122-130). */
sk_status sk_intern_terms(sk_ctx* ctx, const uint64_t* keys, uint64_t n, uint32_t* out_ids);
/* Σ_v d(v)^2 over the id-oriented simple graph (lib.rs:138-143). */
uint64_t sk_binary_plan_intermediate(const uint32_t* a, const uint32_t* b, uint64_t n);

/* ----------------------------------------------- ASP simplification */

/* The grounding simplification filter (crates/strata-gpu/src/asp.rs:53-99,
 * simplify_filter at asp.rs:121-135) over a ground program in CSR form: rule
 * i has head head[i] (an atom id, or -1 for a constraint) and body
 * body_lit[body_start[i] .. body_start[i+1]), literals signed +(atom+1) /
 * -(atom+1).  A rule dies when a positive body atom is impossible (!poss)
 * or a negative one certain (cert); in a kept rule, certain positive and
 * impossible negative literals are dropped.  Outputs (caller buffers of
 * n_rules heads, n_rules + 1 offsets, body_start[n_rules] literals) keep
 * input order.  Every certain atom must be possible (cert[a] implies
 * poss[a]): the reference's CPU filter (asp.rs:53-66) and its device filter
 * disagree on an atom certain but not possible, so such an atom, a literal
 * naming no atom below n_atoms, or decreasing offsets, is
 * SK_ERR_INVALID_ARGUMENT. */
sk_status sk_asp_simplify_filter(sk_ctx* ctx, const int32_t* head, const uint32_t* body_start,
                                 const int32_t* body_lit, uint64_t n_rules, const uint8_t* poss,
                                 const uint8_t* cert, uint64_t n_atoms, int32_t* out_head,
                                 uint32_t* out_body_start, int32_t* out_body_lit, uint64_t* out_n_rules,
                                 uint64_t* out_n_lits);
/* The same filter over device-resident buffers (no reference counterpart:
 * for callers whose ground program is already in HBM, and for
 * benchmarking).  n_lits = body_start[n_rules]; outputs are device buffers
 * of the sizes above; the two counts are returned to the host once the
 * filter has finished. */
sk_status sk_asp_simplify_filter_device(sk_ctx* ctx, const int32_t* head, const uint32_t* body_start,
                                        const int32_t* body_lit, uint64_t n_rules, uint64_t n_lits,
                                        const uint8_t* poss, const uint8_t* cert, uint64_t n_atoms,
                                        int32_t* out_head, uint32_t* out_body_start, int32_t* out_body_lit,
                                        uint64_t* out_n_rules, uint64_t* out_n_lits);

/* The same filter with no device allocation (no counterpart in the
 * reference: for callers that cannot allocate device memory, such as code
 * under HIP graph capture or a serving loop that must not call hipMalloc).
 * Its scratch (block totals, counts, the atom state table) is a static
 * device array of about 6 MiB per device, so at most 2^27 rules and 2^22
 * atoms; larger inputs are SK_ERR_INVALID_ARGUMENT with a message in
 * sk_ctx_error.  Every call on one device shares that array: calls that
 * overlap on one device (two contexts or host threads) race.  Arguments,
 * outputs and results are those of sk_asp_simplify_filter_device. */
sk_status sk_asp_simplify_filter_device_noalloc(sk_ctx* ctx, const int32_t* head, const uint32_t* body_start,
                                                const int32_t* body_lit, uint64_t n_rules, uint64_t n_lits,
                                                const uint8_t* poss, const uint8_t* cert, uint64_t n_atoms,
                                                int32_t* out_head, uint32_t* out_body_start,
                                                int32_t* out_body_lit, uint64_t* out_n_rules,
                                                uint64_t* out_n_lits);

/* ------------------------------------------------------- sort primitives */

/* The reference crate's device sort helpers over host slices:
 *   sk_sort_u64            gpu_sort_u64        (crates/strata-gpu/src/sort.rs:396)
 *   sk_sort_unique_u64     gpu_sort_unique_u64 (sort.rs:409)
 *   sk_sort_unique_pairs   sort_unique_pairs   (sort.rs:518)
 *   sk_sort_min_pairs      sort_min_pairs      (sort.rs:565): per distinct
 *                          (a, b) the minimum weight
 * Outputs are caller buffers of n entries; *out_n receives the row count. */
sk_status sk_sort_u64(sk_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out);
sk_status sk_sort_unique_u64(sk_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out, uint64_t* out_n);
sk_status sk_sort_unique_pairs(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n,
                               uint32_t* out_a, uint32_t* out_b, uint64_t* out_n);
sk_status sk_sort_min_pairs(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, const int64_t* w,
                            uint64_t n, uint32_t* out_a, uint32_t* out_b, int64_t* out_w,
                            uint64_t* out_n);
/* Device-resident stable LSD radix sort of n u64 keys on their low key_bits
 * bits, carrying optional i64 weights (w and w_alt both given or both
 * NULL).  keys_alt / w_alt are scratch of n entries; on return the sorted
 * data is in the alternate buffers when *result_in_alt is 1.  Runs on the
 * context's stream and returns once it is enqueued (sk_ctx_synchronize to
 * wait).  Counterpart of radix_sort_u64 (sort.rs:337-365). */
sk_status sk_radix_sort_device(sk_ctx* ctx, uint64_t* keys, uint64_t* keys_alt, int64_t* w,
                               int64_t* w_alt, uint64_t n, int key_bits, int* result_in_alt);

/* Radix sort autotuner of a context (no reference counterpart).  Each sort
 * of at least 2^16 records picks a candidate configuration: digit width (6,
 * 7 or 8 bits), tile (items per thread), threads per block, and the pass
 * structure (a histogram and a scatter launch per pass, or single-sweep
 * passes with one histogram launch for every pass).  The choice comes from
 * two-pass timings of every candidate, measured on the device per key type,
 * weights and size class (2^16 ... 2^26 records).
 * mode 0: a fixed default (8-bit digits); 1 (the default): a class is timed
 * on its first sort; 2: time every class now. */
sk_status sk_ctx_set_autotune(sk_ctx* ctx, int mode);
/* Pin every sort of the context to candidate c (0 .. sk_radix_candidates()
 * - 1, described by sk_radix_candidate), or -1 to return to the tuner.  A
 * testing knob. */
sk_status sk_ctx_force_radix(sk_ctx* ctx, int candidate);
/* The configuration a sort of n records with key_bits significant bits would
 * take now; *est_ms the tuner's estimate (-1 when the class is untimed). */
sk_status sk_ctx_radix_choice(sk_ctx* ctx, int wide, int weighted, uint64_t n, int key_bits,
                              int* digit_bits, int* items, double* est_ms);
/* The candidate index behind sk_ctx_radix_choice's answer. */
sk_status sk_ctx_radix_pick(sk_ctx* ctx, int wide, int weighted, uint64_t n, int key_bits, int* candidate,
                            double* est_ms);
/* Number of radix candidates, and candidate c's digit bits, items per thread
 * (for 128-bit keys when wide), threads per block and pass structure
 * (*sweep = 1: single-sweep passes). */
int sk_radix_candidates(void);
/* The tuner's measured time of one pass (ms) of every candidate for size
 * class cls (0 .. 5: 2^16, 2^18, ..., 2^26 records), written to ms[0 ..
 * n_ms - 1]; SK_ERR_INVALID_ARGUMENT when the class is untimed. */
sk_status sk_ctx_radix_timing(sk_ctx* ctx, int wide, int weighted, int cls, float* ms, int n_ms);
sk_status sk_radix_candidate(int c, int wide, int* digit_bits, int* items, int* block, int* sweep);

/* ------------------------------------------------------------ batched runs */

/* Many independent programs evaluated together.  Programs whose Core-IR is
 * identical up to predicate names share one device fixpoint: their relations
 * are keyed by program id, every round is one set of launches over all of
 * them, and each program keeps its own round count, cap and error. */
typedef struct sk_batch sk_batch;
sk_status sk_batch_create(sk_ctx* ctx, sk_batch** out);
void sk_batch_free(sk_batch* b);
/* Add one program with all-Sym EDB facts (one entry per EDB predicate).  The
 * fact buffers (columns and weights) are read at sk_batch_load, not copied
 * here: they must stay valid and unchanged until sk_batch_load returns. */
sk_status sk_batch_add(sk_batch* b, const sk_program* prog, uint32_t n_edb,
                       const char* const* edb_preds, const uint64_t* edb_n,
                       const uint32_t* const* const* edb_cols, const int64_t* const* edb_w,
                       uint32_t* out_index);
/* Device-driven rounds (hipGraph) for the batch's strata that allow them:
 * the Bool groups' word-parallel pull rounds and the Trop groups' push rounds
 * (default on).  Set before sk_batch_load.  Results, statuses and round
 * counts are identical either way. */
sk_status sk_batch_set_graph(sk_batch* b, int enabled);
sk_status sk_batch_load(sk_batch* b);
sk_status sk_batch_execute(sk_batch* b);
/* Per-program status, round counts and results. */
sk_status sk_batch_status(const sk_batch* b, uint32_t index, sk_error_info* out);
uint64_t sk_batch_iterations(const sk_batch* b, uint32_t index, uint32_t stratum);
int64_t sk_batch_result_size(sk_batch* b, uint32_t index, const char* pred);
sk_status sk_batch_result_read_sym(sk_batch* b, uint32_t index, const char* pred,
                                   uint32_t* const* cols, int64_t* weights, uint8_t* inf);
double sk_batch_execute_ms(const sk_batch* b);
/* Per-kernel records of the last sk_batch_execute summed over the batch's
 * groups (profiling on in the batch's context), as sk_run_kernel_record. */
uint32_t sk_batch_kernel_count(const sk_batch* b);
sk_status sk_batch_kernel_record(const sk_batch* b, uint32_t i, char* name, size_t name_len,
                                 uint32_t* kernel_class, sk_kernel_stats* out);

/* -------------------------------------------------------------- serving */

/* A resident, pre-sorted Trop EDB `edge(node, node)` on the device and a
 * request loop answering multi-source cheapest-cost queries
 *   cost(S,Y) :- src(S), edge(S,Y).  cost(S,Z) :- cost(S,Y), edge(Y,Z).
 * optionally filtered to a target set
 *   ans(S,Z) :- cost(S,Z), tgt(Z).
 * Every source runs as an instance in a lane of a 64-lane group of dense
 * cells; the number of groups comes from query_budget_bytes (0: half the
 * context's budget).  Lanes go to sources as they free up, so a query of any
 * number of sources runs, in pieces when it has more sources than free
 * lanes, and all instances advance together one round per step (continuous
 * batching).  Answers are the same alone, in any mix and in any pieces. */
typedef struct sk_server sk_server;
sk_status sk_server_create(sk_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                           const int64_t* w, uint64_t n_edges, uint64_t query_budget_bytes,
                           uint32_t max_batch, sk_server** out);
void sk_server_free(sk_server* s);
/* Enqueue a query; returns a ticket.  n_targets == 0 means no filter. */
sk_status sk_server_submit(sk_server* s, const uint32_t* sources, uint32_t n_sources,
                           const uint32_t* targets, uint32_t n_targets, uint64_t* ticket);
/* Advance the request loop by one device round (admits, runs, retires). */
sk_status sk_server_step(sk_server* s, uint32_t* n_active, uint32_t* n_waiting);
/* SK_OK with the answer size once finished, SK_ERR_NOT_READY before. */
sk_status sk_server_poll(sk_server* s, uint64_t ticket, uint64_t* n_rows, uint64_t* iterations,
                         double* latency_ms);
/* Copy a finished answer (rows sorted by (S, Z)) and release it. */
sk_status sk_server_take(sk_server* s, uint64_t ticket, uint32_t* src, uint32_t* dst,
                         int64_t* w);
/* Release a finished answer without copying. */
sk_status sk_server_release(sk_server* s, uint64_t ticket);
/* Streaming, one source at a time in source order, while the query runs:
 * SK_OK with the next source's position in the sorted source set, its id and
 * its row count once its answer is ready (SK_ERR_NOT_READY before); after
 * the last source, *index = the number of sources and *n_rows = 0.
 * sk_server_take_source copies that source's rows (Z, w), sorted by Z, and
 * releases them (with dst and w null: releases them without a copy).
 * sk_server_take then returns only the sources not yet taken. */
sk_status sk_server_poll_source(sk_server* s, uint64_t ticket, uint32_t* index, uint32_t* source,
                                uint64_t* n_rows);
sk_status sk_server_take_source(sk_server* s, uint64_t ticket, uint32_t* dst, int64_t* w);
/* A query whose ready answers not yet taken hold more than `bytes` of device
 * memory gets no new lanes until they are taken (0, the default: no limit). */
sk_status sk_server_set_hold(sk_server* s, uint64_t bytes);
/* Counters: out[0] steps, [1] pull rounds (one per group), [2] push joins,
 * [3] lanes, [4] 1 while the cells are narrow, [5] groups, [6] distinct
 * edges, [7] (edge, delta cell) pairs of the delta rounds that pulled, [8]
 * of those that pushed; the first n are written. */
sk_status sk_server_stats(const sk_server* s, uint64_t* out, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif /* STRATA_K_H */
