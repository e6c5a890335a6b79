// sk_engine.h — host-side engine: Core-IR model, value dictionary, device
// relations, rule plans and the semi-naive driver.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "sk_dev.h"
#include "sk_join.h"
#include "strata_k.h"

namespace sk {

// Evaluation error: aborts the run with an sk_status (EvalError mirror).
struct EvalErr {
  int code = SK_OK;
  std::string pred;
  u32 slot = 0;
  i64 lhs = 0, rhs = 0;
  std::string msg;
};

// ------------------------------------------------------------ Core-IR

struct Term {
  u32 kind = SK_T_VAR;
  u32 a = 0;
  i64 b = 0;
};

struct Atom {
  std::string pred;
  bool neg = false;
  std::vector<Term> args;
  int rel = -1;  // relation index, -1 = undeclared
};

struct Rule {
  Atom head;
  std::vector<Atom> body;
  u32 stratum = 0, var_count = 0;
  bool agg = false;
};

struct PredInfo {
  std::string name;
  u32 arity = 0, sem = SK_BOOL, stratum = 0;
};

struct Program {
  std::vector<PredInfo> preds;  // one per relation (last declaration of a name wins)
  std::map<std::string, int> by_name;
  std::vector<Rule> rules;
  u32 num_strata = 0;
  static Program from_c(const sk_program* p);
  int rel_of(const std::string& name) const {
    auto it = by_name.find(name);
    return it == by_name.end() ? -1 : it->second;
  }
};

// --------------------------------------------------------- dictionary

// Order-preserving value codes: Sym(s) -> s, Int(v) -> sym_span + rank(v).
// GroundVal orders Sym before Int (crates/strata-ir/src/value.rs:15-24), so
// code order equals value order and packed keys sort in BTreeMap order.
struct Dict {
  u64 sym_span = 1;
  std::vector<i64> ints;
  u64 nvals() const { return sym_span + ints.size(); }
  int bits() const {
    u64 n = nvals();
    int b = 1;
    while (b < 64 && (1ull << b) < n) b++;
    return b;
  }
  bool int_code(i64 v, u32* c) const;
  u32 encode(u32 kind, i64 v, bool* ok) const;
  void decode(u32 code, u32* kind, i64* v) const;
};

// ------------------------------------------------------ device relations

struct Keys {
  bool wide = false;
  DBuf<u64> a;
  DBuf<u128> b;
  void ensure(Pool* p, u64 n, bool w) {
    if (w != wide) {
      a.release();
      b.release();
    }
    wide = w;
    if (w) b.ensure(p, n ? n : 1);
    else a.ensure(p, n ? n : 1);
  }
  void* ptr() const { return wide ? (void*)b.get() : (void*)a.get(); }
  void release() {
    a.release();
    b.release();
  }
};

struct RelData {
  u64 n = 0;
  Keys k;
  DBuf<i64> w;
  DBuf<u8> inf;
  void release() {
    n = 0;
    k.release();
    w.release();
    inf.release();
  }
};

// A view of another relation state's buffers: it never frees them (its
// buffers have no pool), so it must not outlive them or be written.
inline RelData view_of(const RelData& s) {
  RelData v;
  v.n = s.n;
  v.k.wide = s.k.wide;
  v.k.a.p = s.k.a.p;
  v.k.a.n = s.k.a.n;
  v.k.b.p = s.k.b.p;
  v.k.b.n = s.k.b.n;
  v.w.p = s.w.p;
  v.w.n = s.w.n;
  v.inf.p = s.inf.p;
  v.inf.n = s.inf.n;
  return v;
}

// A permuted (and CSR-indexed) copy of a relation state.
struct Index {
  u64 version = ~0ull;
  std::vector<int> perm;
  bool direct = false;  // the relation state itself is in index order (data unused)
  RelData data;
  DBuf<u64> csr;
  u64 csr_version = ~0ull;
  DBuf<u32> adj;  // second index column, for leapfrog lists
  u64 adj_version = ~0ull;
  DBuf<PullChunk> pull;  // pull-round chunk list over the CSR's rows
  u64 pull_n = 0, pull_version = ~0ull;
  u64 pull_rows = 0;  // destination rows with an index row
};


// Allocator whose value-initialisation is a no-op: staging columns are
// resized and then filled by parallel copies, so zeroing them first would
// be a wasted pass over the memory.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
using SymCol = std::vector<u32, NoInitAlloc<u32>>;

struct DRel {
  int arity = 0;
  u32 sem = SK_BOOL;
  int bits = 1;       // bits per packed column (codes < 2^bits)
  bool wide = false;  // arity * bits > 64: 128-bit keys
  RelData full, delta;
  // column order the delta list is sorted by (empty = key order); a dense
  // relation emits its delta in cell order
  std::vector<int> delta_order;
  u64 version = 1, dversion = 1;
  std::map<std::string, Index> idx;
  // staged EDB facts (host, before sk_run_load)
  std::vector<u32> st_kind;  // row-major kinds (empty when all Sym)
  std::vector<i64> st_val;   // row-major values
  std::vector<SymCol> st_sym;  // all-Sym column stage
  // batched staging straight to the device (stage_dev): u32 code columns and weights
  bool st_dev = false;
  std::vector<DBuf<u32>> st_dcol;
  DBuf<i64> st_dw;
  std::vector<i64> st_w;
  std::vector<u8> st_inf;
  bool st_weighted = false;
  u64 st_rows = 0;
  u64 st_reserve = 0;  // expected staged rows (capacity hint)
  int st_narrow = -1;  // staged weights all in [0, 2^32 - 1): 1 / 0, or -1 when not known
};

// A rule `H(r, x) :- H(j, x), E(j, r)` (any body order, either column
// order) that a round may evaluate from the destination side (k_pull).
struct PullRule {
  int rule = -1, dpos = -1;  // the rule and its delta literal
  int E = -1, eR = 0, eJ = 1;  // the fixed relation and its row / join columns
  int R = 0, L = 1;          // the head's row column and lane column
  bool delta_first = true;   // ⊗ order in the body
  u64 cost = ~0ull;          // push join output count of the delta of version cost_dv
  u64 cost_dv = ~0ull;
};

// A Bool rule `H(.., x) :- H(.., x), E(..)` (either body order) whose lane
// variable x sits in the cells' least significant column, with rows of
// whole words: a round may evaluate it word-parallel from the destination
// side (k_bitpull).  Every other column of the head and of the delta literal
// holds a distinct variable bound by E, and every variable of E is one of
// those, so each tuple of E names one (head row, delta row) pair.
struct BitPullRule {
  int rule = -1, dpos = -1;
  int E = -1;
  int eh[kMaxCols] = {}, ed[kMaxCols] = {};  // per H column: E column of the head's / delta's
                                             // variable; -1 = the lane column
  DBuf<u64> graph, csr;  // the row graph, built once per version of E
  DBuf<BitChunk> chunks;
  u64 nchunks = 0, nedges = 0;
  u64 ev = ~0ull;
};

// Dense cell state of one relation while its stratum runs (see sk_cells.h).
struct DenseRel {
  bool keep_full = false;  // some rule of the stratum reads the relation's full
                           // state: each round's delta is also merged into it
  std::vector<int> order;  // cell column order, most significant first
  bool key_order = true;   // order is the key's column order
  u64 ncells = 0;
  CellDesc cd;
  DBuf<u64> cell;
  DBuf<u32> cell32;  // narrow Trop cells
  DBuf<u32> present, maxf, infb, touched, bad;
  DBuf<u32> dom_bits[kMaxCols];
  DBuf<u64> dom_pre[kMaxCols];
  DBuf<u32> dom_list[kMaxCols];
  DBuf<u64> pdom_off[kMaxCols];    // batched: per-program domains (see CellDesc)
  DBuf<u32> pdom_codes[kMaxCols];
  // pull evaluation: per row, the last delta's lanes and list position
  std::vector<PullRule> pulls;
  DBuf<u64> rowrec;  // per row: {delta lanes, delta-list position}
  DBuf<u8> dlane;    // per delta entry: its cell's lane
  DBuf<u32> dpack;   // narrow cells below 2^26: per delta entry (weight << 6) | lane
  DBuf<u64> rowrec8; // with dpack: per row, delta-list position | count << 40
  // the delta is compacted into capacity buffers (ncells records) with no
  // host read between passes; `spare` keeps them while the delta is empty
  bool into = false;
  RelData spare;
  u64 nrows = 0;
  int lbits = 0;
  // word-parallel Bool pulls: the last delta as a bitmap and per-row stamps
  std::vector<BitPullRule> bpulls;
  DBuf<u32> dbits, rstamp;
  u64 brows = 0;  // rows of 2^blbits cells
  int blbits = 0;
  u32 stamp = 0;
};

// Distinct codes of one relation column, as a bitmap over the dictionary.
struct ColCodes {
  u64 version = ~0ull, nvals = 0, count = 0;
  DBuf<u32> bm;
  int rel = -1, col = -1;
};

// ------------------------------------------------------------- plans

struct StepPlan {
  int kind = 0;  // 0 scan, 1 join, 2 anti
  int lit = -1;
  int rel = -1;
  bool delta = false;
  bool never = false;    // the literal cannot match (arity / value mismatch)
  std::vector<int> perm; // join: index column order (probe columns first)
  int nprobe = 0;
  std::vector<int> probe_slot;     // per probe column: var slot or -1 (constant)
  std::vector<u32> probe_kind;     // term kind of the probe constant
  std::vector<i64> probe_val;      // constant value
  std::vector<std::pair<int, int>> filt;  // index-order column pairs that must be equal
  std::vector<std::pair<int, int>> newvars;  // (index-order column, slot)
  // scan
  std::vector<std::pair<int, u32>> scan_const_kind;  // (column, term kind)
  std::vector<i64> scan_const_val;
  std::vector<std::pair<int, int>> scan_eq;
  // anti
  std::vector<int> anti_slot;  // per column: slot or -1
  std::vector<u32> anti_kind;
  std::vector<i64> anti_val;
  // static error raised when the step is reached with a non-empty input
  int err = 0;
  u32 err_slot = 0;
  std::vector<int> in_slots, out_slots;
};

struct RulePlan {
  int rule = -1;
  int dpos = -1;
  std::vector<StepPlan> steps;
  bool unit_source = false;  // body without positive literals
  // leapfrog triangle
  bool tri = false;
  int tri_lit[3] = {-1, -1, -1};  // driver, lit indexed on v, lit indexed on u
  int tri_slot[3] = {0, 0, 0};    // slots of u, v, w
  int tri_col_v = 0, tri_col_w2 = 1, tri_col_u = 0, tri_col_w3 = 1;
  bool sorted_out = false;
  int head_err = 0;  // UnsafeHeadVar / Unsupported after a non-empty solve
  u32 head_err_slot = 0;
  bool trop = false;
};

// ------------------------------------------------------------ run state

// One device, one stream and everything that lives on it: a caching pool
// (blocks are reused in stream order, so a pool serves one stream only), the
// device error record and pinned words for small reads.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<Pool> pool;
  Timer timer;
  DevErr* derr = nullptr;
  u64* host_words = nullptr;
  hipEvent_t read_ev = nullptr;
  // pinned fault word of the single-sweep radix passes (Exec::fault_*)
  u32* fault_h = nullptr;
  u32* fault_d = nullptr;
  EvalErr last;
  RadixTune tune;
  // 4-clique counter test knobs: every edge to the wave path, and the wave
  // path's LDS table size (1 .. 1024).
  bool k4_force_heavy = false;
  u32 k4_tcap = 1024;
  // this device's copy of the ASP filter's static scratch (resolved at init)
  u8* asp_scratch = nullptr;
  // Two pinned staging buffers for large device-to-host result copies,
  // allocated on first use (d2h).
  static constexpr u64 kStageBytes = 64ull << 20;
  u8* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  Exec exec() { return Exec{pool.get(), stream, &timer, derr, host_words, read_ev, &tune, fault_d, fault_h}; }
  void init(int dev, u64 budget);  // throws DeviceError
  void release();
  // Device-to-host copy into caller memory.  Small copies are enqueued on
  // the stream (the caller synchronises); large ones go through the pinned
  // staging buffers in chunks, each drained into `dst` by several host
  // threads while the next chunk is in flight, and have finished on return.
  void d2h(void* dst, const void* src, u64 bytes);
  // Host-to-device copy from caller memory, the mirror of d2h: large copies
  // are packed into the pinned buffers by several host threads while the
  // previous chunk is in flight; `src` may be reused on return.
  void h2d(void* dst, const void* src, u64 bytes);
  // Host-to-device copy of host pieces laid end to end from dst (the same
  // staging as h2d); the pieces may be reused once it has synchronised the
  // stream (large copies return with every piece read).
  struct HostSeg {
    const void* p;
    u64 bytes;
  };
  void h2d_gather(void* dst, const std::vector<HostSeg>& segs);
};

struct Chunk {
  u64 n = 0;
  Keys k;
  DBuf<i64> w;
  DBuf<u8> inf;
  bool sorted_unique = false;
};

class Engine {
 public:
  Engine(Ctx* c, Program p) : ctx_(c), P_(std::move(p)) {}
  void validate();           // static checks done at sk_prepare
  void stage_rows(const std::string& pred, u64 n, const sk_val* rows, const i64* w, const u8* inf);
  void stage_sym(const std::string& pred, u64 n, u32 arity, const u32* const* cols, const i64* w, const u8* inf);
  void load();
  void execute();
  // Restore every relation to its state right after load (re-run support).
  void reset();
  std::vector<bool> snapshot_set() const;
  // results
  int64_t size(const std::string& pred);
  void read(const std::string& pred, sk_val* rows, i64* w, u8* inf, u8* is_w);
  void read_sym(const std::string& pred, u32* const* cols, i64* w, u8* inf);
  u64 iterations(u32 k) const { return k < iters_.size() ? iters_[k] : 0; }
  u64 universe() const { return universe_; }
  double exec_ms = 0;
  EvalErr pending;  // staged error (reported by load / execute)
  bool loaded = false;
  const Program& program() const { return P_; }
  DRel& rel(int i) { return R_[i]; }
  const Dict& dict() const { return D_; }
  bool has_inf() const { return has_inf_; }

  // Batched mode: column 0 of every relation is a program id in [0, np).
  // Each program keeps its own round counts, Bellman-Ford cap and errors.
  void set_partitions(u32 np);
  void set_partition_universe(u32 p, u64 u) { part_universe_[p] = u; }
  EvalErr partition_error(u32 p) const { return part_err_[p]; }
  u64 partition_iterations(u32 p, u32 k) const { return k < part_iters_[p].size() ? part_iters_[p][k] : 0; }
  u64 partition_size(const std::string& pred, u32 p);
  void partition_read_sym(const std::string& pred, u32 p, u32* const* cols, i64* w, u8* inf);
  // Attach a device-resident relation owned by the caller (not freed here).
  void attach(const std::string& pred, u64 n, void* keys, bool wide, int bits, i64* w, u8* inf, u64* csr,
              u64 max_sym, u64 universe_syms);
  // Use a given universe instead of scanning the relations at load.
  void set_universe(u64 u) { universe_override_ = u; }
  // Expected number of staged rows of a relation (avoids regrowth).
  void reserve_staging(const std::string& pred, u64 rows) {
    int r = P_.rel_of(pred);
    if (r >= 0) R_[r].st_reserve = rows;
  }
  // Batched staging straight into device columns: `total` rows of an
  // all-Sym relation whose column 0 is the program id, given as pieces laid
  // end to end (piece s: rows [at, at + n) of program pid, its other columns
  // and weights in host memory that must stay valid until load()).  The
  // callers' arrays are gathered through the pinned staging buffers; no host
  // copy of the relation is made.  set_max_sym_hint bounds every Sym id, so
  // the dictionary needs no scan of the stage.
  struct DevSeg {
    u32 pid;
    u64 at, n;
    const u32* const* cols;  // arity - 1 columns
    const i64* w;
  };
  void stage_dev(const std::string& pred, u64 total, bool weighted, const std::vector<DevSeg>& segs);
  void set_max_sym_hint(u64 max_sym) { max_sym_hint_ = max_sym; }
  // Some staged weight does not fit a narrow cell.
  void mark_not_narrow() {
    for (auto& R : R_)
      if (R.st_narrow == 1) R.st_narrow = 0;
  }
  // Keep eligible relations as dense cells during their stratum (default on).
  void set_dense(bool on) { dense_enabled_ = on; }
  // Destination-side (pull) rounds: 0 never, 1 when the cost model prefers
  // them (default), 2 whenever the rule's shape allows.
  void set_pull(int mode) { pull_mode_ = mode; }
  // Device-driven rounds captured in a hipGraph for strata whose every delta
  // variant takes the word-parallel pull (default on).
  void set_graph(bool on) { graph_enabled_ = on; }
  void set_pack(bool on) { pack_enabled_ = on; }

 private:
  std::vector<RelData> snap_;  // post-load state of relations written by rules
  Dict snap_dict_;
  std::vector<int> snap_bits_;
  std::vector<u64> snap_version_;  // each relation's version when its snapshot was (last) equal to it
  u32 nparts_ = 0;
  std::vector<u64> part_universe_;
  std::vector<EvalErr> part_err_;
  std::vector<std::vector<u64>> part_iters_;
  std::vector<u8> part_dead_;
  DBuf<DevErr> part_derr_;
  std::map<int, std::pair<u64, std::vector<u64>>> part_full_counts_;  // by relation: (version, counts)
  DBuf<u8> part_dead_dev_;
  u64 attached_max_sym_ = 0;
  u64 max_sym_hint_ = ~0ull;
  u64 universe_override_ = ~0ull;
  std::vector<bool> attached_;
  void part_round_bookkeeping(u32 k, bool is_trop, std::vector<u8>& active_prev, bool count_round);
  std::vector<u64> part_counts(int rel, bool delta);
  void drop_dead_from_deltas();

  Ctx* ctx_;
  Program P_;
  Dict D_;
  std::vector<DRel> R_;
  std::vector<u64> iters_;
  u64 universe_ = 0;
  bool has_inf_ = false;
  std::vector<bool> in_idb_;  // relation changes in the stratum being run
  DBuf<i64> int_vals_;        // device copy of Dict::ints
  std::vector<std::unique_ptr<DenseRel>> dense_;  // per relation, while its stratum runs
  std::map<std::pair<int, int>, ColCodes> colcodes_;  // column code sets, by (relation, column)
  bool dense_enabled_ = true;
  int pull_mode_ = 1;
  bool graph_enabled_ = true;
  bool pack_enabled_ = true;
  int round_variants_ = 0, round_bitpulls_ = 0, round_pulls_ = 0;  // delta variants of the current round,
                                                                    // and those pulled
  bool narrow_ok_ = true;   // every staged Trop weight is finite-or-inf in [0, 2^32-1)
  bool pack_ok_ = true;     // ... and below 2^26 (pull deltas may be packed; see CellDesc::dpack)

  void build_dict();
  void upload_rel(int r);
  void compute_universe();
  void grow_dict(const Dict& nd, u32 stratum);
  void run_stratum(u32 k);
  void plan_dense(u32 k);
  void end_dense();
  bool any_narrow() const;
  bool narrow_off_ = false;  // this stratum runs with wide cells only
  bool fold_req_ = false, fold_done_ = false;  // read the status with the next capacity compaction
  DevErr folded_{};
  void widen_dense();
  DenseRel* dense_of(int rel) { return rel >= 0 && (size_t)rel < dense_.size() ? dense_[rel].get() : nullptr; }
  void eval_aggregate(const Rule& r);
  RulePlan plan_rule(int ri, int dpos);
  void exec_plan(const RulePlan& pl, std::map<int, std::vector<Chunk>>& out);
  bool try_pull(const RulePlan& pl);
  bool try_bitpull(const RulePlan& pl);
  void ensure_bitpull_graph(BitPullRule& bp, DenseRel& dn);
  int graph_loop_rel(u32 k);
  PullDesc pull_desc(PullRule& pr, DenseRel& dn, int hr);
  int trop_graph_rel(u32 k);
  int run_trop_graph_loop(u32 k, int rel, u64 cap, u64* iters, DevErr* st, const RulePlan* pl);
  // Batched Trop strata: descriptors of a delta variant's push round, and
  // the remaining rounds device-driven (see run_trop_part_graph_loop).
  struct PushCapture {
    ScanDesc sd;
    int tcols = 0;  // table columns after the delta scan
    int erel = -1;  // the probed (fixed) relation
    ProbeDesc pd;
    OutDesc o;
    const void* skeys = nullptr;
    const i64* sw = nullptr;
    const u64* csr = nullptr;
    int part_col = -1;
  };
  bool capture_push(const RulePlan& pl, PushCapture* pc, bool parts = true);
  int trop_part_graph_rel(u32 k, int* rule, int* dpos);
  int run_trop_part_graph_loop(u32 k, int rel, const RulePlan& pl, u64* iters, std::vector<u8>& active_prev,
                               DevErr* st);
  void run_graph_loop(u32 k, int rel, u64* iters, std::vector<u8>& active_prev);
  void plan_bitpulls(u32 k, int i, DenseRel& dn);
  u64 finalize(int rel, std::vector<Chunk>& chunks);
  const RelData& source(int rel, bool delta) { return delta ? R_[rel].delta : R_[rel].full; }
  Index& get_index(int rel, bool delta, const std::vector<int>& perm, bool want_csr, bool want_adj);
  u32 code_or_never(u32 kind, i64 v, bool* never) const;
  void check_device_error();
  DevErr read_device_status();  // one copy: overflow record + narrow widen flag
  void raise_device_error(const DevErr& h);
};

}  // namespace sk

// The C ABI's context handle.
struct sk_ctx {
  sk::Ctx c;
  std::vector<sk::KRecord> kernels;  // snapshot taken by sk_ctx_kernel_count
};
