// sk_engine.cpp — semi-naive evaluation driver for gfx950.
//
// Restates the reference interpreter's evaluation order on device-resident
// relations (crates/strata-eval/src/seminaive.rs):
//   run_semi_naive (seminaive.rs:29-67)    EDB seeding with ⊕, universe, strata in order
//   saturate_stratum (seminaive.rs:70-106) aggregates once, bootstrap round, delta
//                                          rounds while any delta is non-empty, the
//                                          Trop cap universe+2 checked after each round
//   apply_round (seminaive.rs:114-210)     per rule in declaration order, per IDB
//                                          positive literal with a non-empty delta;
//                                          Jacobi: all derivations combined at the end
// A relation is a sorted, unique array of packed keys (+ Trop weight and
// +inf flag); the round's derivations per head are sorted, collapsed (Bool
// dedup / Trop min) and merged into the full relation, which yields the next
// delta (new tuples, or strictly lower weights) in the same pass.
#include "sk_engine.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <set>
#include <thread>
#include <unordered_set>

namespace sk {

// ================================================================ program

Program Program::from_c(const sk_program* p) {
  Program P;
  P.num_strata = p->num_strata;
  for (u32 i = 0; i < p->n_preds; i++) {
    const sk_pred& d = p->preds[i];
    std::string nm = d.name ? d.name : "";
    auto it = P.by_name.find(nm);
    PredInfo pi{nm, d.arity, d.semiring, d.stratum};
    if (it == P.by_name.end()) {
      P.by_name[nm] = (int)P.preds.size();
      P.preds.push_back(pi);
    } else {
      P.preds[it->second] = pi;  // a repeated declaration keeps the last one (BTreeMap insert)
    }
  }
  auto mk_atom = [&](u32 ai) {
    Atom a;
    const sk_atom& s = p->atoms[ai];
    a.pred = s.pred ? s.pred : "";
    a.neg = s.negated != 0;
    for (u32 t = 0; t < s.n_terms; t++) {
      const sk_term& tm = p->terms[s.first_term + t];
      a.args.push_back(Term{tm.kind, tm.a, tm.b});
    }
    a.rel = P.rel_of(a.pred);
    return a;
  };
  for (u32 i = 0; i < p->n_rules; i++) {
    const sk_rule& r = p->rules[i];
    Rule R;
    R.head = mk_atom(r.head);
    for (u32 l = 0; l < r.n_lits; l++) R.body.push_back(mk_atom(r.first_lit + l));
    R.stratum = r.stratum;
    R.var_count = r.var_count;
    for (auto& t : R.head.args)
      if (t.kind == SK_T_AGG) R.agg = true;
    P.rules.push_back(std::move(R));
  }
  return P;
}

// ============================================================= dictionary

bool Dict::int_code(i64 v, u32* c) const {
  auto it = std::lower_bound(ints.begin(), ints.end(), v);
  if (it == ints.end() || *it != v) return false;
  *c = (u32)(sym_span + (it - ints.begin()));
  return true;
}

u32 Dict::encode(u32 kind, i64 v, bool* ok) const {
  *ok = true;
  if (kind == SK_VAL_SYM) {
    if (v < 0 || (u64)v >= sym_span) {
      *ok = false;
      return 0;
    }
    return (u32)v;
  }
  u32 c = 0;
  *ok = int_code(v, &c);
  return c;
}

void Dict::decode(u32 code, u32* kind, i64* v) const {
  if (code < sym_span) {
    *kind = SK_VAL_SYM;
    *v = code;
  } else {
    *kind = SK_VAL_INT;
    *v = ints[code - sym_span];
  }
}

// ================================================================= helpers

static EvalErr mkerr(int code, const std::string& pred = "", const std::string& msg = "") {
  EvalErr e;
  e.code = code;
  e.pred = pred;
  e.msg = msg;
  return e;
}

void Ctx::init(int dev, u64 budget) {
  device = dev;
  SK_HIP(hipSetDevice(dev));
  pool = std::make_unique<Pool>(budget);
  // default-priority, non-blocking stream
  SK_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  timer.stream = stream;
  SK_HIP(hipMalloc((void**)&derr, sizeof(DevErr)));
  SK_HIP(hipMemsetAsync(derr, 0, sizeof(DevErr), stream));
  SK_HIP(hipHostMalloc((void**)&host_words, 64, hipHostMallocDefault));
  SK_HIP(hipHostMalloc((void**)&fault_h, 64, hipHostMallocMapped));
  *fault_h = 0;
  SK_HIP(hipHostGetDevicePointer((void**)&fault_d, fault_h, 0));
  SK_HIP(hipEventCreateWithFlags(&read_ev, hipEventDisableTiming));
  asp_scratch = asp_scratch_address(dev);
  SK_HIP(hipStreamSynchronize(stream));
}

namespace {

constexpr u64 kStageDirect = 8ull << 20;  // copies below this are one enqueued hipMemcpyAsync

// memcpy of len bytes split over up to 8 host threads (small copies: one)
void par_copy(u8* to, const u8* from, u64 len) {
  if (len < (4ull << 20)) {
    if (len) std::memcpy(to, from, len);
    return;
  }
  const unsigned nthreads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const u64 part = (len + nthreads - 1) / nthreads;
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < nthreads; t++) {
    const u64 a = std::min(len, t * part), e = std::min(len, a + part);
    if (a < e) ts.emplace_back([=] { std::memcpy(to + a, from + a, e - a); });
  }
  std::memcpy(to, from, std::min(len, part));
  for (auto& th : ts) th.join();
}

// largest of n u32 values, over up to 8 host threads
u32 par_max(const u32* v, u64 n) {
  const unsigned nthreads = n < (1u << 20) ? 1u : std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const u64 part = (n + nthreads - 1) / nthreads;
  std::vector<u32> best(nthreads, 0);
  auto scan = [&](unsigned t) {
    const u64 a = std::min(n, t * part), e = std::min(n, a + part);
    u32 m = 0;
    for (u64 i = a; i < e; i++) m = std::max(m, v[i]);
    best[t] = m;
  };
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < nthreads; t++) ts.emplace_back(scan, t);
  scan(0);
  for (auto& th : ts) th.join();
  return *std::max_element(best.begin(), best.end());
}

}  // namespace

void Ctx::h2d(void* dst, const void* src, u64 bytes) {
  if (bytes < kStageDirect) {
    if (bytes) SK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    return;
  }
  for (int b = 0; b < 2; b++)
    if (!stage[b]) {
      SK_HIP(hipHostMalloc((void**)&stage[b], kStageBytes, hipHostMallocDefault));
      SK_HIP(hipEventCreateWithFlags(&stage_ev[b], hipEventDisableTiming));
    }
  // a buffer is refilled once the copy that read it has finished (events
  // recorded by earlier calls on this stream count too)
  const u64 nchunks = (bytes + kStageBytes - 1) / kStageBytes;
  for (u64 k = 0; k < nchunks; k++) {
    const int b = (int)(k & 1);
    SK_HIP(hipEventSynchronize(stage_ev[b]));
    const u64 off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    par_copy(stage[b], (const u8*)src + off, len);
    SK_HIP(hipMemcpyAsync((u8*)dst + off, stage[b], len, hipMemcpyHostToDevice, stream));
    SK_HIP(hipEventRecord(stage_ev[b], stream));
  }
}

void Ctx::h2d_gather(void* dst, const std::vector<HostSeg>& segs) {
  std::vector<u64> pre(segs.size() + 1, 0);
  for (size_t i = 0; i < segs.size(); i++) pre[i + 1] = pre[i] + segs[i].bytes;
  const u64 total = pre.back();
  if (total < kStageDirect) {
    for (size_t i = 0; i < segs.size(); i++)
      if (segs[i].bytes)
        SK_HIP(hipMemcpyAsync((u8*)dst + pre[i], segs[i].p, segs[i].bytes, hipMemcpyHostToDevice, stream));
    return;
  }
  for (int b = 0; b < 2; b++)
    if (!stage[b]) {
      SK_HIP(hipHostMalloc((void**)&stage[b], kStageBytes, hipHostMallocDefault));
      SK_HIP(hipEventCreateWithFlags(&stage_ev[b], hipEventDisableTiming));
    }
  // bytes [a, e) of the pieces laid end to end, into `to`
  auto gather = [&](u8* to, const u64 a, const u64 e) {
    for (size_t i = std::upper_bound(pre.begin(), pre.end(), a) - pre.begin() - 1; i < segs.size() && pre[i] < e;
         i++) {
      const u64 s0 = std::max(a, pre[i]), s1 = std::min(e, pre[i + 1]);
      if (s0 < s1) std::memcpy(to + (s0 - a), (const u8*)segs[i].p + (s0 - pre[i]), s1 - s0);
    }
  };
  const unsigned nthreads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const u64 nchunks = (total + kStageBytes - 1) / kStageBytes;
  for (u64 k = 0; k < nchunks; k++) {
    const int b = (int)(k & 1);
    SK_HIP(hipEventSynchronize(stage_ev[b]));
    const u64 off = k * kStageBytes, len = std::min(kStageBytes, total - off);
    const u64 part = (len + nthreads - 1) / nthreads;
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < nthreads; t++) {
      const u64 a = std::min(len, t * part), e = std::min(len, a + part);
      if (a < e) ts.emplace_back([&, a, e] { gather(stage[b] + a, off + a, off + e); });
    }
    gather(stage[b], off, off + std::min(len, part));
    for (auto& th : ts) th.join();
    SK_HIP(hipMemcpyAsync((u8*)dst + off, stage[b], len, hipMemcpyHostToDevice, stream));
    SK_HIP(hipEventRecord(stage_ev[b], stream));
  }
}

void Ctx::d2h(void* dst, const void* src, u64 bytes) {
  if (bytes < kStageDirect) {
    if (bytes) SK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    return;
  }
  for (int b = 0; b < 2; b++)
    if (!stage[b]) {
      SK_HIP(hipHostMalloc((void**)&stage[b], kStageBytes, hipHostMallocDefault));
      SK_HIP(hipEventCreateWithFlags(&stage_ev[b], hipEventDisableTiming));
    }
  const u64 nchunks = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](u64 k) {
    const u64 off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    SK_HIP(hipMemcpyAsync(stage[k & 1], (const u8*)src + off, len, hipMemcpyDeviceToHost, stream));
    SK_HIP(hipEventRecord(stage_ev[k & 1], stream));
  };
  issue(0);
  for (u64 k = 0; k < nchunks; k++) {
    // chunk k - 1's buffer was drained in the previous step, so chunk k + 1
    // may fill it while chunk k is drained
    if (k + 1 < nchunks) issue(k + 1);
    SK_HIP(hipEventSynchronize(stage_ev[k & 1]));
    const u64 off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    par_copy((u8*)dst + off, stage[k & 1], len);
  }
}

void Ctx::release() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  for (int b = 0; b < 2; b++) {
    if (stage[b]) (void)hipHostFree(stage[b]);
    if (stage_ev[b]) (void)hipEventDestroy(stage_ev[b]);
    stage[b] = nullptr;
    stage_ev[b] = nullptr;
  }
  timer.reset();
  pool.reset();
  if (derr) (void)hipFree(derr);
  if (host_words) (void)hipHostFree(host_words);
  if (fault_h) (void)hipHostFree(fault_h);
  if (read_ev) (void)hipEventDestroy(read_ev);
  if (stream) (void)hipStreamDestroy(stream);
  derr = nullptr;
  host_words = nullptr;
  fault_h = fault_d = nullptr;
  read_ev = nullptr;
  stream = nullptr;
}

static int key_bits_of(int arity, int bits) { return arity * bits; }

template <class F>
static void with_key(bool wide, F f) {
  if (wide) f((u128*)nullptr);
  else f((u64*)nullptr);
}

// Owners of tables / head chunks, used by the kernel wrappers' allocators.
struct TableOwner {
  Pool* pool = nullptr;
  Table t;
  std::vector<DBuf<u32>> cols;
  DBuf<i64> w;
  DBuf<u8> inf;
  bool trop = false, has_inf = false;
};

static void table_alloc(void* ctx, Table* t, u64 n) {
  TableOwner* o = (TableOwner*)ctx;
  o->cols.resize(t->ncols);
  for (int c = 0; c < t->ncols; c++) {
    o->cols[c].reset(o->pool, n ? n : 1);
    t->col[c] = o->cols[c].get();
  }
  if (o->trop) {
    o->w.reset(o->pool, n ? n : 1);
    t->w = o->w.get();
    if (o->has_inf) {
      o->inf.reset(o->pool, n ? n : 1);
      t->inf = o->inf.get();
    } else {
      t->inf = nullptr;
    }
  } else {
    t->w = nullptr;
    t->inf = nullptr;
  }
  t->n = n;
}

struct HeadOwner {
  Pool* pool = nullptr;
  Chunk* c = nullptr;
  bool wide = false, trop = false, has_inf = false;
};

static void head_alloc(void* ctx, OutDesc* o, u64 n) {
  HeadOwner* h = (HeadOwner*)ctx;
  if (o->cd.on) {
    // rows are ⊕-applied to the head's cells by the emitting kernel
    o->keys = nullptr;
    o->w = nullptr;
    o->inf = nullptr;
    h->c->n = 0;
    return;
  }
  h->c->k.ensure(h->pool, n, h->wide);
  o->keys = h->c->k.ptr();
  if (h->trop) {
    h->c->w.ensure(h->pool, n ? n : 1);
    o->w = h->c->w.get();
    if (h->has_inf) {
      h->c->inf.ensure(h->pool, n ? n : 1);
      o->inf = h->c->inf.get();
    } else {
      o->inf = nullptr;
    }
  } else {
    o->w = nullptr;
    o->inf = nullptr;
  }
  h->c->n = n;
}

// ============================================================ validation

void Engine::validate() {
  // compound terms (@terms) are not evaluated on the device
  auto no_compound = [&](const Atom& a) {
    for (auto& t : a.args)
      if (t.kind == SK_T_COMPOUND)
        throw mkerr(SK_ERR_UNSUPPORTED, a.pred, "compound terms (@terms) are not evaluated by the device engine");
  };
  for (auto& r : P_.rules) {
    no_compound(r.head);
    for (auto& b : r.body) no_compound(b);
    int hr = r.head.rel;
    // the reference panics on an annotation mismatch in ⊗ (crates/strata-eval/src/value.rs:56-62);
    // the engine refuses such rules up front
    u32 hs = hr >= 0 ? P_.preds[hr].sem : SK_BOOL;
    if (r.agg) {
      if (hr >= 0 && hs != SK_BOOL)
        throw mkerr(SK_ERR_SEMIRING_MISMATCH, r.head.pred, "aggregate head must be Bool (the result is inserted as Unit)");
      for (auto& b : r.body)
        if (!b.neg && b.rel >= 0 && P_.preds[b.rel].sem != SK_BOOL)
          throw mkerr(SK_ERR_SEMIRING_MISMATCH, b.pred, "aggregate body over a Trop relation");
    } else {
      for (auto& b : r.body)
        if (!b.neg && b.rel >= 0 && hr >= 0 && P_.preds[b.rel].sem != hs)
          throw mkerr(SK_ERR_SEMIRING_MISMATCH, b.pred, "body literal semiring differs from the head's");
    }
    if (r.head.args.size() > (size_t)kMaxCols)
      throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, r.head.pred, "head arity above 8");
    for (auto& b : r.body)
      if (b.args.size() > (size_t)kMaxCols) throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, b.pred, "atom arity above 8");
  }
  for (auto& p : P_.preds)
    if (p.arity > (u32)kMaxCols) throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, p.name, "predicate arity above 8");
  R_.resize(P_.preds.size());
  for (size_t i = 0; i < P_.preds.size(); i++) {
    R_[i].arity = (int)P_.preds[i].arity;
    R_[i].sem = P_.preds[i].sem;
  }
}

// ================================================================ staging

void Engine::stage_rows(const std::string& pred, u64 n, const sk_val* rows, const i64* w, const u8* inf) {
  if (pending.code) return;
  int r = P_.rel_of(pred);
  if (r < 0) {
    pending = mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
    return;
  }
  DRel& R = R_[r];
  bool weighted = w != nullptr;
  if (weighted != (R.sem == SK_TROP)) {
    pending = mkerr(SK_ERR_SEMIRING_MISMATCH, pred, "fact annotation differs from the predicate's semiring");
    return;
  }
  u64 a = (u64)R.arity;
  for (u64 i = 0; i < n * a; i++) {
    R.st_kind.push_back(rows[i].kind);
    R.st_val.push_back(rows[i].v);
  }
  // keep the generic rows aligned with their weights: generic rows are
  // appended after any all-Sym stage of the same relation at load time
  if (weighted) {
    if (R.st_reserve > R.st_w.capacity()) {
      R.st_w.reserve(R.st_reserve);
      R.st_inf.reserve(R.st_reserve);
    }
    const u64 base = R.st_w.size();
    R.st_w.resize(base + n);
    R.st_inf.resize(base + n);
    for (u64 i = 0; i < n; i++) {
      u8 f = inf ? inf[i] : 0;
      R.st_w[base + i] = f ? 0 : w[i];  // +inf rows carry weight 0, like every derived +inf
      R.st_inf[base + i] = f;
      if (f) has_inf_ = true;
    }
  }
  R.st_weighted = weighted;
  R.st_rows += n;
}

void Engine::stage_sym(const std::string& pred, u64 n, u32 arity, const u32* const* cols, const i64* w,
                       const u8* inf) {
  if (pending.code) return;
  int r = P_.rel_of(pred);
  if (r < 0) {
    pending = mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
    return;
  }
  DRel& R = R_[r];
  if ((int)arity != R.arity) {
    pending = mkerr(SK_ERR_INVALID_ARGUMENT, pred, "fact arity differs from the declaration");
    return;
  }
  bool weighted = w != nullptr;
  if (weighted != (R.sem == SK_TROP)) {
    pending = mkerr(SK_ERR_SEMIRING_MISMATCH, pred, "fact annotation differs from the predicate's semiring");
    return;
  }
  if (!R.st_kind.empty()) {
    // mixed staging: convert to generic rows
    for (u64 i = 0; i < n; i++)
      for (u32 j = 0; j < arity; j++) {
        R.st_kind.push_back(SK_VAL_SYM);
        R.st_val.push_back(cols[j][i]);
      }
  } else {
    R.st_sym.resize(arity);
    for (u32 j = 0; j < arity; j++) {
      if (R.st_reserve > R.st_sym[j].capacity()) R.st_sym[j].reserve(R.st_reserve);
      const u64 at = R.st_sym[j].size();
      R.st_sym[j].resize(at + n);  // uninitialised: filled by the copy
      par_copy((u8*)(R.st_sym[j].data() + at), (const u8*)cols[j], n * 4);
    }
  }
  if (weighted) {
    if (R.st_reserve > R.st_w.capacity()) {
      R.st_w.reserve(R.st_reserve);
      R.st_inf.reserve(R.st_reserve);
    }
    const u64 base = R.st_w.size();
    R.st_w.resize(base + n);
    R.st_inf.resize(base + n);
    for (u64 i = 0; i < n; i++) {
      u8 f = inf ? inf[i] : 0;
      R.st_w[base + i] = f ? 0 : w[i];  // +inf rows carry weight 0, like every derived +inf
      R.st_inf[base + i] = f;
      if (f) has_inf_ = true;
    }
  }
  R.st_weighted = weighted;
  R.st_rows += n;
}

void Engine::stage_dev(const std::string& pred, u64 total, bool weighted, const std::vector<DevSeg>& segs) {
  int r = P_.rel_of(pred);
  if (r < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
  DRel& R = R_[r];
  if (weighted != (R.sem == SK_TROP))
    throw mkerr(SK_ERR_SEMIRING_MISMATCH, pred, "fact annotation differs from the predicate's semiring");
  if (!R.st_kind.empty() || R.st_rows || R.arity < 1)
    throw mkerr(SK_ERR_INVALID_ARGUMENT, pred, "relation already staged");
  Exec x = ctx_->exec();
  R.st_dcol.clear();
  R.st_dcol.resize(R.arity);
  for (int j = 0; j < R.arity; j++) R.st_dcol[j].reset(x.pool, total ? total : 1);
  if (weighted) R.st_dw.reset(x.pool, total ? total : 1);
  u64 at = 0;
  for (const DevSeg& sg : segs) {
    if (sg.at != at) throw mkerr(SK_ERR_INVALID_ARGUMENT, pred, "staged pieces must lie end to end");
    fill_u32(x, R.st_dcol[0].get() + sg.at, sg.pid, sg.n);
    at += sg.n;
  }
  if (at != total) throw mkerr(SK_ERR_INVALID_ARGUMENT, pred, "staged pieces do not cover the relation");
  std::vector<Ctx::HostSeg> hs;
  for (int j = 1; j < R.arity; j++) {
    hs.clear();
    for (const DevSeg& sg : segs) hs.push_back({sg.cols[j - 1], sg.n * 4});
    ctx_->h2d_gather(R.st_dcol[j].get(), hs);
  }
  if (weighted) {
    hs.clear();
    for (const DevSeg& sg : segs) hs.push_back({sg.w, sg.n * 8});
    ctx_->h2d_gather(R.st_dw.get(), hs);
  }
  R.st_dev = true;
  R.st_weighted = weighted;
  R.st_rows = total;
  R.st_narrow = weighted ? 1 : -1;
}

// =================================================================== load

void Engine::build_dict() {
  u64 maxsym = 0;
  std::set<i64> ints;
  for (auto& R : R_) {
    if (max_sym_hint_ == ~0ull)
      for (auto& c : R.st_sym) maxsym = std::max<u64>(maxsym, par_max(c.data(), c.size()));
    for (size_t i = 0; i < R.st_kind.size(); i++) {
      if (R.st_kind[i] == SK_VAL_SYM) {
        if (R.st_val[i] < 0 || R.st_val[i] > 0xffffffffll)
          throw mkerr(SK_ERR_INVALID_ARGUMENT, "", "symbol id outside u32");
        maxsym = std::max<u64>(maxsym, (u64)R.st_val[i]);
      } else {
        ints.insert(R.st_val[i]);
      }
    }
  }
  auto scan_terms = [&](const Atom& a) {
    for (auto& t : a.args) {
      if (t.kind == SK_T_CONST) maxsym = std::max<u64>(maxsym, t.a);
      if (t.kind == SK_T_INT) ints.insert(t.b);
    }
  };
  for (auto& r : P_.rules) {
    scan_terms(r.head);
    for (auto& b : r.body) scan_terms(b);
  }
  maxsym = std::max<u64>(maxsym, attached_max_sym_);
  if (max_sym_hint_ != ~0ull) maxsym = std::max<u64>(maxsym, max_sym_hint_);
  D_.sym_span = maxsym + 1;
  D_.ints.assign(ints.begin(), ints.end());
  if (D_.nvals() > 0xffffffffull) throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, "", "more than 2^32 distinct values");
}

void Engine::upload_rel(int r) {
  DRel& R = R_[r];
  if (!attached_.empty() && attached_[r]) return;
  Exec x = ctx_->exec();
  const int bits = D_.bits();
  R.bits = bits;
  R.wide = key_bits_of(R.arity, bits) > 64;
  if (key_bits_of(R.arity, bits) > 128)
    throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, P_.preds[r].name, "packed tuple wider than 128 bits");
  u64 n = R.st_rows;
  if (n == 0) {
    R.full.n = 0;
    R.st_dcol.clear();
    R.st_dw.release();
    R.st_dev = false;
    return;
  }
  bool trop = R.sem == SK_TROP;
  if (trop && R.st_narrow == 0) narrow_ok_ = false;
  // batched staging records only whether its weights are narrow
  if (trop && R.st_narrow > 0) pack_ok_ = false;
  if (trop && R.st_narrow < 0)
    for (u64 i = 0; i < n && narrow_ok_; i++)
      if (!R.st_inf[i]) {
        const u64 v = (u64)R.st_w[i];
        if (v >= 0xFFFFFFFFull) narrow_ok_ = false;
        if (v >= kPackLim) pack_ok_ = false;
      }
  // encode to codes, one column at a time
  std::vector<DBuf<u32>> dcols(R.arity);
  Cols cols{};
  u64 nsym = R.st_sym.empty() ? 0 : R.st_sym[0].size();
  for (int j = 0; j < R.arity && R.st_dev; j++) {
    dcols[j] = std::move(R.st_dcol[j]);
    cols.c[j] = dcols[j].get();
  }
  for (int j = 0; j < R.arity && !R.st_dev; j++) {
    dcols[j].reset(x.pool, n);
    cols.c[j] = dcols[j].get();
    if (nsym) ctx_->h2d(dcols[j].get(), R.st_sym[j].data(), nsym * 4);
    u64 ng = n - nsym;
    if (ng) {
      std::vector<u32> enc(ng);
      for (u64 i = 0; i < ng; i++) {
        bool ok;
        enc[i] = D_.encode(R.st_kind[i * R.arity + j], R.st_val[i * R.arity + j], &ok);
      }
      SK_HIP(hipMemcpyAsync(dcols[j].get() + nsym, enc.data(), ng * 4, hipMemcpyHostToDevice, x.stream));
      SK_HIP(hipStreamSynchronize(x.stream));
    }
  }
  Keys k1, k2;
  k1.ensure(x.pool, n, R.wide);
  k2.ensure(x.pool, n, R.wide);
  DBuf<i64> w1, w2;
  DBuf<u8> f1, f2;
  if (trop) {
    if (R.st_dev) {
      w1 = std::move(R.st_dw);
    } else {
      w1.reset(x.pool, n);
      ctx_->h2d(w1.get(), R.st_w.data(), n * 8);
    }
    w2.reset(x.pool, n);
    if (has_inf_) {
      f1.reset(x.pool, n);
      f2.reset(x.pool, n);
      if (R.st_dev) memset_async(x, f1.get(), 0, n);
      else ctx_->h2d(f1.get(), R.st_inf.data(), n);
    }
  }
  with_key(R.wide, [&](auto* tag) {
    using K = std::remove_pointer_t<decltype(tag)>;
    K* kp = (K*)k1.ptr();
    pack_cols<K>(x, cols, R.arity, bits, n, kp);
    K* sk;
    i64* sw;
    u8* sf;
    radix_sort<K>(x, kp, (K*)k2.ptr(), trop ? w1.get() : nullptr, trop ? w2.get() : nullptr,
                  (trop && has_inf_) ? f1.get() : nullptr, (trop && has_inf_) ? f2.get() : nullptr, n,
                  key_bits_of(R.arity, bits), &sk, &sw, &sf);
    R.full.k.ensure(x.pool, n, R.wide);
    if (trop) {
      R.full.w.reset(x.pool, n);
      if (has_inf_) R.full.inf.reset(x.pool, n);
    }
    u64 m = unique_runs<K>(x, sk, trop ? sw : nullptr, (trop && has_inf_) ? sf : nullptr, n, (K*)R.full.k.ptr(),
                           trop ? R.full.w.get() : nullptr, (trop && has_inf_) ? R.full.inf.get() : nullptr);
    R.full.n = m;
  });
  SK_HIP(hipStreamSynchronize(x.stream));
  R.version++;
  // release staging
  std::vector<u32>().swap(R.st_kind);
  std::vector<i64>().swap(R.st_val);
  std::vector<SymCol>().swap(R.st_sym);
  std::vector<i64>().swap(R.st_w);
  std::vector<u8>().swap(R.st_inf);
  R.st_dcol.clear();
  R.st_dw.release();
  R.st_dev = false;
  R.st_rows = 0;
  R.st_narrow = -1;
}

void Engine::compute_universe() {
  Exec x = ctx_->exec();
  u64 words = (D_.sym_span + 31) / 32;
  DBuf<u32> bm(x.pool, words ? words : 1);
  memset_async(x, bm.get(), 0, words * 4);
  for (auto& R : R_) {
    const int bits = R.bits;
    if (R.full.n == 0 || R.arity == 0) continue;
    std::vector<DBuf<u32>> cols(R.arity);
    std::vector<u32*> cp(R.arity);
    for (int j = 0; j < R.arity; j++) {
      cols[j].reset(x.pool, R.full.n);
      cp[j] = cols[j].get();
    }
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      unpack_cols<K>(x, (const K*)R.full.k.ptr(), R.full.n, R.arity, bits, cp.data());
    });
    for (int j = 0; j < R.arity; j++) mark_codes(x, cp[j], R.full.n, (u32)std::min<u64>(D_.sym_span, 0xffffffffull), bm.get());
    SK_HIP(hipStreamSynchronize(x.stream));
  }
  universe_ = popcount_bits(x, bm.get(), words);
}

void Engine::load() {
  if (pending.code) throw pending;
  narrow_ok_ = true;
  pack_ok_ = true;
  build_dict();
  for (size_t r = 0; r < R_.size(); r++) upload_rel((int)r);
  // distinct Sym values after the EDB load (crates/strata-eval/src/naive.rs:143-155)
  // (batched runs cap each program by its own universe instead)
  if (universe_override_ != ~0ull) universe_ = universe_override_;
  else if (!nparts_) compute_universe();
  for (size_t r = 0; r < R_.size(); r++)
    if (!attached_.empty() && attached_[r] && R_[r].bits < D_.bits())
      throw mkerr(SK_ERR_INVALID_ARGUMENT, P_.preds[r].name, "attached relation packed narrower than the dictionary");
  Exec x = ctx_->exec();
  if (!D_.ints.empty()) {
    int_vals_.reset(x.pool, D_.ints.size());
    SK_HIP(hipMemcpyAsync(int_vals_.get(), D_.ints.data(), D_.ints.size() * 8, hipMemcpyHostToDevice, x.stream));
  }
  // snapshot relations that rules write (heads), so reset() can restore them
  snap_.clear();
  snap_.resize(R_.size());
  std::vector<bool> written = snapshot_set();
  for (size_t i = 0; i < R_.size(); i++) {
    if (!written[i] || R_[i].full.n == 0) continue;
    const RelData& s = R_[i].full;
    RelData& d = snap_[i];
    d.n = s.n;
    d.k.ensure(x.pool, s.n, R_[i].wide);
    SK_HIP(hipMemcpyAsync(d.k.ptr(), s.k.ptr(), s.n * (R_[i].wide ? 16 : 8), hipMemcpyDeviceToDevice, x.stream));
    if (s.w.get()) {
      d.w.reset(x.pool, s.n);
      SK_HIP(hipMemcpyAsync(d.w.get(), s.w.get(), s.n * 8, hipMemcpyDeviceToDevice, x.stream));
    }
    if (s.inf.get()) {
      d.inf.reset(x.pool, s.n);
      SK_HIP(hipMemcpyAsync(d.inf.get(), s.inf.get(), s.n, hipMemcpyDeviceToDevice, x.stream));
    }
  }
  snap_dict_ = D_;
  snap_bits_.clear();
  snap_version_.clear();
  for (auto& R : R_) {
    snap_bits_.push_back(R.bits);
    snap_version_.push_back(R.version);
  }
  SK_HIP(hipStreamSynchronize(x.stream));
  loaded = true;
}

// Relations reset() must restore: those written by rules; with aggregates
// (which may grow the dictionary and re-pack relations) every relation.
std::vector<bool> Engine::snapshot_set() const {
  bool any_agg = false;
  for (auto& r : P_.rules) any_agg |= r.agg;
  std::vector<bool> s(R_.size(), any_agg);
  for (auto& r : P_.rules)
    if (r.head.rel >= 0) s[r.head.rel] = true;
  if (!attached_.empty())
    for (size_t i = 0; i < R_.size(); i++)
      if (attached_[i]) s[i] = false;
  return s;
}

void Engine::reset() {
  if (!loaded) return;
  Exec x = ctx_->exec();
  dense_.clear();
  std::vector<bool> written = snapshot_set();
  if (snap_dict_.nvals() != D_.nvals()) {
    D_ = snap_dict_;
    int_vals_.release();
    if (!D_.ints.empty()) {
      int_vals_.reset(x.pool, D_.ints.size());
      SK_HIP(hipMemcpyAsync(int_vals_.get(), D_.ints.data(), D_.ints.size() * 8, hipMemcpyHostToDevice, x.stream));
    }
  }
  // a relation no rule writes changes in an execute only when the dictionary
  // grows under it (grow_dict re-encodes it and bumps its version): one whose
  // version is still the snapshot's holds the snapshot's content, and keeps
  // it with its version, so its indexes and column code sets stay cached
  std::vector<bool> head(R_.size(), false);
  for (auto& r : P_.rules)
    if (r.head.rel >= 0) head[r.head.rel] = true;
  for (size_t i = 0; i < R_.size(); i++) {
    DRel& R = R_[i];
    R.delta.release();
    R.delta_order.clear();
    R.dversion++;
    if (!written[i]) continue;
    if (!head[i] && i < snap_version_.size() && R.version == snap_version_[i]) continue;
    R.bits = snap_bits_[i];
    R.wide = key_bits_of(R.arity, R.bits) > 64;
    const RelData& s = snap_[i];
    RelData d;
    d.n = s.n;
    if (s.n) {
      d.k.ensure(x.pool, s.n, R.wide);
      SK_HIP(hipMemcpyAsync(d.k.ptr(), s.k.ptr(), s.n * (R.wide ? 16 : 8), hipMemcpyDeviceToDevice, x.stream));
      if (s.w.get()) {
        d.w.reset(x.pool, s.n);
        SK_HIP(hipMemcpyAsync(d.w.get(), s.w.get(), s.n * 8, hipMemcpyDeviceToDevice, x.stream));
      }
      if (s.inf.get()) {
        d.inf.reset(x.pool, s.n);
        SK_HIP(hipMemcpyAsync(d.inf.get(), s.inf.get(), s.n, hipMemcpyDeviceToDevice, x.stream));
      }
    }
    R.full = std::move(d);
    R.version++;
    R.idx.clear();
    if (i < snap_version_.size()) snap_version_[i] = R.version;
  }
  SK_HIP(hipStreamSynchronize(x.stream));
}

// Grow the dictionary with new Int values (aggregate results).  New codes
// are order-preserving.  When existing codes keep their values (every new
// Int sorts after the old ones) only relations that can still change (stratum
// >= the current one) are re-packed, and only if their column width must
// grow; otherwise every relation is re-encoded through the monotone code map,
// so packed keys stay sorted.
void Engine::grow_dict(const Dict& nd, u32 stratum) {
  Exec x = ctx_->exec();
  bool shifted = false;
  std::vector<u32> map(D_.nvals());
  for (u64 c = 0; c < D_.nvals(); c++) {
    if (c < D_.sym_span) map[c] = (u32)c;
    else {
      u32 nc = 0;
      nd.int_code(D_.ints[c - D_.sym_span], &nc);
      map[c] = nc;
      if (nc != c) shifted = true;
    }
  }
  DBuf<u32> dmap(x.pool, map.size() ? map.size() : 1);
  if (!map.empty()) SK_HIP(hipMemcpyAsync(dmap.get(), map.data(), map.size() * 4, hipMemcpyHostToDevice, x.stream));
  const int nb = nd.bits();
  std::vector<bool> live(R_.size(), false);
  for (size_t i = 0; i < R_.size(); i++) live[i] = P_.preds[i].stratum >= stratum;
  for (auto& r : P_.rules)
    if (r.stratum >= stratum && r.head.rel >= 0) live[r.head.rel] = true;
  for (size_t i = 0; i < R_.size(); i++) {
    DRel& R = R_[i];
    int ob = R.bits;
    int tb = shifted ? std::max(ob, nb) : ((live[i] && ob < nb) ? nb : ob);
    if (!shifted && tb == ob) continue;
    bool nw = key_bits_of(R.arity, tb) > 64;
    if (key_bits_of(R.arity, tb) > 128) throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, "", "packed tuple wider than 128 bits");
    if (R.full.n) {
      Keys nk;
      nk.ensure(x.pool, R.full.n, nw);
      with_key(R.wide, [&](auto* t1) {
        using K = std::remove_pointer_t<decltype(t1)>;
        with_key(nw, [&](auto* t2) {
          using KO = std::remove_pointer_t<decltype(t2)>;
          remap_keys<K, KO>(x, (const K*)R.full.k.ptr(), R.full.n, R.arity, ob, tb, dmap.get(), (KO*)nk.ptr());
        });
      });
      SK_HIP(hipStreamSynchronize(x.stream));
      R.full.k = std::move(nk);
    }
    R.bits = tb;
    R.wide = nw;
    R.delta.release();
    R.idx.clear();
    R.version++;
    R.dversion++;
  }
  D_ = nd;
  int_vals_.release();
  if (!D_.ints.empty()) {
    int_vals_.reset(x.pool, D_.ints.size());
    SK_HIP(hipMemcpyAsync(int_vals_.get(), D_.ints.data(), D_.ints.size() * 8, hipMemcpyHostToDevice, x.stream));
  }
  SK_HIP(hipStreamSynchronize(x.stream));
}

// ================================================================ indexes

Index& Engine::get_index(int rel, bool delta, const std::vector<int>& perm, bool want_csr, bool want_adj) {
  DRel& R = R_[rel];
  std::string key = delta ? "d" : "f";
  for (int p : perm) key += char('0' + p);
  Index& ix = R.idx[key];
  u64 ver = delta ? R.dversion : R.version;
  Exec x = ctx_->exec();
  const int bits = R.bits;
  bool identity = true;
  for (size_t j = 0; j < perm.size(); j++)
    if (perm[j] != (int)j) identity = false;
  // the list's own column order: key order, or a dense relation's cell order
  const bool key_sorted = !delta || R.delta_order.empty();
  const bool in_order = key_sorted ? identity : perm == R.delta_order;
  const RelData& src = source(rel, delta);
  if (ix.version != ver) {
    ix.perm = perm;
    ix.csr_version = ~0ull;
    ix.adj_version = ~0ull;
    ix.direct = identity && key_sorted;
    if (!ix.direct) {
      bool trop = R.sem == SK_TROP;
      u64 n = src.n;
      ix.data.n = n;
      ix.data.k.ensure(x.pool, n, R.wide);
      Keys alt;
      alt.ensure(x.pool, n, R.wide);
      DBuf<i64> w2;
      DBuf<u8> f2;
      if (trop) {
        ix.data.w.ensure(x.pool, n ? n : 1);
        w2.reset(x.pool, n ? n : 1);
        if (has_inf_) {
          ix.data.inf.ensure(x.pool, n ? n : 1);
          f2.reset(x.pool, n ? n : 1);
        }
        if (n) {
          SK_HIP(hipMemcpyAsync(ix.data.w.get(), src.w.get(), n * 8, hipMemcpyDeviceToDevice, x.stream));
          if (has_inf_)
            SK_HIP(hipMemcpyAsync(ix.data.inf.get(), src.inf.get(), n, hipMemcpyDeviceToDevice, x.stream));
        }
      }
      with_key(R.wide, [&](auto* tag) {
        using K = std::remove_pointer_t<decltype(tag)>;
        repack<K, K>(x, (const K*)src.k.ptr(), n, R.arity, bits, perm.data(), (K*)ix.data.k.ptr());
        if (in_order) return;  // already sorted by the index's column order
        K* sk;
        i64* sw;
        u8* sf;
        radix_sort<K>(x, (K*)ix.data.k.ptr(), (K*)alt.ptr(), trop ? ix.data.w.get() : nullptr,
                      trop ? w2.get() : nullptr, (trop && has_inf_) ? ix.data.inf.get() : nullptr,
                      (trop && has_inf_) ? f2.get() : nullptr, n, key_bits_of(R.arity, bits), &sk, &sw, &sf);
        if (sk != (K*)ix.data.k.ptr()) {
          std::swap(ix.data.k, alt);
          if (trop) {
            std::swap(ix.data.w, w2);
            if (has_inf_) std::swap(ix.data.inf, f2);
          }
        }
      });
    }
    ix.version = ver;
  }
  const RelData& d = ix.direct ? src : ix.data;
  if (want_csr && ix.csr_version != ver) {
    u64 nv = D_.nvals();
    ix.csr.ensure(x.pool, nv + 1);
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      build_csr<K>(x, (const K*)d.k.ptr(), d.n, bits * (R.arity - 1), nv, ix.csr.get());
    });
    ix.csr_version = ver;
  }
  if (want_adj && ix.adj_version != ver) {
    ix.adj.ensure(x.pool, d.n ? d.n : 1);
    DBuf<u32> first(x.pool, d.n ? d.n : 1);
    u32* cp[2] = {first.get(), ix.adj.get()};
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      unpack_cols<K>(x, (const K*)d.k.ptr(), d.n, 2, bits, cp);
    });
    SK_HIP(hipStreamSynchronize(x.stream));
    ix.adj_version = ver;
  }
  return ix;
}

// ================================================================= plans

u32 Engine::code_or_never(u32 kind, i64 v, bool* never) const {
  bool ok;
  u32 c = D_.encode(kind == SK_T_CONST ? SK_VAL_SYM : SK_VAL_INT, v, &ok);
  if (!ok) *never = true;
  return c;
}

RulePlan Engine::plan_rule(int ri, int dpos) {
  const Rule& r = P_.rules[ri];
  RulePlan pl;
  pl.rule = ri;
  pl.dpos = dpos;
  int hr = r.head.rel;
  pl.trop = hr >= 0 && P_.preds[hr].sem == SK_TROP;
  const int n = (int)r.body.size();
  const u32 cur = r.stratum;

  // literals that raise a static error when reached keep body order
  bool issue = false;
  {
    std::set<u32> bound;
    for (int j = 0; j < n; j++) {
      const Atom& a = r.body[j];
      if (a.rel < 0) issue = true;
      if (a.neg) {
        if (a.rel >= 0 && P_.preds[a.rel].stratum >= cur) issue = true;
        for (auto& t : a.args)
          if ((t.kind == SK_T_VAR && !bound.count(t.a)) || t.kind == SK_T_AGG) issue = true;
      } else {
        for (auto& t : a.args)
          if (t.kind == SK_T_VAR) bound.insert(t.a);
      }
    }
  }

  // leapfrog triangle: Bool, three positive binary atoms over three vars,
  // each pair of atoms sharing one var, no literal changing in this stratum
  if (!pl.trop && !r.agg && dpos < 0 && n == 3 && !issue) {
    bool ok = true;
    for (auto& a : r.body) {
      if (a.neg || a.args.size() != 2 || a.rel < 0 || P_.preds[a.rel].arity != 2 || in_idb_[a.rel]) ok = false;
      for (auto& t : a.args)
        if (t.kind != SK_T_VAR) ok = false;
      if (ok && a.args[0].a == a.args[1].a) ok = false;
    }
    if (ok) {
      u32 u = r.body[0].args[0].a, v = r.body[0].args[1].a;
      u32 w = ~0u;
      int lv = -1, lu = -1;
      for (int j = 1; j < 3; j++) {
        const Atom& a = r.body[j];
        u32 x0 = a.args[0].a, x1 = a.args[1].a;
        bool hasv = x0 == v || x1 == v, hasu = x0 == u || x1 == u;
        u32 other = (x0 == v || x0 == u) ? x1 : x0;
        if (hasv && !hasu) {
          lv = j;
          if (w == ~0u) w = other;
          else if (w != other) ok = false;
        } else if (hasu && !hasv) {
          lu = j;
          if (w == ~0u) w = other;
          else if (w != other) ok = false;
        } else {
          ok = false;
        }
      }
      if (ok && lv > 0 && lu > 0 && w != u && w != v) {
        pl.tri = true;
        pl.tri_lit[0] = 0;
        pl.tri_lit[1] = lv;
        pl.tri_lit[2] = lu;
        pl.tri_slot[0] = (int)u;
        pl.tri_slot[1] = (int)v;
        pl.tri_slot[2] = (int)w;
        pl.tri_col_v = r.body[lv].args[0].a == v ? 0 : 1;
        pl.tri_col_w2 = 1 - pl.tri_col_v;
        pl.tri_col_u = r.body[lu].args[0].a == u ? 0 : 1;
        pl.tri_col_w3 = 1 - pl.tri_col_u;
        // head (u, v, w) in order: the emitted rows are already sorted and unique
        const Atom& h = r.head;
        pl.sorted_out = h.args.size() == 3 && h.args[0].kind == SK_T_VAR && h.args[0].a == u &&
                        h.args[1].kind == SK_T_VAR && h.args[1].a == v && h.args[2].kind == SK_T_VAR &&
                        h.args[2].a == w;
        for (auto& t : h.args)
          if (t.kind == SK_T_VAR && t.a != u && t.a != v && t.a != w) {
            pl.head_err = SK_ERR_UNSAFE_HEAD_VAR;
            pl.head_err_slot = t.a;
            break;
          }
        if (!pl.head_err) return pl;
        pl.tri = false;
      }
    }
  }

  // evaluation order
  std::vector<int> order;
  bool reorder = !pl.trop && !issue;
  if (reorder) {
    std::vector<bool> used(n, false);
    std::set<u32> bound;
    auto add_negs = [&]() {
      for (int j = 0; j < n; j++) {
        if (used[j] || !r.body[j].neg) continue;
        bool all = true;
        for (auto& t : r.body[j].args)
          if (t.kind == SK_T_VAR && !bound.count(t.a)) all = false;
        if (all) {
          used[j] = true;
          order.push_back(j);
        }
      }
    };
    int first = dpos;
    if (first < 0)
      for (int j = 0; j < n; j++)
        if (!r.body[j].neg) {
          first = j;
          break;
        }
    if (first < 0) add_negs();
    while (first >= 0) {
      used[first] = true;
      order.push_back(first);
      for (auto& t : r.body[first].args)
        if (t.kind == SK_T_VAR) bound.insert(t.a);
      add_negs();
      first = -1;
      for (int j = 0; j < n && first < 0; j++) {
        if (used[j] || r.body[j].neg) continue;
        for (auto& t : r.body[j].args)
          if (t.kind == SK_T_VAR && bound.count(t.a)) first = j;
      }
      for (int j = 0; j < n && first < 0; j++)
        if (!used[j] && !r.body[j].neg) first = j;
    }
    for (int j = 0; j < n; j++)
      if (!used[j]) order.push_back(j);
  } else {
    for (int j = 0; j < n; j++) order.push_back(j);
  }

  pl.unit_source = order.empty() || r.body[order[0]].neg;
  std::set<u32> bound;
  std::vector<int> slots;  // current table layout (all bound vars; pruned by liveness below)
  for (size_t si = 0; si < order.size(); si++) {
    int j = order[si];
    const Atom& a = r.body[j];
    StepPlan st;
    st.lit = j;
    st.rel = a.rel;
    st.delta = (j == dpos);
    st.in_slots = slots;
    if (a.rel < 0) {
      st.err = SK_ERR_UNKNOWN_PRED;
      st.kind = a.neg ? 2 : (si == 0 ? 0 : 1);
      pl.steps.push_back(st);
      break;  // nothing after an unknown predicate can run
    }
    int arity = (int)P_.preds[a.rel].arity;
    if (a.neg) {
      st.kind = 2;
      if (P_.preds[a.rel].stratum >= cur) st.err = SK_ERR_NEGATION_NOT_STRATIFIED;
      for (auto& t : a.args) {
        if (st.err) break;
        if (t.kind == SK_T_VAR && !bound.count(t.a)) {
          st.err = SK_ERR_UNSAFE_NEG_VAR;
          st.err_slot = t.a;
        } else if (t.kind == SK_T_AGG) {
          st.err = SK_ERR_UNSUPPORTED;
        }
      }
      for (auto& t : a.args) {
        st.anti_slot.push_back(t.kind == SK_T_VAR ? (int)t.a : -1);
        st.anti_kind.push_back(t.kind);
        st.anti_val.push_back(t.kind == SK_T_CONST ? (i64)t.a : t.b);
      }
      if ((int)a.args.size() != arity) st.never = true;  // never present: the negation holds
      st.out_slots = slots;
      pl.steps.push_back(st);
      if (st.err) break;
      continue;
    }
    bool first_step = slots.empty() && si == 0 && !pl.unit_source;
    if ((int)a.args.size() != arity) st.never = true;
    for (auto& t : a.args)
      if (t.kind == SK_T_AGG) st.never = true;
    if (first_step) {
      st.kind = 0;
      std::map<u32, int> firstcol;
      for (int c = 0; c < (int)a.args.size(); c++) {
        const Term& t = a.args[c];
        if (t.kind == SK_T_VAR) {
          auto it = firstcol.find(t.a);
          if (it == firstcol.end()) {
            firstcol[t.a] = c;
            st.newvars.push_back({c, (int)t.a});
          } else {
            st.scan_eq.push_back({it->second, c});
          }
        } else if (t.kind == SK_T_CONST || t.kind == SK_T_INT) {
          st.scan_const_kind.push_back({c, t.kind});
          st.scan_const_val.push_back(t.kind == SK_T_CONST ? (i64)t.a : t.b);
        }
      }
    } else {
      st.kind = 1;
      std::vector<int> probe, rest;
      for (int c = 0; c < (int)a.args.size(); c++) {
        const Term& t = a.args[c];
        bool isprobe = (t.kind == SK_T_VAR && bound.count(t.a)) || t.kind == SK_T_CONST || t.kind == SK_T_INT;
        (isprobe ? probe : rest).push_back(c);
      }
      if (st.never) {
        probe.clear();
        rest.clear();
        for (int c = 0; c < arity; c++) rest.push_back(c);
      }
      st.perm = probe;
      st.perm.insert(st.perm.end(), rest.begin(), rest.end());
      st.nprobe = (int)probe.size();
      for (int c : probe) {
        const Term& t = a.args[c];
        st.probe_slot.push_back(t.kind == SK_T_VAR ? (int)t.a : -1);
        st.probe_kind.push_back(t.kind);
        st.probe_val.push_back(t.kind == SK_T_CONST ? (i64)t.a : t.b);
      }
      std::map<u32, int> firstidx;
      for (int q = st.nprobe; q < (int)st.perm.size() && !st.never; q++) {
        const Term& t = a.args[st.perm[q]];
        if (t.kind != SK_T_VAR) continue;
        auto it = firstidx.find(t.a);
        if (it == firstidx.end()) {
          firstidx[t.a] = q;
          st.newvars.push_back({q, (int)t.a});
        } else {
          st.filt.push_back({it->second, q});
        }
      }
    }
    for (auto& nv : st.newvars)
      if (!bound.count((u32)nv.second)) {
        bound.insert((u32)nv.second);
        slots.push_back(nv.second);
      }
    st.out_slots = slots;
    pl.steps.push_back(st);
  }

  // head checks (build_head, crates/strata-eval/src/naive.rs:526-532)
  std::set<u32> need;
  if (!r.agg) {
    for (auto& t : r.head.args) {
      if (t.kind == SK_T_VAR) {
        if (!bound.count(t.a)) {
          pl.head_err = SK_ERR_UNSAFE_HEAD_VAR;
          pl.head_err_slot = t.a;
          break;
        }
        need.insert(t.a);
      } else if (t.kind == SK_T_AGG) {
        pl.head_err = SK_ERR_UNSUPPORTED;
        break;
      }
    }
  } else {
    // aggregate: group keys and (except for count) the aggregand stay live
    for (auto& t : r.head.args) {
      if (t.kind == SK_T_VAR && bound.count(t.a)) need.insert(t.a);
      if (t.kind == SK_T_AGG && t.a != SK_AGG_COUNT && bound.count((u32)t.b)) need.insert((u32)t.b);
    }
  }
  if (pl.head_err) need.clear();

  // liveness: keep only variables used later
  std::set<u32> live = need;
  for (int s = (int)pl.steps.size() - 1; s >= 0; s--) {
    StepPlan& st = pl.steps[s];
    std::vector<int> out;
    for (int sl : st.out_slots)
      if (live.count((u32)sl)) out.push_back(sl);
    st.out_slots = out;
    if (st.kind == 1)
      for (int sl : st.probe_slot)
        if (sl >= 0) live.insert((u32)sl);
    if (st.kind == 2)
      for (int sl : st.anti_slot)
        if (sl >= 0) live.insert((u32)sl);
  }
  for (size_t s = 0; s < pl.steps.size(); s++) {
    pl.steps[s].in_slots = s == 0 ? std::vector<int>() : pl.steps[s - 1].out_slots;
    if (pl.steps[s].out_slots.size() > (size_t)kMaxCols)
      throw mkerr(SK_ERR_UNSUPPORTED_SHAPE, r.head.pred, "more than 8 live variables in a rule body");
  }
  return pl;
}

// ============================================================ plan exec

static int col_index(const std::vector<int>& slots, int slot) {
  for (size_t i = 0; i < slots.size(); i++)
    if (slots[i] == slot) return (int)i;
  return -1;
}

// Destination-side evaluation of a delta variant, when the rule has the
// pull shape and this round's push join would produce at least |E| outputs
// (a pull round reads all of E's index; measured on gfx950, a push join of
// about |E| outputs costs as much).  Returns false to run the push join
// instead.
// A round pulls when the push join would emit at least this many outputs
// (|E| scaled by SK_PULL_RATIO_PCT percent): there a pull round, which reads
// all of E's index, and the push join cost about the same.
#ifndef SK_PULL_RATIO_PCT
#define SK_PULL_RATIO_PCT 100
#endif
static u64 pull_threshold(u64 edges) { return edges / 100 * SK_PULL_RATIO_PCT + edges % 100 * SK_PULL_RATIO_PCT / 100; }

bool Engine::try_pull(const RulePlan& pl) {
  if (pl.dpos < 0 || pull_mode_ == 0) return false;
  const int hr = P_.rules[pl.rule].head.rel;
  DenseRel* dn = dense_of(hr);
  if (!dn) return false;
  PullRule* pr = nullptr;
  for (auto& p : dn->pulls)
    if (p.rule == pl.rule && p.dpos == pl.dpos) pr = &p;
  if (!pr) return false;
  DRel& H = R_[hr];
  DRel& E = R_[pr->E];
  if (H.delta.n == 0 || E.full.n == 0) return false;
  Exec x = ctx_->exec();
  const bool trop = H.sem == SK_TROP;
  Index& ixj = get_index(pr->E, false, {pr->eJ, pr->eR}, true, false);
  u64 cand = 0;
  if (pr->cost_dv == H.dversion) {
    cand = pr->cost;
  } else {
    with_key(H.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      cand = push_cost<K>(x, (const K*)H.delta.k.ptr(), H.delta.n, 2, H.bits, pr->R, ixj.csr.get());
    });
  }
  if (pull_mode_ == 1 && cand < pull_threshold(E.full.n)) return false;
  const PullDesc pd = pull_desc(*pr, *dn, hr);
  // algorithmic bytes: every index row of E (j, weight, its source's row
  // record) once, every candidate's delta entry once, and every destination
  // row's cells once.  Packed (narrow cells below 2^26): 8-byte records and
  // one 4-byte word per candidate; otherwise 16-byte records and the lane
  // byte and weight per candidate.
  const u64 drows = get_index(pr->E, false, {pr->eR, pr->eJ}, true, true).pull_rows;
  const double cell_b = trop ? (dn->cd.narrow ? 4.0 : 8.0) : 0.125;
  const double bytes = pd.dpack ? (double)E.full.n * 20.0 + (double)cand * 4.0 +
                                      (double)drows * cell_b * (double)(1u << dn->lbits)
                                : (double)E.full.n * (4.0 + (trop ? 8.0 : 0.0) + 16.0) +
                                      (double)cand * (1.0 + (trop ? 8.0 : 0.0)) +
                                      (double)drows * cell_b * (double)(1u << dn->lbits);
  pull_apply(x, pd, dn->cd, trop, bytes, cand);
  round_pulls_++;
  return true;
}

// The launch descriptor of a pull rule over relation hr's current delta.
PullDesc Engine::pull_desc(PullRule& pr, DenseRel& dn, int hr) {
  Exec x = ctx_->exec();
  DRel& H = R_[hr];
  DRel& E = R_[pr.E];
  const bool trop = H.sem == SK_TROP;
  Index& ixr = get_index(pr.E, false, {pr.eR, pr.eJ}, true, true);
  const u64 nv = D_.nvals();
  // the chunk list is an index of E, kept with it across runs
  if (ixr.pull_version != ixr.csr_version) {
    ixr.pull_n = pull_chunks(x, ixr.csr.get(), nv, &ixr.pull, &ixr.pull_rows);
    ixr.pull_version = ixr.csr_version;
  }
  PullDesc pd;
  pd.nchunks = ixr.pull_n;
  pd.chunks = ixr.pull.get();
  pd.adj = ixr.adj.get();
  pd.w = trop ? (ixr.direct ? E.full.w.get() : ixr.data.w.get()) : nullptr;
  pd.rowrec = dn.rowrec.get();
  pd.dw = trop ? H.delta.w.get() : nullptr;
  pd.dlane = dn.dlane.get();
  pd.dpack = dn.cd.dpack;
  pd.rowrec8 = dn.cd.rowrec8;
  pd.lbits = dn.lbits;
  pd.delta_first = pr.delta_first ? 1 : 0;
  pd.nrows = dn.nrows;
  return pd;
}

// The relation whose remaining Trop rounds may run device-driven, or -1:
// one derived relation, dense with capacity compaction and a pull rule, the
// stratum's only delta variant, which pulled in the round just run.
int Engine::trop_graph_rel(u32 k) {
  if (!graph_enabled_ || pull_mode_ == 0 || nparts_ || round_variants_ != 1 || round_pulls_ != 1) return -1;
  int rel = -1;
  for (size_t i = 0; i < R_.size(); i++)
    if (in_idb_[i]) {
      if (rel >= 0) return -1;
      rel = (int)i;
    }
  if (rel < 0 || R_[rel].delta.n == 0 || R_[rel].sem != SK_TROP) return -1;
  DenseRel* dn = dense_of(rel);
  if (!dn || dn->pulls.empty() || !dn->into || dn->keep_full || !dn->cd.trop || !dn->cd.pc_csr) return -1;
  int variants = 0;
  for (size_t ri = 0; ri < P_.rules.size(); ri++) {
    const Rule& r = P_.rules[ri];
    if (r.stratum != k || r.agg) continue;
    for (size_t j = 0; j < r.body.size(); j++) {
      const Atom& a = r.body[j];
      if (a.neg || a.rel < 0 || !in_idb_[a.rel]) continue;
      variants++;
      if (dn->pulls[0].rule != (int)ri || dn->pulls[0].dpos != (int)j) return -1;
    }
  }
  return variants == 1 ? rel : -1;
}

// Runs the stratum's next Trop rounds on the device: a hipGraph of
// kGraphRounds rounds (the pull, the touched compaction into the capacity
// buffers with its push cost and a status snapshot, the round's
// bookkeeping, see cells_loop_round_trop) relaunched until the rounds are
// done or leave.  Returns 0 converged, 1 the next round pushes (the delta,
// its size and push cost are in place), 2 a device status to handle (*st),
// 3 the round cap exceeded.  *iters counts the rounds run, as the host loop
// counts them.
int Engine::run_trop_graph_loop(u32 k, int rel, u64 cap, u64* iters, DevErr* st, const RulePlan* pl) {
  (void)k;
  constexpr int kGraphRounds = 4;
  DRel& H = R_[rel];
  DenseRel& dn = *dense_of(rel);
  PullRule& pr = dn.pulls[0];
  Exec x = ctx_->exec();
  PullDesc pd = pull_desc(pr, dn, rel);
  const u64 nb = (((dn.ncells + 31) / 32) + kBlock - 1) / kBlock;
  DBuf<LoopState> ls(x.pool, 1);
  DBuf<u32> cnt(x.pool, nb);
  DBuf<u64> off(x.pool, nb), pcb(x.pool, nb), tot(x.pool, 8);
  const u64 cap_left = cap > *iters ? cap - *iters : 0;
  const u64 push_limit = pull_mode_ == 2 ? 0 : pull_threshold(R_[pr.E].full.n);
  // Rounds whose push cost is below push_limit push on the device when the
  // delta variant has the capture_push shape and the delta fits the push
  // table (push_cap rows); otherwise such a round leaves for the host.
  PushCapture pc;
  const bool dev_push = pl && pull_mode_ != 2 && capture_push(*pl, &pc, false);
  const u64 push_cap = dev_push ? std::min<u64>(std::max<u64>(dn.ncells, 1), kDevPushRows) : 0;
  std::vector<DBuf<u32>> tcols;
  Table t;
  DBuf<i64> tw;
  DBuf<u64> plo, poff, pbsum, pjtot;
  DBuf<u32> pcnt;
  OutDesc o;
  if (dev_push) {
    tcols.resize(pc.tcols);
    t.ncols = pc.tcols;
    for (int c = 0; c < pc.tcols; c++) {
      tcols[c].reset(x.pool, push_cap);
      t.col[c] = tcols[c].get();
    }
    tw.reset(x.pool, push_cap);
    t.w = tw.get();
    plo.reset(x.pool, push_cap);
    poff.reset(x.pool, push_cap);
    pcnt.reset(x.pool, push_cap);
    pbsum.reset(x.pool, kDevScanBlocks);
    pjtot.reset(x.pool, 1);
    o = pc.o;
    o.cd = dn.cd;
  }
  LoopState h0{};
  h0.n = H.delta.n;
  // the first round pushes when the host's round left a delta that should
  const bool cost_ok = pr.cost_dv == H.dversion;
  h0.pc = cost_ok ? pr.cost : ~0ull;
  h0.stamp = dev_push && cost_ok && pr.cost < push_limit && H.delta.n <= push_cap ? 1u : 0u;
  SK_HIP(hipMemcpyAsync(ls.get(), &h0, sizeof h0, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  CellDesc cdl = dn.cd;
  cdl.gate = ls.get();
  pd.gate = ls.get();
  const bool hf = has_inf_;
  Exec xg = x;
  xg.timer = nullptr;  // no event records inside the capture
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  SK_HIP(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
  try {
    with_key(H.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      for (int r = 0; r < kGraphRounds; r++) {
        pull_apply(xg, pd, cdl, true, 0, 0);  // returns at once in a push round
        if (dev_push) {
          with_key(R_[pc.erel].wide, [&](auto* tage) {
            using KE = std::remove_pointer_t<decltype(tage)>;
            push_round_dev<K, KE>(xg, (const K*)H.delta.k.ptr(), H.delta.w.get(), H.arity, H.bits, pc.sd, t, pc.pd,
                                  (const KE*)pc.skeys, pc.sw, pc.csr, o, plo.get(), pcnt.get(), poff.get(),
                                  pbsum.get(), pjtot.get(), push_cap, ls.get());
          });
        }
        cells_loop_round_trop<K>(xg, cdl, dn.ncells, H.bits, (K*)H.delta.k.ptr(), H.delta.w.get(),
                                 hf ? H.delta.inf.get() : nullptr, cnt.get(), off.get(), pcb.get(), tot.get(),
                                 ctx_->derr, ls.get(), cap_left, push_limit, push_cap);
      }
    });
  } catch (...) {
    hipGraph_t bad = nullptr;
    (void)hipStreamEndCapture(x.stream, &bad);
    if (bad) (void)hipGraphDestroy(bad);
    throw;
  }
  SK_HIP(hipStreamEndCapture(x.stream, &g));
  hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    SK_HIP(ie);
  }
  LoopState hs{};
  try {
    for (;;) {
      {
        Scope sc(x.timer, "graph_rounds", KC_JOIN, 0.0, kGraphRounds);
        SK_HIP(hipGraphLaunch(ge, x.stream));
      }
      read_u64s(x, (const u64*)ls.get(), (u64*)&hs, 8);
      if (hs.done || hs.leave) break;
    }
  } catch (...) {
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    throw;
  }
  SK_HIP(hipGraphExecDestroy(ge));
  SK_HIP(hipGraphDestroy(g));
  *iters += hs.iters;
  if (hs.leave == 2) {
    *st = hs.st;
    return 2;
  }
  if (hs.leave == 3) return 3;
  // the last round's delta sits in the relation's capacity buffers
  H.dversion++;
  H.delta_order.clear();
  if (hs.done || hs.n == 0) {
    dn.spare = std::move(H.delta);
    H.delta.release();
    return 0;
  }
  H.delta.n = hs.n;
  if (!dn.key_order) H.delta_order = dn.order;
  pr.cost = hs.pc;
  pr.cost_dv = H.dversion;
  return 1;
}

// Word-parallel destination-side evaluation of a Bool delta variant (see
// BitPullRule), when the rule has that shape and the pull's bytes (every
// row-graph entry, plus the delta rows of the active share of them) are
// below the push join's (each output a random cell access), both estimated
// from the delta's size.  Returns false to run the push join instead.
bool Engine::try_bitpull(const RulePlan& pl) {
  if (pl.dpos < 0 || pull_mode_ == 0) return false;
  const int hr = P_.rules[pl.rule].head.rel;
  DenseRel* dn = dense_of(hr);
  if (!dn || dn->bpulls.empty()) return false;
  BitPullRule* bp = nullptr;
  for (auto& p : dn->bpulls)
    if (p.rule == pl.rule && p.dpos == pl.dpos) bp = &p;
  if (!bp) return false;
  DRel& H = R_[hr];
  DRel& E = R_[bp->E];
  if (H.delta.n == 0 || E.full.n == 0) return false;
  Exec x = ctx_->exec();
  ensure_bitpull_graph(*bp, *dn);
  if (bp->nedges == 0) return false;
  const double W = (double)(1ull << (dn->blbits - 5));
  const double nE = (double)bp->nedges, rows = (double)std::max<u64>(dn->brows, 1);
  const double active = std::min(1.0, (double)H.delta.n / rows);
  const double pull_b = nE * 12.0 + nE * active * 4.0 * W;
  const double push_b = (double)H.delta.n * (nE / rows) * 16.0;
  if (pull_mode_ == 1 && pull_b > push_b) return false;
  BitPullDesc bd;
  bd.nchunks = bp->nchunks;
  bd.chunks = bp->chunks.get();
  bd.graph = bp->graph.get();
  bd.dbits = dn->dbits.get();
  bd.rstamp = dn->rstamp.get();
  bd.stamp = dn->stamp;
  bd.wshift = dn->blbits - 5;
  // algorithmic bytes: every graph entry (8) and its source row's stamp (4),
  // the active entries' delta rows, and each chunk's destination row read
  // and written once
  bitpull_apply(x, bd, dn->cd, nE * 12.0 + nE * active * 4.0 * W + (double)bp->nchunks * 8.0 * W, bp->nedges);
  round_bitpulls_++;
  return true;
}

// The row graph and chunk list of a bit-pull rule, built once per version of
// its fixed relation.
void Engine::ensure_bitpull_graph(BitPullRule& bp, DenseRel& dn) {
  DRel& E = R_[bp.E];
  if (bp.ev == E.version) return;
  Exec x = ctx_->exec();
  const DRel& H = R_[P_.rules[bp.rule].head.rel];
  with_key(E.wide, [&](auto* tag) {
    using K = std::remove_pointer_t<decltype(tag)>;
    bp.nedges = bitpull_graph<K>(x, (const K*)E.full.k.ptr(), E.full.n, E.arity, E.bits, dn.cd, H.arity, bp.eh, bp.ed,
                                 dn.blbits, dn.brows, &bp.graph, &bp.csr);
  });
  bp.nchunks = bitpull_chunks(x, bp.csr.get(), dn.brows, &bp.chunks);
  bp.ev = E.version;
}

// The relation whose remaining rounds may run device-driven, or -1: the
// stratum derives one relation, kept as Bool cells, and every delta variant of the stratum's rules is one of its
// bit-pull rules, all of which took the pull in the round just run.
int Engine::graph_loop_rel(u32 k) {
  if (!graph_enabled_ || pull_mode_ == 0 || round_variants_ == 0 || round_bitpulls_ != round_variants_) return -1;
  int rel = -1;
  for (size_t i = 0; i < R_.size(); i++)
    if (in_idb_[i]) {
      if (rel >= 0) return -1;
      rel = (int)i;
    }
  if (rel < 0 || R_[rel].delta.n == 0) return -1;
  DenseRel* dn = dense_of(rel);
  if (!dn || dn->bpulls.empty() || dn->keep_full || dn->cd.trop) return -1;
  for (size_t ri = 0; ri < P_.rules.size(); ri++) {
    const Rule& r = P_.rules[ri];
    if (r.stratum != k || r.agg) continue;
    for (size_t j = 0; j < r.body.size(); j++) {
      const Atom& a = r.body[j];
      if (a.neg || a.rel < 0 || !in_idb_[a.rel]) continue;
      bool found = false;
      for (auto& bp : dn->bpulls) found = found || (bp.rule == (int)ri && bp.dpos == (int)j && bp.nchunks);
      if (!found) return -1;
    }
  }
  return rel;
}

// The push round of a delta variant whose plan is a scan of the stratum's
// Trop relation's delta, then one join to a fixed relation writing the head
// into cells, as launch descriptors (the same ones exec_plan builds), or
// false for any other shape.
bool Engine::capture_push(const RulePlan& pl, PushCapture* pc, bool parts) {
  const Rule& r = P_.rules[pl.rule];
  if (pl.tri || pl.unit_source || pl.head_err || r.agg || !pl.trop || pl.steps.size() != 2 || has_inf_) return false;
  const StepPlan& s0 = pl.steps[0];
  const StepPlan& s1 = pl.steps[1];
  if (s0.kind != 0 || !s0.delta || s0.never || s0.err || !s0.scan_const_kind.empty() || !s0.scan_eq.empty())
    return false;
  if (s1.kind != 1 || s1.delta || s1.never || s1.err || in_idb_[s1.rel] || !s1.filt.empty() || s1.nprobe < 1)
    return false;
  for (int j = 0; j < s1.nprobe; j++)
    if (s1.probe_slot[j] < 0) return false;
  const int hr = r.head.rel;
  DenseRel* dh = dense_of(hr);
  if (!dh || s0.rel != hr || R_[s1.rel].sem != SK_TROP) return false;
  ScanDesc& sd = pc->sd;
  sd = ScanDesc{};
  sd.arity = R_[s0.rel].arity;
  sd.bits = R_[s0.rel].bits;
  pc->tcols = (int)s0.out_slots.size();
  for (int c = 0; c < pc->tcols; c++)
    for (auto& nv : s0.newvars)
      if (nv.second == s0.out_slots[c]) sd.out_col[c] = nv.first;
  ProbeDesc& pd = pc->pd;
  pd = ProbeDesc{};
  pd.arity = R_[s1.rel].arity;
  pd.bits = R_[s1.rel].bits;
  pd.nprobe = s1.nprobe;
  for (int j = 0; j < s1.nprobe; j++) pd.psrc[j] = col_index(s1.in_slots, s1.probe_slot[j]);
  OutDesc& o = pc->o;
  o = OutDesc{};
  o.ncols = (int)r.head.args.size();
  o.bits = R_[hr].bits;
  for (int c = 0; c < o.ncols; c++) {
    const Term& t = r.head.args[c];
    if (t.kind != SK_T_VAR) return false;
    const int ci = col_index(s1.in_slots, (int)t.a);
    if (ci >= 0) {
      o.kind[c] = 0;
      o.idx[c] = ci;
    } else {
      o.kind[c] = 1;
      for (auto& nv : s1.newvars)
        if (nv.second == (int)t.a) o.idx[c] = nv.first;
    }
  }
  o.cd = dh->cd;
  const bool want_csr = s1.nprobe == 1;
  Index& ix = get_index(s1.rel, false, s1.perm, want_csr, false);
  const RelData& d = ix.direct ? R_[s1.rel].full : ix.data;
  pd.sn = d.n;
  pc->erel = s1.rel;
  pc->skeys = d.k.ptr();
  pc->sw = d.w.get();
  pc->csr = want_csr ? ix.csr.get() : nullptr;
  pc->part_col = parts ? col_index(s1.in_slots, (int)r.var_count - 1) : -1;
  return (!parts || pc->part_col >= 0) && pc->sw && d.n > 0;
}

// The relation of a batched Trop stratum whose remaining rounds may run on
// the device, or -1: one derived relation, dense with capacity compaction
// (program ids leading the cell order, a word never spanning programs),
// whose only delta variant in the round just run is a push round of the
// capture_push shape.
int Engine::trop_part_graph_rel(u32 k, int* rule, int* dpos) {
  if (!graph_enabled_ || !nparts_ || has_inf_ || round_variants_ != 1) return -1;
  int rel = -1;
  for (size_t i = 0; i < R_.size(); i++)
    if (in_idb_[i]) {
      if (rel >= 0) return -1;
      rel = (int)i;
    }
  if (rel < 0 || R_[rel].delta.n == 0 || R_[rel].sem != SK_TROP) return -1;
  DenseRel* dn = dense_of(rel);
  if (!dn || !dn->into || dn->keep_full || !dn->cd.trop || !cells_into_fits(dn->ncells)) return -1;
  if (dn->cd.shift[0] < 5) return -1;
  for (int c = 1; c < dn->cd.ncols; c++)
    if (dn->cd.shift[c] >= dn->cd.shift[0]) return -1;
  int found = 0;
  for (size_t ri = 0; ri < P_.rules.size(); ri++) {
    const Rule& r = P_.rules[ri];
    if (r.stratum != k || r.agg) continue;
    for (size_t j = 0; j < r.body.size(); j++) {
      const Atom& a = r.body[j];
      if (a.neg || a.rel < 0 || !in_idb_[a.rel]) continue;
      found++;
      *rule = (int)ri;
      *dpos = (int)j;
    }
  }
  return found == 1 ? rel : -1;
}

// Runs a batched Trop stratum's next rounds on the device: a hipGraph of
// kGraphRounds rounds (the push round from device-side counts,
// push_round_dev; the capacity compaction marking each program with a
// delta cell and the per-program bookkeeping, cells_loop_round_trop_parts)
// relaunched until the rounds are done or leave.  Per-program round counts
// are added to part_iters_ and active_prev is left as the device has it.
// Returns 0 converged; 2 the round's join ran but a status or a program's
// overflow stopped it before its compaction (*st; the previous delta is in
// place, n restored); 3 a program passed its cap after the round (the new
// delta is in place).
int Engine::run_trop_part_graph_loop(u32 k, int rel, const RulePlan& pl, u64* iters, std::vector<u8>& active_prev,
                                     DevErr* st) {
  constexpr int kGraphRounds = 4;
  DRel& H = R_[rel];
  DenseRel& dn = *dense_of(rel);
  PushCapture pc;
  if (!capture_push(pl, &pc)) return -1;
  Exec x = ctx_->exec();
  const u32 np = nparts_;
  const u64 cap = std::max<u64>(dn.ncells, 1);
  // the rounds' compactions write up to ncells delta entries in place; a
  // delta that a dropped program shrank (drop_dead_from_deltas sizes it to
  // its rows) moves to buffers of that capacity first
  const u64 kcap = H.wide ? H.delta.k.b.n : H.delta.k.a.n;
  if (kcap < cap || H.delta.w.n < cap) {
    RelData nd;
    nd.n = H.delta.n;
    nd.k.ensure(x.pool, cap, H.wide);
    nd.w.reset(x.pool, cap);
    if (nd.n) {
      SK_HIP(hipMemcpyAsync(nd.k.ptr(), H.delta.k.ptr(), nd.n * (H.wide ? 16 : 8), hipMemcpyDeviceToDevice,
                            x.stream));
      SK_HIP(hipMemcpyAsync(nd.w.get(), H.delta.w.get(), nd.n * 8, hipMemcpyDeviceToDevice, x.stream));
    }
    H.delta = std::move(nd);
  }
  std::vector<DBuf<u32>> tcols(pc.tcols);
  Table t;
  t.ncols = pc.tcols;
  for (int c = 0; c < pc.tcols; c++) {
    tcols[c].reset(x.pool, cap);
    t.col[c] = tcols[c].get();
  }
  DBuf<i64> tw(x.pool, cap);
  t.w = tw.get();
  t.part_col = pc.part_col;
  t.perr = part_derr_.get();
  DBuf<u64> lo(x.pool, cap), off(x.pool, cap), bsum(x.pool, kDevScanBlocks), jtot(x.pool, 1);
  DBuf<u32> cnt(x.pool, cap);
  const u64 nb = (((dn.ncells + 31) / 32) + kBlock - 1) / kBlock;
  DBuf<u32> ccnt(x.pool, nb ? nb : 1);
  DBuf<u64> coff(x.pool, nb ? nb : 1), ctot(x.pool, 8);
  DBuf<LoopState> ls(x.pool, 1);
  DBuf<u8> ap(x.pool, np), an(x.pool, np);
  DBuf<u64> pit(x.pool, np), capl(x.pool, np);
  LoopState h0{};
  h0.n = H.delta.n;
  h0.stamp = 1;  // every round pushes
  std::vector<u64> hcap(np);
  for (u32 p = 0; p < np; p++) {
    const u64 lim = part_universe_[p] + 2, used = part_iters_[p][k];
    hcap[p] = part_dead_[p] ? ~0ull : (lim > used ? lim - used : 0);
  }
  if (active_prev.size() != np) active_prev.assign(np, 0);
  SK_HIP(hipMemcpyAsync(ls.get(), &h0, sizeof h0, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipMemcpyAsync(ap.get(), active_prev.data(), np, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipMemcpyAsync(capl.get(), hcap.data(), np * 8, hipMemcpyHostToDevice, x.stream));
  memset_async(x, an.get(), 0, np);
  memset_async(x, pit.get(), 0, np * 8);
  SK_HIP(hipStreamSynchronize(x.stream));
  CellDesc cdl = dn.cd;
  cdl.gate = ls.get();
  cdl.pact = an.get();
  cdl.pact_list = cdl.dom_list[0];
  cdl.pact_wshift = cdl.shift[0] - 5;
  OutDesc o = pc.o;
  o.cd = dn.cd;
  Exec xg = x;
  xg.timer = nullptr;  // no event records inside the capture
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  SK_HIP(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
  try {
    with_key(H.wide, [&](auto* tagd) {
      using KD = std::remove_pointer_t<decltype(tagd)>;
      with_key(R_[pc.erel].wide, [&](auto* tage) {
        using KE = std::remove_pointer_t<decltype(tage)>;
        for (int r = 0; r < kGraphRounds; r++) {
          push_round_dev<KD, KE>(xg, (const KD*)H.delta.k.ptr(), H.delta.w.get(), H.arity, H.bits, pc.sd, t, pc.pd,
                                 (const KE*)pc.skeys, pc.sw, pc.csr, o, lo.get(), cnt.get(), off.get(), bsum.get(),
                                 jtot.get(), cap, ls.get());
          cells_loop_round_trop_parts<KD>(xg, cdl, dn.ncells, H.bits, (KD*)H.delta.k.ptr(), H.delta.w.get(),
                                          ccnt.get(), coff.get(), ctot.get(), ctx_->derr, ls.get(), ap.get(),
                                          an.get(), pit.get(), capl.get(), part_derr_.get(), np);
        }
      });
    });
  } catch (...) {
    hipGraph_t bad = nullptr;
    (void)hipStreamEndCapture(x.stream, &bad);
    if (bad) (void)hipGraphDestroy(bad);
    throw;
  }
  SK_HIP(hipStreamEndCapture(x.stream, &g));
  hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    SK_HIP(ie);
  }
  LoopState hl{};
  try {
    for (;;) {
      {
        Scope sc(x.timer, "graph_rounds_trop", KC_JOIN, 0.0, kGraphRounds);
        SK_HIP(hipGraphLaunch(ge, x.stream));
      }
      u64 hv[3] = {0, 0, 0};
      read_u64s(x, (const u64*)ls.get(), hv, 3);
      if ((u32)hv[0] || (u32)(hv[2] >> 32)) break;  // done, or leave
    }
  } catch (...) {
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    throw;
  }
  SK_HIP(hipGraphExecDestroy(ge));
  SK_HIP(hipGraphDestroy(g));
  std::vector<u64> pi(np);
  SK_HIP(hipMemcpyAsync(&hl, ls.get(), sizeof hl, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipMemcpyAsync(pi.data(), pit.get(), np * 8, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipMemcpyAsync(active_prev.data(), ap.get(), np, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  *iters += hl.iters;
  for (u32 p = 0; p < np; p++) part_iters_[p][k] += pi[p];
  if (hl.done) return 0;
  // the delta in the capacity buffers: the previous one (2) or the new one (3)
  H.delta.n = hl.n;
  H.dversion++;
  *st = hl.st;
  return (int)hl.leave;
}

// Runs the stratum's remaining delta rounds on the device: a hipGraph of
// kGraphRounds rounds (each: every bit pull, then the bitmap-only
// compaction and the round's bookkeeping, see cells_loop_round) is launched until the
// device's done flag is set, with one small read per launch.  Rounds after
// convergence do nothing.  Round counts (per program in batched runs) are
// the host loop's: every round that ran, including the final empty one.
void Engine::run_graph_loop(u32 k, int rel, u64* iters, std::vector<u8>& active_prev) {
  constexpr int kGraphRounds = 4;
  DRel& H = R_[rel];
  DenseRel& dn = *dense_of(rel);
  Exec x = ctx_->exec();
  std::vector<BitPullDesc> bds;
  for (auto& bp : dn.bpulls) {
    ensure_bitpull_graph(bp, dn);
    if (!bp.nchunks) continue;
    BitPullDesc bd;
    bd.nchunks = bp.nchunks;
    bd.chunks = bp.chunks.get();
    bd.graph = bp.graph.get();
    bd.dbits = dn.dbits.get();
    bd.rstamp = dn.rstamp.get();
    bd.wshift = dn.blbits - 5;
    bds.push_back(bd);
  }
  const u32 np = nparts_;
  DBuf<LoopState> ls(x.pool, 1);
  DBuf<u8> ap(x.pool, np ? np : 1), an(x.pool, np ? np : 1);
  DBuf<u64> pit(x.pool, np ? np : 1);
  LoopState h0{};
  h0.stamp = dn.stamp;
  SK_HIP(hipMemcpyAsync(ls.get(), &h0, sizeof h0, hipMemcpyHostToDevice, x.stream));
  if (np) {
    if (active_prev.size() != np) active_prev.assign(np, 0);
    SK_HIP(hipMemcpyAsync(ap.get(), active_prev.data(), np, hipMemcpyHostToDevice, x.stream));
    memset_async(x, an.get(), 0, np);
    memset_async(x, pit.get(), 0, np * 8);
  }
  SK_HIP(hipStreamSynchronize(x.stream));
  CellDesc cdl = dn.cd;
  if (np) {
    cdl.pact = an.get();
    cdl.pact_shift = cdl.shift[0] - dn.blbits;
    cdl.pact_list = cdl.dom_list[0];
  }
  for (auto& bd : bds) bd.loop = ls.get();
  Exec xg = x;
  xg.timer = nullptr;  // no event records inside the capture
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  SK_HIP(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int r = 0; r < kGraphRounds; r++) {
      for (auto& bd : bds) bitpull_apply(xg, bd, cdl, 0, 0);
      cells_loop_round(xg, cdl, dn.ncells, ls.get(), ap.get(), an.get(), pit.get(), np);
    }
  } catch (...) {
    hipGraph_t bad = nullptr;
    (void)hipStreamEndCapture(x.stream, &bad);
    if (bad) (void)hipGraphDestroy(bad);
    throw;
  }
  SK_HIP(hipStreamEndCapture(x.stream, &g));
  hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    SK_HIP(ie);
  }
  u64 hv[2] = {0, 0};
  try {
    for (;;) {
      {
        Scope sc(x.timer, "graph_rounds", KC_JOIN, 0.0, kGraphRounds);
        SK_HIP(hipGraphLaunch(ge, x.stream));
      }
      read_u64s(x, (const u64*)ls.get(), hv, 2);
      if ((u32)hv[0]) break;
    }
  } catch (...) {
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    throw;
  }
  SK_HIP(hipGraphExecDestroy(ge));
  SK_HIP(hipGraphDestroy(g));
  *iters += hv[1];
  dn.stamp = (u32)(hv[0] >> 32);
  if (np) {
    std::vector<u64> pi(np);
    SK_HIP(hipMemcpyAsync(pi.data(), pit.get(), np * 8, hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    for (u32 p = 0; p < np; p++) part_iters_[p][k] += pi[p];
    active_prev.assign(np, 0);
  }
  H.delta.release();
  H.delta_order.clear();
  H.dversion++;
}

void Engine::exec_plan(const RulePlan& pl, std::map<int, std::vector<Chunk>>& out) {
  if (try_bitpull(pl) || try_pull(pl)) return;
  const Rule& r = P_.rules[pl.rule];
  Exec x = ctx_->exec();
  const int hr = r.head.rel;
  const bool trop = pl.trop;
  DRel& H = R_[hr];
  const int hbits = H.bits;

  Chunk ch;
  HeadOwner ho;
  ho.pool = x.pool;
  ho.c = &ch;
  ho.wide = H.wide;
  ho.trop = trop;
  ho.has_inf = has_inf_ && trop;

  // head projection descriptor over a table layout (or the leapfrog slots)
  auto head_desc = [&](const std::vector<int>& slots, OutDesc* o) {
    o->ncols = (int)r.head.args.size();
    o->bits = hbits;
    for (int c = 0; c < o->ncols; c++) {
      const Term& t = r.head.args[c];
      if (t.kind == SK_T_VAR) {
        o->kind[c] = 0;
        o->idx[c] = col_index(slots, (int)t.a);
      } else {
        bool never = false;
        o->kind[c] = 2;
        o->cnst[c] = code_or_never(t.kind, t.kind == SK_T_CONST ? (i64)t.a : t.b, &never);
      }
    }
  };

  if (pl.tri) {
    const Atom& d = r.body[pl.tri_lit[0]];
    std::vector<int> pv = {pl.tri_col_v, pl.tri_col_w2};
    std::vector<int> pu = {pl.tri_col_u, pl.tri_col_w3};
    Index& iv = get_index(r.body[pl.tri_lit[1]].rel, false, pv, true, true);
    Index& iu = get_index(r.body[pl.tri_lit[2]].rel, false, pu, true, true);
    const RelData& dr = R_[d.rel].full;
    TriDesc td;
    td.bits = R_[d.rel].bits;
    td.u_col = 0;
    OutDesc o;
    std::vector<int> tslots = {pl.tri_slot[0], pl.tri_slot[1], pl.tri_slot[2]};
    head_desc(tslots, &o);
    with_key(R_[d.rel].wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      double ll = tri_list_len<K>(x, (const K*)dr.k.ptr(), dr.n, td, iv.csr.get(), iu.csr.get());
      triangles<K>(x, (const K*)dr.k.ptr(), dr.n, td, iv.csr.get(), iv.adj.get(), iu.csr.get(), iu.adj.get(), &o,
                   H.wide, false, head_alloc, &ho, &ll);
    });
    ch.sorted_unique = pl.sorted_out;
    out[hr].push_back(std::move(ch));
    return;
  }

  TableOwner cur;
  cur.pool = x.pool;
  cur.trop = trop;
  cur.has_inf = has_inf_ && trop;
  if (pl.unit_source) {
    cur.t.ncols = 0;
    table_alloc(&cur, &cur.t, 1);
    if (trop) {
      memset_async(x, cur.t.w, 0, 8);  // otimes_id = Finite(0)
      if (cur.t.inf) memset_async(x, cur.t.inf, 0, 1);
    }
  }
  bool fused = false;
  const size_t ns = pl.steps.size();
  for (size_t s = 0; s < ns; s++) {
    const StepPlan& st = pl.steps[s];
    bool reached = (s == 0 && !pl.unit_source) || cur.t.n > 0;
    if (st.err) {
      if (!reached) {
        cur.t.n = 0;
        break;
      }
      EvalErr e = mkerr(st.err, r.body[st.lit].pred);
      e.slot = st.err_slot;
      if (st.err == SK_ERR_UNKNOWN_PRED) e.msg = "unknown predicate";
      if (st.err == SK_ERR_NEGATION_NOT_STRATIFIED) e.msg = "negated predicate is not in a strictly lower stratum";
      if (st.err == SK_ERR_UNSAFE_NEG_VAR) e.msg = "unbound variable in negated literal";
      throw e;
    }
    TableOwner nxt;
    nxt.pool = x.pool;
    nxt.trop = trop;
    nxt.has_inf = has_inf_ && trop;
    nxt.t.ncols = (int)st.out_slots.size();
    const bool last = (s + 1 == ns);
    if (st.kind == 0) {
      const RelData& src = source(st.rel, st.delta);
      ScanDesc sd;
      sd.arity = R_[st.rel].arity;
      sd.bits = R_[st.rel].bits;
      bool never = st.never;
      for (size_t q = 0; q < st.scan_const_kind.size(); q++) {
        sd.const_col[sd.nconst] = st.scan_const_kind[q].first;
        sd.const_val[sd.nconst] = code_or_never(st.scan_const_kind[q].second, st.scan_const_val[q], &never);
        sd.nconst++;
      }
      for (auto& e : st.scan_eq) {
        sd.eq_a[sd.neq] = e.first;
        sd.eq_b[sd.neq] = e.second;
        sd.neq++;
      }
      for (int c = 0; c < nxt.t.ncols; c++) {
        int sl = st.out_slots[c];
        for (auto& nv : st.newvars)
          if (nv.second == sl) sd.out_col[c] = nv.first;
      }
      if (never || src.n == 0) {
        table_alloc(&nxt, &nxt.t, 0);
      } else {
        with_key(R_[st.rel].wide, [&](auto* tag) {
          using K = std::remove_pointer_t<decltype(tag)>;
          scan_rel<K>(x, (const K*)src.k.ptr(), trop ? src.w.get() : nullptr, (trop && has_inf_) ? src.inf.get() : nullptr,
                      src.n, sd, &nxt.t, table_alloc, &nxt);
        });
      }
    } else if (st.kind == 1) {
      bool never = st.never;
      ProbeDesc pd;
      pd.arity = R_[st.rel].arity;
      pd.bits = R_[st.rel].bits;
      pd.nprobe = st.nprobe;
      for (int j = 0; j < st.nprobe; j++) {
        if (st.probe_slot[j] >= 0) {
          pd.psrc[j] = col_index(st.in_slots, st.probe_slot[j]);
        } else {
          pd.psrc[j] = -1;
          pd.pconst[j] = code_or_never(st.probe_kind[j], st.probe_val[j], &never);
        }
      }
      for (auto& f : st.filt) {
        pd.fa[pd.nfilt] = f.first;
        pd.fb[pd.nfilt] = f.second;
        pd.nfilt++;
      }
      const RelData& src = source(st.rel, st.delta);
      bool static_rel = !st.delta && !in_idb_[st.rel];
      bool final_step = last && !pl.head_err && !r.agg;
      OutDesc o;
      if (final_step) {
        // fused head projection: table columns or new relation columns
        o.ncols = (int)r.head.args.size();
        o.bits = hbits;
        for (int c = 0; c < o.ncols; c++) {
          const Term& t = r.head.args[c];
          if (t.kind == SK_T_VAR) {
            int ci = col_index(st.in_slots, (int)t.a);
            if (ci >= 0) {
              o.kind[c] = 0;
              o.idx[c] = ci;
            } else {
              o.kind[c] = 1;
              for (auto& nv : st.newvars)
                if (nv.second == (int)t.a) o.idx[c] = nv.first;
            }
          } else {
            o.kind[c] = 2;
            bool nv2 = false;
            o.cnst[c] = code_or_never(t.kind, t.kind == SK_T_CONST ? (i64)t.a : t.b, &nv2);
          }
        }
        if (DenseRel* dh = dense_of(hr)) o.cd = dh->cd;
      } else {
        o.ncols = nxt.t.ncols;
        o.table = &nxt.t;
        for (int c = 0; c < o.ncols; c++) {
          int sl = st.out_slots[c];
          int ci = col_index(st.in_slots, sl);
          if (ci >= 0) {
            o.kind[c] = 0;
            o.idx[c] = ci;
          } else {
            o.kind[c] = 1;
            for (auto& nv : st.newvars)
              if (nv.second == sl) o.idx[c] = nv.first;
          }
        }
      }
      if (never || src.n == 0 || cur.t.n == 0) {
        if (final_step) head_alloc(&ho, &o, 0);
        else table_alloc(&nxt, &nxt.t, 0);
      } else {
        bool want_csr = static_rel && st.nprobe == 1 && pd.nfilt == 0;
        Index& ix = get_index(st.rel, st.delta, st.perm, want_csr, false);
        const RelData& d = ix.direct ? src : ix.data;
        pd.sn = d.n;
        Table tin = cur.t;
        if (nparts_ && trop) {
          // batched mode: an overflow is charged to the row's program
          tin.part_col = col_index(st.in_slots, (int)r.var_count - 1);
          tin.perr = part_derr_.get();
        }
        with_key(R_[st.rel].wide, [&](auto* tag) {
          using K = std::remove_pointer_t<decltype(tag)>;
          join<K>(x, tin, pd, (const K*)d.k.ptr(), trop ? d.w.get() : nullptr, (trop && has_inf_) ? d.inf.get() : nullptr,
                  want_csr ? ix.csr.get() : nullptr, &o, final_step, H.wide, trop, table_alloc, &nxt, head_alloc, &ho);
        });
      }
      if (final_step) fused = true;
    } else {
      // anti-join against the (lower-stratum, frozen) relation
      const RelData& tgt = R_[st.rel].full;
      AntiDesc ad;
      ad.arity = R_[st.rel].arity;
      ad.bits = R_[st.rel].bits;
      bool never_present = st.never;
      for (int c = 0; c < ad.arity && !st.never; c++) {
        if (st.anti_slot[c] >= 0) {
          ad.src[c] = col_index(st.in_slots, st.anti_slot[c]);
        } else {
          ad.src[c] = -1;
          ad.cnst[c] = code_or_never(st.anti_kind[c], st.anti_val[c], &never_present);
        }
      }
      ad.never_present = never_present || tgt.n == 0;
      for (int c = 0; c < nxt.t.ncols; c++) nxt.t.srccol[c] = col_index(st.in_slots, st.out_slots[c]);
      with_key(R_[st.rel].wide, [&](auto* tag) {
        using K = std::remove_pointer_t<decltype(tag)>;
        anti_join<K>(x, cur.t, ad, (const K*)tgt.k.ptr(), tgt.n, &nxt.t, table_alloc, &nxt);
      });
    }
    cur = std::move(nxt);
    if (!fused) {
      // the table's column pointers live in `cur` now
      for (int c = 0; c < cur.t.ncols; c++) cur.t.col[c] = cur.cols[c].get();
      cur.t.w = cur.w.get();
      cur.t.inf = cur.inf.get();
    }
  }

  if (!fused) {
    if (pl.head_err) {
      if (cur.t.n > 0) {
        EvalErr e = mkerr(pl.head_err, r.head.pred,
                          pl.head_err == SK_ERR_UNSAFE_HEAD_VAR ? "unbound head variable" : "aggregate in atom position");
        e.slot = pl.head_err_slot;
        throw e;
      }
      return;
    }
    OutDesc o;
    head_desc(pl.steps.empty() ? std::vector<int>() : pl.steps.back().out_slots, &o);
    if (DenseRel* dh = dense_of(hr)) o.cd = dh->cd;
    project(x, cur.t, &o, H.wide, head_alloc, &ho);
  }
  if (ch.n) out[hr].push_back(std::move(ch));
}

// ============================================================== finalize

// ⊕-combine one round's derivations of relation `rel` into it (Jacobi
// application, crates/strata-eval/src/seminaive.rs:192-208); the next delta
// is the set of changed tuples with their merged values.
u64 Engine::finalize(int rel, std::vector<Chunk>& chunks) {
  DRel& R = R_[rel];
  Exec x = ctx_->exec();
  const int bits = R.bits;
  const bool trop = R.sem == SK_TROP;
  const bool hf = trop && has_inf_;
  u64 total = 0;
  for (auto& c : chunks) total += c.n;
  // the last round's delta; a dense relation compacting into capacity
  // buffers reuses its buffers for the new one
  RelData old = std::move(R.delta);
  R.delta.release();
  R.delta_order.clear();
  R.dversion++;
  {
    DenseRel* d0 = dense_of(rel);
    if (!(d0 && d0->into)) old.release();
  }
  if (DenseRel* dn = dense_of(rel)) {
    // rows not applied by their emitting kernel (leapfrog output) go in now;
    // the delta is the set of cells this round improved, in key order
    u64 nd = 0;
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      for (auto& c : chunks)
        if (c.n) cells_insert<K>(x, dn->cd, (const K*)c.k.ptr(), trop ? c.w.get() : nullptr, hf ? c.inf.get() : nullptr,
                                 c.n, R.arity, bits);
      chunks.clear();
      DBuf<K> dk;
      DBuf<i64> dw;
      DBuf<u8> df;
      // a new stamp for the rows of the delta this compaction writes
      if (dn->cd.dbits) dn->cd.stamp = ++dn->stamp;
      // the first pull rule's push cost for the new delta comes with the count
      u64 pcost = ~0ull;
      if (!dn->pulls.empty()) {
        PullRule& p0 = dn->pulls[0];
        Index& ixj = get_index(p0.E, false, {p0.eJ, p0.eR}, true, false);
        dn->cd.pc_csr = ixj.csr.get();
        dn->cd.pc_nv = D_.nvals();
      }
      if (dn->into) {
        // no host read between the compaction's passes; the round's device
        // status may ride along with its one read
        RelData buf = old.k.ptr() ? std::move(old) : std::move(dn->spare);
        buf.k.ensure(x.pool, dn->ncells, R.wide);
        if (trop) buf.w.ensure(x.pool, dn->ncells);
        if (hf) buf.inf.ensure(x.pool, dn->ncells);
        const bool fold = fold_req_ && !fold_done_;
        DevErr st{};
        nd = cells_compact_into<K>(x, dn->cd, dn->ncells, bits, (K*)buf.k.ptr(), trop ? buf.w.get() : nullptr,
                                   hf ? buf.inf.get() : nullptr, &pcost, fold ? ctx_->derr : nullptr, &st);
        if (fold) {
          fold_done_ = true;
          folded_ = st;
        }
        if (!dn->pulls.empty() && pcost != ~0ull) {
          dn->pulls[0].cost = pcost;
          dn->pulls[0].cost_dv = R.dversion;
        }
        if (!nd) {
          dn->spare = std::move(buf);
          return;
        }
        buf.n = nd;
        R.delta = std::move(buf);
      } else {
        nd = cells_compact<K>(x, dn->cd, dn->ncells, true, bits, &dk, &dw, &df, hf, 0, &pcost);
        if (!dn->pulls.empty() && pcost != ~0ull) {
          dn->pulls[0].cost = pcost;
          dn->pulls[0].cost_dv = R.dversion;
        }
        if (!nd) return;
        R.delta.n = nd;
        R.delta.k.wide = sizeof(K) == 16;
        if constexpr (sizeof(K) == 16) R.delta.k.b = std::move(dk);
        else R.delta.k.a = std::move(dk);
        R.delta.w = std::move(dw);
        R.delta.inf = std::move(df);
      }
      if (!dn->key_order) R.delta_order = dn->order;
      if (dn->keep_full) {
        DBuf<K> u_k, f_k;
        DBuf<i64> u_w, f_w;
        DBuf<u8> u_f, f_f;
        u64 nu = 0, nf = 0;
        merge_diff<K>(x, (const K*)R.delta.k.ptr(), trop ? R.delta.w.get() : nullptr, hf ? R.delta.inf.get() : nullptr,
                      nd, (const K*)R.full.k.ptr(), trop ? R.full.w.get() : nullptr, hf ? R.full.inf.get() : nullptr,
                      R.full.n, &u_k, &u_w, &u_f, &nu, &f_k, &f_w, &f_f, &nf, trop, hf);
        // every delta row strictly improves its cell, so all of them are
        // fresh and the union is always written
        RelData nfull;
        nfull.n = nu;
        nfull.k.wide = sizeof(K) == 16;
        if constexpr (sizeof(K) == 16) nfull.k.b = std::move(u_k);
        else nfull.k.a = std::move(u_k);
        nfull.w = std::move(u_w);
        nfull.inf = std::move(u_f);
        R.full = std::move(nfull);
        R.version++;
      }
    });
    return nd;
  }
  if (total == 0) return 0;
  Chunk cand;
  if (chunks.size() == 1) {
    cand = std::move(chunks[0]);
  } else {
    cand.k.ensure(x.pool, total, R.wide);
    if (trop) {
      cand.w.reset(x.pool, total);
      if (hf) cand.inf.reset(x.pool, total);
    }
    u64 off = 0;
    size_t kb = R.wide ? 16 : 8;
    for (auto& c : chunks) {
      if (!c.n) continue;
      SK_HIP(hipMemcpyAsync((char*)cand.k.ptr() + off * kb, c.k.ptr(), c.n * kb, hipMemcpyDeviceToDevice, x.stream));
      if (trop) SK_HIP(hipMemcpyAsync(cand.w.get() + off, c.w.get(), c.n * 8, hipMemcpyDeviceToDevice, x.stream));
      if (hf) SK_HIP(hipMemcpyAsync(cand.inf.get() + off, c.inf.get(), c.n, hipMemcpyDeviceToDevice, x.stream));
      off += c.n;
    }
    cand.n = total;
    cand.sorted_unique = false;
  }
  chunks.clear();
  u64 nfresh = 0;
  with_key(R.wide, [&](auto* tag) {
    using K = std::remove_pointer_t<decltype(tag)>;
    K* ck = (K*)cand.k.ptr();
    i64* cw = trop ? cand.w.get() : nullptr;
    u8* cf = hf ? cand.inf.get() : nullptr;
    u64 cn = cand.n;
    Keys uk, alt;
    DBuf<i64> uw, w2;
    DBuf<u8> uf, f2;
    if (!cand.sorted_unique) {
      alt.ensure(x.pool, cn, R.wide);
      if (trop) w2.reset(x.pool, cn);
      if (hf) f2.reset(x.pool, cn);
      K* sk;
      i64* sw;
      u8* sf;
      radix_sort<K>(x, ck, (K*)alt.ptr(), cw, trop ? w2.get() : nullptr, cf, hf ? f2.get() : nullptr, cn,
                    key_bits_of(R.arity, bits), &sk, &sw, &sf);
      uk.ensure(x.pool, cn, R.wide);
      if (trop) uw.reset(x.pool, cn);
      if (hf) uf.reset(x.pool, cn);
      cn = unique_runs<K>(x, sk, trop ? sw : nullptr, hf ? sf : nullptr, cn, (K*)uk.ptr(), trop ? uw.get() : nullptr,
                          hf ? uf.get() : nullptr);
      ck = (K*)uk.ptr();
      cw = trop ? uw.get() : nullptr;
      cf = hf ? uf.get() : nullptr;
    }
    if (R.full.n == 0) {
      // full is empty: every candidate is fresh and the union is the
      // candidate set.  Full takes the candidate buffers and the delta is a
      // view of them: nothing writes a delta in place, and the next round
      // reads it before any merge replaces full (both on this stream).
      RelData nf;
      nf.n = cn;
      if (!cand.sorted_unique) {
        nf.k = std::move(uk);
        nf.w = std::move(uw);
        nf.inf = std::move(uf);
      } else {
        nf.k = std::move(cand.k);
        nf.w = std::move(cand.w);
        nf.inf = std::move(cand.inf);
      }
      R.full = std::move(nf);
      R.delta = view_of(R.full);
      nfresh = cn;
      return;
    }
    DBuf<K> u_k, f_k;
    DBuf<i64> u_w, f_w;
    DBuf<u8> u_f, f_f;
    u64 nu = 0;
    merge_diff<K>(x, ck, cw, cf, cn, (const K*)R.full.k.ptr(), trop ? R.full.w.get() : nullptr,
                  hf ? R.full.inf.get() : nullptr, R.full.n, &u_k, &u_w, &u_f, &nu, &f_k, &f_w, &f_f, &nfresh, trop, hf);
    if (nfresh) {
      RelData nf;
      nf.n = nu;
      if constexpr (sizeof(K) == 16) {
        nf.k.wide = true;
        nf.k.b = std::move(u_k);
      } else {
        nf.k.wide = false;
        nf.k.a = std::move(u_k);
      }
      nf.w = std::move(u_w);
      nf.inf = std::move(u_f);
      R.full = std::move(nf);
      R.delta.n = nfresh;
      if constexpr (sizeof(K) == 16) {
        R.delta.k.wide = true;
        R.delta.k.b = std::move(f_k);
      } else {
        R.delta.k.wide = false;
        R.delta.k.a = std::move(f_k);
      }
      R.delta.w = std::move(f_w);
      R.delta.inf = std::move(f_f);
    }
  });
  if (nfresh) R.version++;
  R.delta.n = nfresh;
  return nfresh;
}

DevErr Engine::read_device_status() {
  Exec x = ctx_->exec();
  DevErr h{};
  SK_HIP(hipMemcpyAsync(&h, ctx_->derr, sizeof(DevErr), hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  check_fault(x);
  return h;
}

void Engine::check_device_error() { raise_device_error(read_device_status()); }

void Engine::raise_device_error(const DevErr& h) {
  Exec x = ctx_->exec();
  if (h.code == SK_ERR_OVERFLOW) {
    DevErr z{};
    SK_HIP(hipMemcpyAsync(ctx_->derr, &z, sizeof(DevErr), hipMemcpyHostToDevice, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    EvalErr e = mkerr(SK_ERR_OVERFLOW, "", "tropical weight overflow");
    e.lhs = h.lhs;
    e.rhs = h.rhs;
    throw e;
  }
}

// ============================================================ aggregates

// eval_aggregate_rule + fold_aggregate (crates/strata-eval/src/naive.rs:300-403).
void Engine::eval_aggregate(const Rule& r) {
  for (auto& b : r.body) {
    if (b.rel < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, b.pred, "unknown predicate");
    if (P_.preds[b.rel].stratum >= r.stratum)
      throw mkerr(SK_ERR_AGGREGATE_IN_RECURSION, b.pred, "aggregate body predicate is not in a strictly lower stratum");
  }
  int agg_index = -1;
  u32 op = 0, slot = 0;
  for (size_t i = 0; i < r.head.args.size(); i++) {
    const Term& t = r.head.args[i];
    if (t.kind == SK_T_AGG) {
      if (agg_index >= 0) throw mkerr(SK_ERR_MALFORMED_AGGREGATE_HEAD, "", "aggregate head must contain exactly one agg term");
      agg_index = (int)i;
      op = t.a;
      slot = (u32)t.b;
    }
  }
  if (agg_index < 0) throw mkerr(SK_ERR_MALFORMED_AGGREGATE_HEAD, "", "aggregate head must contain exactly one agg term");
  if (slot >= r.var_count) throw mkerr(SK_ERR_UNSUPPORTED, "", "aggregate slot out of range");
  int ri = (int)(&r - &P_.rules[0]);
  RulePlan pl = plan_rule(ri, -1);
  Exec x = ctx_->exec();
  const int bits = D_.bits();  // group keys are packed with the dictionary width

  // run the body plan to a table
  std::map<int, std::vector<Chunk>> dummy;
  // Execute steps (no head projection): reuse exec_plan's machinery by
  // walking the steps here.
  TableOwner cur;
  cur.pool = x.pool;
  if (pl.unit_source) {
    cur.t.ncols = 0;
    table_alloc(&cur, &cur.t, 1);
  }
  for (size_t s = 0; s < pl.steps.size(); s++) {
    const StepPlan& st = pl.steps[s];
    bool reached = (s == 0 && !pl.unit_source) || cur.t.n > 0;
    if (st.err) {
      if (!reached) {
        cur.t.n = 0;
        break;
      }
      EvalErr e = mkerr(st.err, r.body[st.lit].pred);
      e.slot = st.err_slot;
      throw e;
    }
    TableOwner nxt;
    nxt.pool = x.pool;
    nxt.t.ncols = (int)st.out_slots.size();
    if (st.kind == 0) {
      const RelData& src = R_[st.rel].full;
      ScanDesc sd;
      sd.arity = R_[st.rel].arity;
      sd.bits = R_[st.rel].bits;
      bool never = st.never;
      for (size_t q = 0; q < st.scan_const_kind.size(); q++) {
        sd.const_col[sd.nconst] = st.scan_const_kind[q].first;
        sd.const_val[sd.nconst] = code_or_never(st.scan_const_kind[q].second, st.scan_const_val[q], &never);
        sd.nconst++;
      }
      for (auto& e : st.scan_eq) {
        sd.eq_a[sd.neq] = e.first;
        sd.eq_b[sd.neq] = e.second;
        sd.neq++;
      }
      for (int c = 0; c < nxt.t.ncols; c++)
        for (auto& nv : st.newvars)
          if (nv.second == st.out_slots[c]) sd.out_col[c] = nv.first;
      if (never || src.n == 0) {
        table_alloc(&nxt, &nxt.t, 0);
      } else if (nxt.t.ncols == 0 && sd.nconst == 0 && sd.neq == 0) {
        // no live columns and no filter: the table is just a row count
        table_alloc(&nxt, &nxt.t, 0);
        nxt.t.n = src.n;
      } else {
        with_key(R_[st.rel].wide, [&](auto* tag) {
          using K = std::remove_pointer_t<decltype(tag)>;
          scan_rel<K>(x, (const K*)src.k.ptr(), nullptr, nullptr, src.n, sd, &nxt.t, table_alloc, &nxt);
        });
      }
    } else if (st.kind == 1) {
      bool never = st.never;
      ProbeDesc pd;
      pd.arity = R_[st.rel].arity;
      pd.bits = R_[st.rel].bits;
      pd.nprobe = st.nprobe;
      for (int j = 0; j < st.nprobe; j++) {
        if (st.probe_slot[j] >= 0) pd.psrc[j] = col_index(st.in_slots, st.probe_slot[j]);
        else {
          pd.psrc[j] = -1;
          pd.pconst[j] = code_or_never(st.probe_kind[j], st.probe_val[j], &never);
        }
      }
      for (auto& f : st.filt) {
        pd.fa[pd.nfilt] = f.first;
        pd.fb[pd.nfilt] = f.second;
        pd.nfilt++;
      }
      OutDesc o;
      o.ncols = nxt.t.ncols;
      o.table = &nxt.t;
      for (int c = 0; c < o.ncols; c++) {
        int sl = st.out_slots[c];
        int ci = col_index(st.in_slots, sl);
        if (ci >= 0) {
          o.kind[c] = 0;
          o.idx[c] = ci;
        } else {
          o.kind[c] = 1;
          for (auto& nv : st.newvars)
            if (nv.second == sl) o.idx[c] = nv.first;
        }
      }
      const RelData& src = R_[st.rel].full;
      if (never || src.n == 0 || cur.t.n == 0) {
        table_alloc(&nxt, &nxt.t, 0);
      } else {
        bool want_csr = st.nprobe == 1 && pd.nfilt == 0;
        Index& ix = get_index(st.rel, false, st.perm, want_csr, false);
        bool identity = true;
        for (size_t q = 0; q < st.perm.size(); q++)
          if (st.perm[q] != (int)q) identity = false;
        const RelData& d = identity ? src : ix.data;
        pd.sn = d.n;
        with_key(R_[st.rel].wide, [&](auto* tag) {
          using K = std::remove_pointer_t<decltype(tag)>;
          join<K>(x, cur.t, pd, (const K*)d.k.ptr(), nullptr, nullptr, want_csr ? ix.csr.get() : nullptr, &o, false,
                  false, false, table_alloc, &nxt, head_alloc, nullptr);
        });
      }
    } else {
      const RelData& tgt = R_[st.rel].full;
      AntiDesc ad;
      ad.arity = R_[st.rel].arity;
      ad.bits = R_[st.rel].bits;
      bool never_present = st.never;
      for (int c = 0; c < ad.arity && !st.never; c++) {
        if (st.anti_slot[c] >= 0) ad.src[c] = col_index(st.in_slots, st.anti_slot[c]);
        else {
          ad.src[c] = -1;
          ad.cnst[c] = code_or_never(st.anti_kind[c], st.anti_val[c], &never_present);
        }
      }
      ad.never_present = never_present || tgt.n == 0;
      for (int c = 0; c < nxt.t.ncols; c++) nxt.t.srccol[c] = col_index(st.in_slots, st.out_slots[c]);
      with_key(R_[st.rel].wide, [&](auto* tag) {
        using K = std::remove_pointer_t<decltype(tag)>;
        anti_join<K>(x, cur.t, ad, (const K*)tgt.k.ptr(), tgt.n, &nxt.t, table_alloc, &nxt);
      });
    }
    u64 keep_n = nxt.t.n;
    cur = std::move(nxt);
    for (int c = 0; c < cur.t.ncols; c++) cur.t.col[c] = cur.cols[c].get();
    cur.t.w = nullptr;
    cur.t.inf = nullptr;
    cur.t.n = keep_n;
  }
  u64 nsol = cur.t.n;
  if (nsol == 0) return;
  // unbound group key / aggregand (naive.rs:350-362)
  std::set<int> boundset(pl.steps.empty() ? std::set<int>() : std::set<int>());
  {
    std::set<u32> b;
    for (auto& a : r.body)
      if (!a.neg)
        for (auto& t : a.args)
          if (t.kind == SK_T_VAR) b.insert(t.a);
    for (size_t i = 0; i < r.head.args.size(); i++) {
      if ((int)i == agg_index) continue;
      const Term& t = r.head.args[i];
      if (t.kind == SK_T_VAR && !b.count(t.a)) {
        EvalErr e = mkerr(SK_ERR_UNSAFE_HEAD_VAR, r.head.pred, "unbound head variable");
        e.slot = t.a;
        throw e;
      }
    }
    if (!b.count(slot)) {
      EvalErr e = mkerr(SK_ERR_UNSAFE_HEAD_VAR, r.head.pred, "unbound head variable");
      e.slot = slot;
      throw e;
    }
  }
  if (op == SK_AGG_PROB_OR) throw mkerr(SK_ERR_UNSUPPORTED, "", "prob-or aggregate");
  // group keys and values
  std::vector<std::vector<u32>> gkeys;  // per group: key codes
  std::vector<i64> gvals;
  std::vector<int> keycols;
  AggDesc ad;
  ad.bits = bits;
  ad.op = (int)op;
  ad.first_int = (u32)D_.sym_span;
  ad.n_ints = (u32)D_.ints.size();
  std::vector<int> last_slots = pl.steps.empty() ? std::vector<int>() : pl.steps.back().out_slots;
  for (size_t i = 0; i < r.head.args.size(); i++) {
    if ((int)i == agg_index) continue;
    const Term& t = r.head.args[i];
    if (t.kind == SK_T_VAR) {
      ad.key_kind[ad.nkey] = 0;
      ad.key_idx[ad.nkey] = col_index(last_slots, (int)t.a);
    } else {
      bool nv = false;
      ad.key_kind[ad.nkey] = 2;
      ad.key_cnst[ad.nkey] = code_or_never(t.kind, t.kind == SK_T_CONST ? (i64)t.a : t.b, &nv);
    }
    ad.nkey++;
  }
  ad.val_col = op == SK_AGG_COUNT ? -1 : col_index(last_slots, (int)slot);
  int kbits = ad.nkey * bits;
  bool kwide = kbits > 64;
  if (ad.nkey == 0 && op == SK_AGG_COUNT) {
    gkeys.push_back({});
    gvals.push_back((i64)nsol);
  } else {
    bool bad = false;
    auto run = [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      std::vector<K> ks;
      std::vector<i64> vs;
      aggregate<K>(x, cur.t, ad, int_vals_.get(), kbits, &ks, &vs, &bad);
      for (size_t g = 0; g < ks.size(); g++) {
        std::vector<u32> kc(ad.nkey);
        K kk = ks[g];
        for (int c = ad.nkey - 1; c >= 0; c--) {
          kc[c] = (u32)(kk & (((K)1 << bits) - 1));
          kk >>= bits;
        }
        gkeys.push_back(kc);
        gvals.push_back(vs[g]);
      }
    };
    if (kwide) run((u128*)nullptr);
    else run((u64*)nullptr);
    if (bad && op != SK_AGG_COUNT) throw mkerr(SK_ERR_NON_INTEGER_AGGREGAND, "", "min/max/sum require an integer aggregand");
  }
  // new Int values enter the dictionary (order-preserving remap)
  Dict nd = D_;
  std::set<i64> all(nd.ints.begin(), nd.ints.end());
  size_t before = all.size();
  for (i64 v : gvals) all.insert(v);
  std::vector<u32> old_to_new;
  if (all.size() != before) {
    nd.ints.assign(all.begin(), all.end());
    Dict od = D_;
    grow_dict(nd, r.stratum);
    for (auto& k : gkeys)
      for (auto& c : k)
        if (c >= od.sym_span) {
          u32 cc = 0;
          nd.int_code(od.ints[c - od.sym_span], &cc);
          c = cc;
        }
  }
  // insert the group tuples (Db::insert with Unit)
  int hr = r.head.rel;
  if (hr < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, r.head.pred, "insert into unknown relation");
  DRel& H = R_[hr];
  const int nb = H.bits;
  std::vector<u32> hc(gkeys.size() * H.arity);
  for (size_t g = 0; g < gkeys.size(); g++) {
    int ki = 0;
    for (int c = 0; c < H.arity; c++) {
      if (c == agg_index) {
        u32 cc = 0;
        D_.int_code(gvals[g], &cc);
        hc[g * H.arity + c] = cc;
      } else {
        hc[g * H.arity + c] = gkeys[g][ki++];
      }
    }
  }
  u64 ng = gkeys.size();
  std::vector<DBuf<u32>> dc(H.arity);
  Cols cols{};
  for (int c = 0; c < H.arity; c++) {
    std::vector<u32> colv(ng);
    for (u64 g = 0; g < ng; g++) colv[g] = hc[g * H.arity + c];
    dc[c].reset(x.pool, ng ? ng : 1);
    SK_HIP(hipMemcpyAsync(dc[c].get(), colv.data(), ng * 4, hipMemcpyHostToDevice, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    cols.c[c] = dc[c].get();
  }
  Chunk chk;
  chk.k.ensure(x.pool, ng, H.wide);
  chk.n = ng;
  with_key(H.wide, [&](auto* tag) {
    using K = std::remove_pointer_t<decltype(tag)>;
    pack_cols<K>(x, cols, H.arity, nb, ng, (K*)chk.k.ptr());
  });
  std::vector<Chunk> v;
  v.push_back(std::move(chk));
  finalize(hr, v);
  H.delta.release();
  H.delta.n = 0;
}

// ================================================================ strata

// ============================================================ dense cells

static int ceil_log2(u64 n) {
  int b = 0;
  while (b < 63 && (1ull << b) < n) b++;
  return b;
}

// Decide which relations of stratum k run as dense cells, and build their
// column domains.  A head column takes its values from head constants, the
// relation's own facts and the body columns that bind its variable: a fixed
// column (EDB, or a relation of a lower stratum) bounds it by that column's
// distinct codes (the smallest such column is taken); otherwise it inherits
// the domain of a column of the stratum (the union over everything reachable
// in this provenance graph).  A domain covering more than half of the codes
// becomes the identity.  Relations whose cells would exceed a quarter of the
// memory budget (at most 16 GiB) keep the sorted-list path.
void Engine::plan_dense(u32 k) {
  dense_.clear();
  dense_.resize(R_.size());
  if (!dense_enabled_) return;
  std::vector<bool> cand(R_.size(), false);
  for (size_t i = 0; i < R_.size(); i++) {
    if (!in_idb_[i] || R_[i].arity > kMaxCols) continue;
    if (i < attached_.size() && attached_[i]) continue;
    cand[i] = R_[i].sem == SK_TROP || R_[i].sem == SK_BOOL;
  }
  for (auto& r : P_.rules)
    if (r.stratum == k && r.agg && r.head.rel >= 0) cand[r.head.rel] = false;
  bool any = false;
  for (size_t i = 0; i < R_.size(); i++) any = any || cand[i];
  if (!any) return;

  const u64 nvals = D_.nvals();
  if (nvals > (1ull << 31)) return;  // domain bitmaps would not be modest
  Exec x = ctx_->exec();
  const u64 nw = (nvals + 31) / 32;

  // distinct codes of a column (kept across executes while the relation's
  // version and the dictionary are unchanged, like the relation indexes)
  auto col_codes = [&](int rel, int col) -> const ColCodes* {
    ColCodes& cc = colcodes_[{rel, col}];
    const DRel& Rr = R_[rel];
    if (cc.version == Rr.version && cc.nvals == nvals && cc.bm.get()) return &cc;
    cc.bm.reset(x.pool, nw);
    memset_async(x, cc.bm.get(), 0, nw * 4);
    const RelData& d = Rr.full;
    with_key(Rr.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      cells_mark_col<K>(x, (const K*)d.k.ptr(), d.n, Rr.arity, Rr.bits, col, cc.bm.get());
    });
    cc.count = d.n ? popcount_bits(x, cc.bm.get(), nw) : 0;
    cc.rel = rel;
    cc.col = col;
    cc.version = Rr.version;
    cc.nvals = nvals;
    return &cc;
  };

  struct Node {
    bool all = false;
    std::vector<const ColCodes*> fixed;
    std::vector<u32> consts;
    std::vector<std::pair<int, int>> deps;
  };
  std::map<std::pair<int, int>, Node> g;
  std::vector<bool> keep_full(R_.size(), false);
  for (auto& r : P_.rules) {
    if (r.stratum != k || r.agg || r.head.rel < 0) continue;
    int nidb = 0;
    for (auto& a : r.body)
      if (!a.neg && a.rel >= 0 && in_idb_[a.rel]) nidb++;
    if (nidb >= 2)
      for (auto& a : r.body)
        if (!a.neg && a.rel >= 0 && in_idb_[a.rel]) keep_full[a.rel] = true;
    if (!cand[r.head.rel]) continue;
    for (size_t j = 0; j < r.head.args.size(); j++) {
      Node& nd = g[{r.head.rel, (int)j}];
      const Term& t = r.head.args[j];
      if (t.kind == SK_T_CONST || t.kind == SK_T_INT) {
        bool never = false;
        u32 c = code_or_never(t.kind, t.kind == SK_T_CONST ? (i64)t.a : t.b, &never);
        if (never) nd.all = true;
        else nd.consts.push_back(c);
        continue;
      }
      if (t.kind != SK_T_VAR) {
        nd.all = true;
        continue;
      }
      // fixed occurrences by row count: a column with no more rows than the
      // best distinct count so far cannot beat it and is not scanned
      std::vector<std::pair<u64, std::pair<int, int>>> occ;
      std::pair<int, int> dep{-1, -1};
      for (auto& a : r.body) {
        if (a.neg || a.rel < 0 || (int)a.args.size() != R_[a.rel].arity) continue;
        for (size_t c = 0; c < a.args.size(); c++) {
          if (a.args[c].kind != SK_T_VAR || a.args[c].a != t.a) continue;
          if (in_idb_[a.rel]) {
            if (dep.first < 0) dep = {a.rel, (int)c};
            continue;
          }
          occ.push_back({R_[a.rel].full.n, {a.rel, (int)c}});
        }
      }
      std::sort(occ.begin(), occ.end());
      const ColCodes* best = nullptr;
      for (auto& o : occ) {
        if (best && o.first >= best->count) break;
        const ColCodes* cc = col_codes(o.second.first, o.second.second);
        if (!best || cc->count < best->count) best = cc;
      }
      if (best) nd.fixed.push_back(best);
      else if (dep.first >= 0) nd.deps.push_back(dep);
      else nd.all = true;
    }
  }

  const u64 cap_bytes = std::min<u64>(ctx_->pool->budget() / 4, 16ull << 30);
  u64 used_bytes = 0;
  const u64 into_cap = std::min<u64>(ctx_->pool->budget() / 8, 8ull << 30);
  u64 into_used = 0;
  for (size_t i = 0; i < R_.size(); i++) {
    if (!cand[i]) continue;
    DRel& R = R_[i];
    const bool trop = R.sem == SK_TROP;
    auto dn = std::make_unique<DenseRel>();
    dn->keep_full = keep_full[i];
    CellDesc& cd = dn->cd;
    cd.trop = trop;
    cd.ncols = R.arity;
    cd.ncodes = nvals;
    cd.nparts = nparts_;
    int tot_bits = 0;
    bool ok = true;
    for (int j = 0; j < R.arity && ok; j++) {
      std::vector<std::pair<int, int>> stack{{(int)i, j}};
      std::set<std::pair<int, int>> seen{{(int)i, j}};
      bool all = false;
      std::vector<const ColCodes*> parts;  // column-code bitmaps to unite
      std::vector<u32> consts;
      while (!stack.empty() && !all) {
        auto at = stack.back();
        stack.pop_back();
        auto it = g.find(at);
        if (it == g.end() || !cand[at.first]) {
          all = true;  // a stratum column whose domain is not tracked
          break;
        }
        const Node& nd = it->second;
        all = nd.all;
        for (const ColCodes* f : nd.fixed) parts.push_back(f);
        for (u32 c : nd.consts) consts.push_back(c);
        if (R_[at.first].full.n) parts.push_back(col_codes(at.first, at.second));
        for (auto& d : nd.deps)
          if (seen.insert(d).second) stack.push_back(d);
      }
      if (all) {
        cd.ebits[j] = R.bits;
      } else {
        DBuf<u32> bm(x.pool, nw);
        memset_async(x, bm.get(), 0, nw * 4);
        for (const ColCodes* pc : parts) cells_or(x, bm.get(), pc->bm.get(), nw);
        for (u32 c : consts) cells_mark_code(x, bm.get(), c);
        u64 cnt = popcount_bits(x, bm.get(), nw);
        if (2 * cnt > nvals) {
          cd.ebits[j] = R.bits;
        } else {
          dn->dom_pre[j].reset(x.pool, nw);
          cells_domain(x, bm.get(), nw, dn->dom_pre[j].get(), &dn->dom_list[j]);
          dn->dom_bits[j] = std::move(bm);
          cd.dom_bits[j] = dn->dom_bits[j].get();
          cd.dom_pre[j] = dn->dom_pre[j].get();
          cd.dom_list[j] = dn->dom_list[j].get();
          cd.ebits[j] = ceil_log2(cnt ? cnt : 1);
        }
        // batched runs: a column whose values differ from program to
        // program (each program's own sources, say) is ranked within its
        // program when that needs fewer bits
        if (nparts_ && j > 0 && consts.empty() && !parts.empty()) {
          std::vector<PartCol> pcs;
          for (const ColCodes* pc : parts) {
            const DRel& F = R_[pc->rel];
            pcs.push_back(PartCol{F.full.k.ptr(), F.wide, F.full.n, F.arity, F.bits, pc->col});
          }
          DBuf<u64> poff;
          DBuf<u32> pcodes;
          const u64 mx = cells_part_domain(x, pcs, nparts_, &poff, &pcodes);
          const int pb = ceil_log2(mx ? mx : 1);
          if (pb < cd.ebits[j]) {
            dn->dom_bits[j].release();
            dn->dom_pre[j].release();
            dn->dom_list[j].release();
            cd.dom_bits[j] = nullptr;
            cd.dom_pre[j] = nullptr;
            cd.dom_list[j] = nullptr;
            dn->pdom_off[j] = std::move(poff);
            dn->pdom_codes[j] = std::move(pcodes);
            cd.pdom_off[j] = dn->pdom_off[j].get();
            cd.pdom_codes[j] = dn->pdom_codes[j].get();
            cd.ebits[j] = pb;
          }
        }
      }
      tot_bits += cd.ebits[j];
      if (tot_bits > 40) ok = false;
    }
    if (!ok) continue;
    // Cell order: when a rule scans this relation's delta and joins it on a
    // variable held in column c, column c goes first, so the next delta (the
    // touched cells, compacted in cell order) groups the join's driver rows by
    // that column: rows sharing it read the same relation range and, with
    // this layout, touch neighbouring cells.  Otherwise key order.
    dn->order.resize(R.arity);
    for (int j = 0; j < R.arity; j++) dn->order[j] = j;
    // (batched runs keep the program id, column 0, first: their deltas are
    // split by it; the join column goes second)
    const size_t first = nparts_ ? 1 : 0;
    if (!dn->keep_full && R.arity >= (int)first + 2) {
      int lead = -1;
      for (auto& r : P_.rules) {
        if (r.stratum != k || r.agg || lead >= 0) continue;
        for (size_t j = 0; j < r.body.size() && lead < 0; j++) {
          const Atom& a = r.body[j];
          if (a.neg || a.rel != (int)i) continue;
          for (size_t c = first; c < a.args.size() && lead < 0; c++) {
            if (a.args[c].kind != SK_T_VAR) continue;
            for (size_t q = 0; q < r.body.size() && lead < 0; q++) {
              if (q == j || r.body[q].neg) continue;
              for (auto& t : r.body[q].args)
                if (t.kind == SK_T_VAR && t.a == a.args[c].a) lead = (int)c;
            }
          }
        }
      }
      if (lead > (int)first) {
        dn->order.erase(dn->order.begin() + lead);
        dn->order.insert(dn->order.begin() + first, lead);
      }
    }
    dn->key_order = true;
    for (int j = 0; j < R.arity; j++)
      if (dn->order[j] != j) dn->key_order = false;
    for (int q = R.arity - 1, sh = 0; q >= 0; q--) {
      cd.shift[dn->order[q]] = sh;
      sh += cd.ebits[dn->order[q]];
    }
    dn->ncells = 1ull << tot_bits;
    const u64 nc = std::max<u64>(dn->ncells, 32);
    const u64 nwc = nc / 32;
    const bool narrow = trop && narrow_ok_ && !narrow_off_;
    const u64 bytes = trop ? nc * (narrow ? 4 : 8) + nwc * 4 * (has_inf_ ? 3 : 2) : nwc * 4 * 2;
    if (used_bytes + bytes > cap_bytes) continue;
    used_bytes += bytes;
    // capacity buffers for the delta (ncells records) when they fit too
    const u64 into_b = dn->ncells * ((R.wide ? 16 : 8) + (trop ? 8 : 0) + (trop && has_inf_ ? 1 : 0));
    if (cells_into_fits(dn->ncells) && into_used + into_b <= into_cap) {
      dn->into = true;
      into_used += into_b;
    }
    if (narrow) {
      memset_async(x, &ctx_->derr->widen, 0, 4);
      dn->cell32.reset(x.pool, nc);
      memset_async(x, dn->cell32.get(), 0xFF, nc * 4);
      cd.narrow = 1;
      cd.cell32 = dn->cell32.get();
      cd.widen = &ctx_->derr->widen;
    } else if (trop) {
      dn->cell.reset(x.pool, nc);
      memset_async(x, dn->cell.get(), 0xFF, nc * 8);
    }
    if (trop) {
      dn->maxf.reset(x.pool, nwc);
      memset_async(x, dn->maxf.get(), 0, nwc * 4);
      if (has_inf_) {
        dn->infb.reset(x.pool, nwc);
        memset_async(x, dn->infb.get(), 0, nwc * 4);
      }
    } else {
      dn->present.reset(x.pool, nwc);
      memset_async(x, dn->present.get(), 0, nwc * 4);
    }
    dn->touched.reset(x.pool, nwc);
    memset_async(x, dn->touched.get(), 0, nwc * 4);
    dn->bad.reset(x.pool, 1);
    memset_async(x, dn->bad.get(), 0, 4);
    cd.cell = dn->cell.get();
    cd.present = dn->present.get();
    cd.maxf = dn->maxf.get();
    cd.infb = dn->infb.get();
    cd.touched = dn->touched.get();
    cd.bad = dn->bad.get();
    cd.on = 1;
    if (R.full.n) {
      with_key(R.wide, [&](auto* tag) {
        using K = std::remove_pointer_t<decltype(tag)>;
        cells_load<K>(x, cd, (const K*)R.full.k.ptr(), trop ? R.full.w.get() : nullptr,
                      (trop && has_inf_) ? R.full.inf.get() : nullptr, R.full.n, R.arity, R.bits);
      });
    }
    // Rules a round may evaluate from the destination side: two positive
    // literals, this relation's delta and a fixed binary relation E of the
    // same semiring, in the shape H(r, x) :- H(j, x), E(j, r); the head's row
    // column must be the cells' major column with an identity domain.
    if (R.arity == 2 && !nparts_ && !(trop && has_inf_)) {
      for (size_t ri = 0; ri < P_.rules.size(); ri++) {
        const Rule& ru = P_.rules[ri];
        if (ru.stratum != k || ru.agg || ru.head.rel != (int)i || ru.body.size() != 2) continue;
        for (int dpos = 0; dpos < 2; dpos++) {
          const Atom& A = ru.body[dpos];
          const Atom& B = ru.body[1 - dpos];
          if (A.neg || B.neg || A.rel != (int)i || B.rel < 0 || B.rel == (int)i || in_idb_[B.rel]) continue;
          if (R_[B.rel].arity != 2 || R_[B.rel].sem != R.sem || A.args.size() != 2 || B.args.size() != 2) continue;
          bool vars = ru.head.args.size() == 2;
          for (auto* at : {&ru.head, &A, &B})
            for (auto& t : at->args) vars = vars && t.kind == SK_T_VAR;
          if (!vars) continue;
          auto hv = [&](int c) { return ru.head.args[c].a; };
          auto av = [&](int c) { return A.args[c].a; };
          auto bv = [&](int c) { return B.args[c].a; };
          for (int Lc = 0; Lc < 2; Lc++) {
            const int Rc = 1 - Lc;
            if (hv(Lc) != av(Lc) || hv(Rc) == av(Rc)) continue;
            const u32 J = av(Rc);
            if (J == hv(Lc) || hv(Rc) == hv(Lc)) continue;
            for (int eJ = 0; eJ < 2; eJ++) {
              const int eR = 1 - eJ;
              if (bv(eJ) != J || bv(eR) != hv(Rc)) continue;
              if (cd.dom_bits[Rc] || cd.shift[Lc] != 0 || cd.shift[Rc] != cd.ebits[Lc]) continue;
              if (cd.ebits[Lc] > 6) continue;  // rows of at most one wave of lanes
              if (!trop && cd.ebits[Lc] >= 5) continue;  // whole-word Bool rows: bit pull
              PullRule pr;
              pr.rule = (int)ri;
              pr.dpos = dpos;
              pr.E = B.rel;
              pr.eR = eR;
              pr.eJ = eJ;
              pr.R = Rc;
              pr.L = Lc;
              pr.delta_first = dpos == 0;
              dn->pulls.push_back(std::move(pr));
              dn->lbits = cd.ebits[Lc];
            }
          }
        }
      }
      if (!dn->pulls.empty()) {
        dn->nrows = dn->ncells >> dn->lbits;
        dn->rowrec.reset(x.pool, 2 * (dn->nrows ? dn->nrows : 1));
        dn->dlane.reset(x.pool, std::max<u64>(dn->ncells, 1));
        cd.rowrec = dn->rowrec.get();
        cd.dlane = dn->dlane.get();
        cd.rnrows = dn->nrows;
        cd.rlbits = dn->lbits;
        // narrow cells whose weights stay below 2^26 also get the packed
        // delta (one word per entry) and 8-byte row records; a derivation at
        // or above 2^26 widens the cells as any narrow overflow does
        if (trop && cd.narrow && pack_ok_ && pack_enabled_) {
          dn->dpack.reset(x.pool, std::max<u64>(dn->ncells, 1));
          dn->rowrec8.reset(x.pool, dn->nrows ? dn->nrows : 1);
          cd.dpack = dn->dpack.get();
          cd.rowrec8 = dn->rowrec8.get();
          cd.narrow_lim = (u32)kPackLim;
        }
      }
    }
    if (!trop && R.arity >= 2) plan_bitpulls(k, (int)i, *dn);
    dense_[i] = std::move(dn);
  }
}

// Bool rules of relation i that a round may evaluate word-parallel (see
// BitPullRule): H(h..) :- H(d..), E(e..) in either body order, all terms
// variables, no variable repeated within an atom, the lane variable in the
// cells' least significant column of both H atoms and absent from E, every
// other H column's variable found in E, and every E variable one of those.
void Engine::plan_bitpulls(u32 k, int i, DenseRel& dn) {
  DRel& R = R_[i];
  CellDesc& cd = dn.cd;
  const int Lc = dn.order.back();
  const int lb = cd.ebits[Lc];
  if (lb < 5 || cd.shift[Lc] != 0 || (dn.ncells >> lb) >= (1ull << 32)) return;
  for (size_t ri = 0; ri < P_.rules.size(); ri++) {
    const Rule& ru = P_.rules[ri];
    if (ru.stratum != k || ru.agg || ru.head.rel != i || ru.body.size() != 2) continue;
    for (int dpos = 0; dpos < 2; dpos++) {
      const Atom& A = ru.body[dpos];
      const Atom& B = ru.body[1 - dpos];
      if (A.neg || B.neg || A.rel != i || B.rel < 0 || B.rel == i || in_idb_[B.rel]) continue;
      if (R_[B.rel].sem != SK_BOOL || (int)A.args.size() != R.arity || (int)ru.head.args.size() != R.arity ||
          (int)B.args.size() != R_[B.rel].arity)
        continue;
      bool ok = true;
      auto distinct_vars = [&](const Atom& at) {
        for (size_t a = 0; a < at.args.size(); a++) {
          if (at.args[a].kind != SK_T_VAR) return false;
          for (size_t b = 0; b < a; b++)
            if (at.args[b].a == at.args[a].a) return false;
        }
        return true;
      };
      ok = distinct_vars(ru.head) && distinct_vars(A) && distinct_vars(B);
      if (!ok) continue;
      const u32 xv = ru.head.args[Lc].a;
      if (A.args[Lc].a != xv) continue;
      auto in_e = [&](u32 v) {
        for (size_t a = 0; a < B.args.size(); a++)
          if (B.args[a].a == v) return (int)a;
        return -1;
      };
      if (in_e(xv) >= 0) continue;
      BitPullRule bp;
      bp.rule = (int)ri;
      bp.dpos = dpos;
      bp.E = B.rel;
      std::vector<bool> covered(B.args.size(), false);
      for (int c = 0; c < R.arity && ok; c++) {
        if (c == Lc) {
          bp.eh[c] = bp.ed[c] = -1;
          continue;
        }
        bp.eh[c] = in_e(ru.head.args[c].a);
        bp.ed[c] = in_e(A.args[c].a);
        ok = bp.eh[c] >= 0 && bp.ed[c] >= 0;
        if (ok) covered[bp.eh[c]] = covered[bp.ed[c]] = true;
      }
      for (bool cv : covered) ok = ok && cv;
      if (!ok) continue;
      dn.bpulls.push_back(std::move(bp));
    }
  }
  if (dn.bpulls.empty()) return;
  Exec x = ctx_->exec();
  dn.blbits = lb;
  dn.brows = dn.ncells >> lb;
  const u64 nwc = std::max<u64>(dn.ncells, 32) / 32;
  dn.dbits.reset(x.pool, nwc);
  dn.rstamp.reset(x.pool, dn.brows);
  memset_async(x, dn.rstamp.get(), 0, dn.brows * 4);
  dn.stamp = 0;
  cd.dbits = dn.dbits.get();
  cd.rstamp = dn.rstamp.get();
  cd.dwshift = lb - 5;
}

bool Engine::any_narrow() const {
  for (auto& d : dense_)
    if (d && d->cd.narrow) return true;
  return false;
}

void Engine::widen_dense() {
  Exec x = ctx_->exec();
  for (auto& d : dense_) {
    if (!d || !d->cd.narrow) continue;
    const u64 nc = std::max<u64>(d->ncells, 32);
    d->cell.reset(x.pool, nc);
    cells_widen(x, d->cell32.get(), d->cell.get(), nc);
    d->cell32.release();
    d->cd.narrow = 0;
    d->cd.cell32 = nullptr;
    d->cd.cell = d->cell.get();
    // wide cells pull from the unpacked delta (weights and lanes), which
    // every compaction writes as well
    d->cd.dpack = nullptr;
    d->cd.rowrec8 = nullptr;
    d->cd.narrow_lim = kCellTop32;
    d->dpack.release();
    d->rowrec8.release();
  }
  memset_async(x, &ctx_->derr->widen, 0, 4);
}

// Materialise every dense relation of the stratum as its sorted full state
// and drop the cells.
void Engine::end_dense() {
  Exec x = ctx_->exec();
  for (size_t i = 0; i < dense_.size(); i++) {
    DenseRel* dn = dense_[i].get();
    if (!dn) continue;
    DRel& R = R_[i];
    u32 bad = 0;
    SK_HIP(hipMemcpyAsync(&bad, dn->bad.get(), 4, hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    if (bad) throw mkerr(SK_ERR_DEVICE, P_.preds[i].name, "a derived value lies outside its dense column domain");
    if (dn->keep_full) continue;
    const bool trop = R.sem == SK_TROP;
    const bool hf = trop && has_inf_;
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      DBuf<K> fk;
      DBuf<i64> fw;
      DBuf<u8> ff;
      u64 n = cells_compact<K>(x, dn->cd, dn->ncells, false, R.bits, &fk, &fw, &ff, hf);
      if (!dn->key_order && n) {
        // cell order differs from key order: a stable sort of the compacted
        // state by its leading key columns 0..k-1, where k is the least count
        // such that cell order already ranks the columns k.. in key order
        const int a = R.arity;
        int k = a;
        for (int kk = 0; kk < a && k == a; kk++) {
          std::vector<int> rest;
          for (int c : dn->order)
            if (c >= kk) rest.push_back(c);
          bool inc = true;
          for (size_t q = 1; q < rest.size(); q++)
            if (rest[q] < rest[q - 1]) inc = false;
          if (inc) k = kk;
        }
        DBuf<K> ak(x.pool, n);
        DBuf<i64> aw;
        DBuf<u8> af;
        if (trop) aw.reset(x.pool, n);
        if (hf) af.reset(x.pool, n);
        K* sk;
        i64* sw;
        u8* sf;
        radix_sort<K>(x, fk.get(), ak.get(), trop ? fw.get() : nullptr, trop ? aw.get() : nullptr,
                      hf ? ff.get() : nullptr, hf ? af.get() : nullptr, n, R.bits * k, &sk, &sw, &sf,
                      R.bits * (a - k));
        if (sk != fk.get()) {
          std::swap(fk, ak);
          if (trop) std::swap(fw, aw);
          if (hf) std::swap(ff, af);
        }
      }
      RelData nf;
      nf.n = n;
      nf.k.wide = sizeof(K) == 16;
      if constexpr (sizeof(K) == 16) nf.k.b = std::move(fk);
      else nf.k.a = std::move(fk);
      nf.w = std::move(fw);
      nf.inf = std::move(ff);
      R.full = std::move(nf);
    });
    R.version++;
  }
  dense_.clear();
}

void Engine::run_stratum(u32 k) {
  // aggregate rules first, once (seminaive.rs:79-84)
  in_idb_.assign(R_.size(), false);
  for (size_t i = 0; i < P_.rules.size(); i++)
    if (P_.rules[i].stratum == k && P_.rules[i].agg) eval_aggregate(P_.rules[i]);
  for (auto& r : P_.rules)
    if (r.stratum == k && !r.agg && r.head.rel >= 0) in_idb_[r.head.rel] = true;
  bool is_trop = false;
  for (auto& r : P_.rules)
    if (r.stratum == k && r.head.rel >= 0 && P_.preds[r.head.rel].sem == SK_TROP) is_trop = true;
  for (size_t i = 0; i < R_.size(); i++)
    if (in_idb_[i]) {
      R_[i].delta.release();
      R_[i].delta_order.clear();
      R_[i].dversion++;
    }
  narrow_off_ = false;
  plan_dense(k);
  const u64 cap = universe_ + 2;
  u64 iters = 0;
  std::vector<u8> active_prev;  // batched mode: program had a non-empty delta after the last round
  std::map<std::pair<int, int>, RulePlan> plans;
  // The round's device status (overflow record, narrow widen flag) is read
  // with the first compaction's count when every relation the stratum
  // derives is dense with capacity compaction and none merges its deltas
  // into its full state; a widen seen that way restarts the stratum with
  // wide cells (the compaction has consumed the round's touched set).
  auto fold_ok = [&]() {
    if (nparts_) return false;
    for (size_t i = 0; i < R_.size(); i++) {
      if (!in_idb_[i]) continue;
      DenseRel* d = dense_of((int)i);
      if (!d || !d->into || d->keep_full) return false;
    }
    return true;
  };
  bool fold = fold_ok();
  auto plan_of = [&](int ri, int dpos) -> const RulePlan& {
    auto key = std::make_pair(ri, dpos);
    auto it = plans.find(key);
    if (it == plans.end()) it = plans.emplace(key, plan_rule(ri, dpos)).first;
    return it->second;
  };
  bool bootstrap = true;
  bool joined = false;  // the round's joins already ran on the device (device-driven rounds left mid-round)
  for (;;) {
    std::map<int, std::vector<Chunk>> chunks;
    round_variants_ = round_bitpulls_ = round_pulls_ = 0;
    auto run_rules = [&]() {
      for (size_t ri = 0; ri < P_.rules.size(); ri++) {
        const Rule& r = P_.rules[ri];
        if (r.stratum != k || r.agg) continue;
        if (r.head.rel < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, r.head.pred, "unknown predicate");
        if (bootstrap) {
          exec_plan(plan_of((int)ri, -1), chunks);
          continue;
        }
        for (size_t j = 0; j < r.body.size(); j++) {
          const Atom& a = r.body[j];
          if (a.neg || a.rel < 0 || !in_idb_[a.rel]) continue;
          if (R_[a.rel].delta.n == 0) continue;
          round_variants_++;
          exec_plan(plan_of((int)ri, (int)j), chunks);
        }
      }
    };
    if (!joined) run_rules();
    joined = false;
    const bool check_err = is_trop && !nparts_;
    const bool need_status = check_err || any_narrow();
    if (need_status && !fold) {
      DevErr st = read_device_status();
      if (st.widen) {
        // a derivation did not fit a narrow cell: widen every narrow relation
        // and apply the whole round again (re-applying is a no-op for ⊕)
        widen_dense();
        chunks.clear();
        run_rules();
        if (check_err) st = read_device_status();
      }
      if (check_err) raise_device_error(st);
    }
    bool any = false;
    fold_req_ = need_status && fold;
    fold_done_ = false;
    for (size_t i = 0; i < R_.size(); i++) {
      if (!in_idb_[i]) continue;
      auto it = chunks.find((int)i);
      if (it == chunks.end() && !dense_of((int)i)) {
        R_[i].delta.release();
        R_[i].delta_order.clear();
        R_[i].dversion++;
      } else if (it == chunks.end()) {
        std::vector<Chunk> none;
        if (finalize((int)i, none) > 0) any = true;
      } else if (finalize((int)i, it->second) > 0) {
        any = true;
      }
    }
    fold_req_ = false;
    // restart the stratum from its initial state with wide cells: the
    // relations it derives are dense and never merged into their full
    // state, so that state is still the stratum's input
    auto restart_wide = [&]() {
      dense_.clear();
      for (size_t i = 0; i < R_.size(); i++)
        if (in_idb_[i]) {
          R_[i].delta.release();
          R_[i].delta_order.clear();
          R_[i].dversion++;
        }
      memset_async(ctx_->exec(), &ctx_->derr->widen, 0, 4);
      narrow_off_ = true;
      plan_dense(k);
      fold = fold_ok();
      plans.clear();
      iters = 0;
      bootstrap = true;
    };
    if (need_status && fold) {
      if (!fold_done_) folded_ = read_device_status();
      if (folded_.widen) {
        restart_wide();
        continue;
      }
      if (check_err) raise_device_error(folded_);
    }
    if (nparts_) {
      part_round_bookkeeping(k, is_trop, active_prev, !bootstrap);
      any = false;
      for (size_t i = 0; i < R_.size(); i++)
        if (in_idb_[i] && R_[i].delta.n) any = true;
      if (!bootstrap) iters++;
    } else if (!bootstrap) {
      iters++;
      if (is_trop && iters > cap)
        throw mkerr(SK_ERR_NEGATIVE_WEIGHT_CYCLE, "", "negative-weight cycle: tropical stratum diverges");
    }
    bootstrap = false;
    if (!any) break;
    // the remaining rounds device-driven, when every delta variant pulled
    const int grel = graph_loop_rel(k);
    if (grel >= 0) {
      run_graph_loop(k, grel, &iters, active_prev);
      break;
    }
    const int trel = is_trop ? trop_graph_rel(k) : -1;
    if (trel >= 0) {
      DevErr gst{};
      const PullRule& tpr = dense_of(trel)->pulls[0];
      const int rc = run_trop_graph_loop(k, trel, cap, &iters, &gst, &plan_of(tpr.rule, tpr.dpos));
      if (rc == 0) break;
      if (rc == 3) throw mkerr(SK_ERR_NEGATIVE_WEIGHT_CYCLE, "", "negative-weight cycle: tropical stratum diverges");
      if (rc == 2) {
        if (gst.widen) {
          restart_wide();
          continue;
        }
        raise_device_error(gst);
      }
      // rc == 1: the next round pushes, on the host
    }
    int pri = -1, pdp = -1;
    const int prel = is_trop && nparts_ ? trop_part_graph_rel(k, &pri, &pdp) : -1;
    if (prel >= 0) {
      DevErr gst{};
      const int rc = run_trop_part_graph_loop(k, prel, plan_of(pri, pdp), &iters, active_prev, &gst);
      if (rc == 0) break;
      if (rc == 2) {
        // the round's join ran; the host finishes it (widening first when
        // the status says so) with the previous delta in place
        joined = true;
        continue;
      }
      if (rc == 3) {
        // a program passed its cap (crates/strata-eval/src/seminaive.rs:98-104)
        bool newly_dead = false;
        for (u32 p = 0; p < nparts_; p++)
          if (!part_dead_[p] && part_iters_[p][k] > part_universe_[p] + 2) {
            part_dead_[p] = 1;
            newly_dead = true;
            part_err_[p] = mkerr(SK_ERR_NEGATIVE_WEIGHT_CYCLE, "", "negative-weight cycle: tropical stratum diverges");
            active_prev[p] = 0;
          }
        if (newly_dead) drop_dead_from_deltas();
        bool more = false;
        for (size_t i = 0; i < R_.size(); i++)
          if (in_idb_[i] && R_[i].delta.n) more = true;
        if (!more) break;
      }
    }
  }
  iters_[k] = iters;
  end_dense();
  for (size_t i = 0; i < R_.size(); i++)
    if (in_idb_[i]) {
      R_[i].delta.release();
      R_[i].delta_order.clear();
      R_[i].dversion++;
    }
}

void Engine::execute() {
  if (!loaded) load();
  dense_.clear();
  iters_.assign(P_.num_strata, 0);
  auto t0 = std::chrono::steady_clock::now();
  for (u32 k = 0; k < P_.num_strata; k++) run_stratum(k);
  SK_HIP(hipStreamSynchronize(ctx_->stream));
  exec_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ========================================================= batched mode

void Engine::set_partitions(u32 np) {
  nparts_ = np;
  part_universe_.assign(np, 0);
  part_err_.assign(np, EvalErr());
  part_iters_.assign(np, std::vector<u64>(P_.num_strata, 0));
  part_dead_.assign(np, 0);
  Exec x = ctx_->exec();
  part_derr_.reset(x.pool, np ? np : 1);
  part_dead_dev_.reset(x.pool, np ? np : 1);
  memset_async(x, part_derr_.get(), 0, sizeof(DevErr) * (np ? np : 1));
  memset_async(x, part_dead_dev_.get(), 0, np ? np : 1);
}

// Rows per program id of a relation state (column 0 is the program id).
std::vector<u64> Engine::part_counts(int rel, bool delta) {
  DRel& R = R_[rel];
  const RelData& d = delta ? R.delta : R.full;
  // the full state's counts are read once per partition when results are
  // taken: keep them while the relation is unchanged
  auto cached = part_full_counts_.find(rel);
  if (!delta && cached != part_full_counts_.end() && cached->second.first == R.version) return cached->second.second;
  std::vector<u64> cnt(nparts_, 0);
  if (d.n == 0) return cnt;
  Exec x = ctx_->exec();
  DBuf<u64> csr(x.pool, nparts_ + 1);
  with_key(R.wide, [&](auto* tag) {
    using K = std::remove_pointer_t<decltype(tag)>;
    build_csr<K>(x, (const K*)d.k.ptr(), d.n, R.bits * (R.arity - 1), nparts_, csr.get());
  });
  std::vector<u64> h(nparts_ + 1);
  SK_HIP(hipMemcpyAsync(h.data(), csr.get(), (nparts_ + 1) * 8, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  for (u32 p = 0; p < nparts_; p++) cnt[p] = h[p + 1] - h[p];
  if (!delta) part_full_counts_[rel] = {R.version, cnt};
  return cnt;
}

void Engine::drop_dead_from_deltas() {
  Exec x = ctx_->exec();
  SK_HIP(hipMemcpyAsync(part_dead_dev_.get(), part_dead_.data(), nparts_, hipMemcpyHostToDevice, x.stream));
  for (size_t i = 0; i < R_.size(); i++) {
    if (!in_idb_[i] || R_[i].delta.n == 0) continue;
    DRel& R = R_[i];
    bool trop = R.sem == SK_TROP, hf = trop && has_inf_;
    RelData nd;
    nd.k.ensure(x.pool, R.delta.n, R.wide);
    if (trop) nd.w.reset(x.pool, R.delta.n);
    if (hf) nd.inf.reset(x.pool, R.delta.n);
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      nd.n = drop_partitions<K>(x, (const K*)R.delta.k.ptr(), trop ? R.delta.w.get() : nullptr,
                                hf ? R.delta.inf.get() : nullptr, R.delta.n, R.arity, R.bits, part_dead_dev_.get(),
                                (K*)nd.k.ptr(), trop ? nd.w.get() : nullptr, hf ? nd.inf.get() : nullptr);
    });
    R.delta = std::move(nd);
    R.dversion++;
  }
}

// After a round in batched mode: charge overflows to their programs, count
// each program's delta round and apply its own cap (crates/strata-eval/src/seminaive.rs:98-104).
void Engine::part_round_bookkeeping(u32 k, bool is_trop, std::vector<u8>& active_prev, bool count_round) {
  Exec x = ctx_->exec();
  bool newly_dead = false;
  if (is_trop) {
    std::vector<DevErr> h(nparts_);
    SK_HIP(hipMemcpyAsync(h.data(), part_derr_.get(), nparts_ * sizeof(DevErr), hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    bool any = false;
    for (u32 p = 0; p < nparts_; p++)
      if (h[p].code) {
        any = true;
        if (!part_dead_[p]) {
          part_dead_[p] = 1;
          newly_dead = true;
          EvalErr e = mkerr(SK_ERR_OVERFLOW, "", "tropical weight overflow");
          e.lhs = h[p].lhs;
          e.rhs = h[p].rhs;
          part_err_[p] = e;
        }
      }
    if (any) memset_async(x, part_derr_.get(), 0, sizeof(DevErr) * nparts_);
  }
  if (count_round) {
    for (u32 p = 0; p < nparts_; p++) {
      if (part_dead_[p] || !active_prev[p]) continue;
      part_iters_[p][k]++;
      if (is_trop && part_iters_[p][k] > part_universe_[p] + 2) {
        part_dead_[p] = 1;
        newly_dead = true;
        part_err_[p] = mkerr(SK_ERR_NEGATIVE_WEIGHT_CYCLE, "", "negative-weight cycle: tropical stratum diverges");
      }
    }
  }
  if (newly_dead) drop_dead_from_deltas();
  active_prev.assign(nparts_, 0);
  for (size_t i = 0; i < R_.size(); i++) {
    if (!in_idb_[i] || R_[i].delta.n == 0) continue;
    std::vector<u64> c = part_counts((int)i, true);
    for (u32 p = 0; p < nparts_; p++)
      if (c[p]) active_prev[p] = 1;
  }
}

u64 Engine::partition_size(const std::string& pred, u32 p) {
  int r = P_.rel_of(pred);
  if (r < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
  return part_counts(r, false)[p];
}

void Engine::partition_read_sym(const std::string& pred, u32 p, u32* const* cols, i64* w, u8* inf) {
  int r = P_.rel_of(pred);
  if (r < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
  DRel& R = R_[r];
  std::vector<u64> c = part_counts(r, false);
  u64 lo = 0;
  for (u32 q = 0; q < p; q++) lo += c[q];
  u64 n = c[p];
  if (!n) return;
  Exec x = ctx_->exec();
  std::vector<DBuf<u32>> dc(R.arity);
  std::vector<u32*> cp(R.arity);
  for (int j = 0; j < R.arity; j++) {
    dc[j].reset(x.pool, n);
    cp[j] = dc[j].get();
  }
  with_key(R.wide, [&](auto* tag) {
    using K = std::remove_pointer_t<decltype(tag)>;
    unpack_cols<K>(x, (const K*)R.full.k.ptr() + lo, n, R.arity, R.bits, cp.data());
  });
  for (int j = 1; j < R.arity; j++)
    if (cols[j - 1]) ctx_->d2h(cols[j - 1], cp[j], n * 4);
  bool trop = R.sem == SK_TROP;
  if (w && trop) ctx_->d2h(w, R.full.w.get() + lo, n * 8);
  if (inf) {
    if (trop && has_inf_) ctx_->d2h(inf, R.full.inf.get() + lo, n);
    else std::memset(inf, 0, n);
  }
  SK_HIP(hipStreamSynchronize(x.stream));
}

void Engine::attach(const std::string& pred, u64 n, void* keys, bool wide, int bits, i64* w, u8* inf, u64* csr,
                    u64 max_sym, u64 universe_syms) {
  (void)universe_syms;
  int r = P_.rel_of(pred);
  if (r < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
  if (attached_.empty()) attached_.assign(R_.size(), false);
  attached_[r] = true;
  DRel& R = R_[r];
  R.bits = bits;
  R.wide = wide;
  R.full.n = n;
  R.full.k.wide = wide;
  // non-owning views: a DBuf without a pool never frees its pointer
  if (wide) {
    R.full.k.b.p = (u128*)keys;
    R.full.k.b.n = n;
  } else {
    R.full.k.a.p = (u64*)keys;
    R.full.k.a.n = n;
  }
  R.full.w.p = w;
  R.full.w.n = w ? n : 0;
  R.full.inf.p = inf;
  R.full.inf.n = inf ? n : 0;
  attached_max_sym_ = std::max(attached_max_sym_, max_sym);
  if (csr) {
    Index& ix = R.idx["f01"];
    ix.perm = {0, 1};
    ix.version = R.version;
    ix.csr.p = csr;
    ix.csr.n = max_sym + 2;
    ix.csr_version = R.version;
  }
}

// ================================================================ results

int64_t Engine::size(const std::string& pred) {
  int r = P_.rel_of(pred);
  return r < 0 ? -1 : (int64_t)R_[r].full.n;
}

void Engine::read(const std::string& pred, sk_val* rows, i64* w, u8* inf, u8* is_w) {
  int r = P_.rel_of(pred);
  if (r < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
  DRel& R = R_[r];
  Exec x = ctx_->exec();
  u64 n = R.full.n;
  if (n == 0) return;
  const int bits = R.bits;
  if (rows && R.arity) {
    std::vector<DBuf<u32>> dc(R.arity);
    std::vector<u32*> cp(R.arity);
    for (int j = 0; j < R.arity; j++) {
      dc[j].reset(x.pool, n);
      cp[j] = dc[j].get();
    }
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      unpack_cols<K>(x, (const K*)R.full.k.ptr(), n, R.arity, bits, cp.data());
    });
    std::vector<u32> hcol(n);
    for (int j = 0; j < R.arity; j++) {
      SK_HIP(hipMemcpyAsync(hcol.data(), cp[j], n * 4, hipMemcpyDeviceToHost, x.stream));
      SK_HIP(hipStreamSynchronize(x.stream));
      for (u64 i = 0; i < n; i++) {
        u32 kind;
        i64 v;
        D_.decode(hcol[i], &kind, &v);
        rows[i * R.arity + j].kind = kind;
        rows[i * R.arity + j].pad = 0;
        rows[i * R.arity + j].v = v;
      }
    }
  }
  bool trop = R.sem == SK_TROP;
  if (w) {
    if (trop) SK_HIP(hipMemcpyAsync(w, R.full.w.get(), n * 8, hipMemcpyDeviceToHost, x.stream));
    else std::memset(w, 0, n * 8);
  }
  if (inf) {
    if (trop && has_inf_) SK_HIP(hipMemcpyAsync(inf, R.full.inf.get(), n, hipMemcpyDeviceToHost, x.stream));
    else std::memset(inf, 0, n);
  }
  if (is_w) std::memset(is_w, trop ? 1 : 0, n);
  SK_HIP(hipStreamSynchronize(x.stream));
}

void Engine::read_sym(const std::string& pred, u32* const* cols, i64* w, u8* inf) {
  int r = P_.rel_of(pred);
  if (r < 0) throw mkerr(SK_ERR_UNKNOWN_PRED, pred, "unknown predicate");
  DRel& R = R_[r];
  Exec x = ctx_->exec();
  u64 n = R.full.n;
  if (n == 0) return;
  const int bits = R.bits;
  if (cols && R.arity) {
    std::vector<DBuf<u32>> dc(R.arity);
    std::vector<u32*> cp(R.arity);
    for (int j = 0; j < R.arity; j++) {
      dc[j].reset(x.pool, n);
      cp[j] = dc[j].get();
    }
    with_key(R.wide, [&](auto* tag) {
      using K = std::remove_pointer_t<decltype(tag)>;
      unpack_cols<K>(x, (const K*)R.full.k.ptr(), n, R.arity, bits, cp.data());
    });
    for (int j = 0; j < R.arity; j++)
      if (cols[j]) ctx_->d2h(cols[j], cp[j], n * 4);
    SK_HIP(hipStreamSynchronize(x.stream));
    if (!D_.ints.empty())
      for (int j = 0; j < R.arity; j++)
        if (cols[j])
          for (u64 i = 0; i < n; i++)
            if (cols[j][i] >= D_.sym_span) throw mkerr(SK_ERR_INVALID_ARGUMENT, pred, "relation holds Int values");
  }
  bool trop = R.sem == SK_TROP;
  if (w) {
    if (trop) ctx_->d2h(w, R.full.w.get(), n * 8);
    else std::memset(w, 0, n * 8);
  }
  if (inf) {
    if (trop && has_inf_) ctx_->d2h(inf, R.full.inf.get(), n);
    else std::memset(inf, 0, n);
  }
  SK_HIP(hipStreamSynchronize(x.stream));
}

}  // namespace sk
