// sk_server.cpp — resident-EDB query serving with continuous batching.
//
// The EDB edge(node, node) (Trop) stays on the device, sorted by (src, dst)
// with the minimum weight per pair and a CSR over src.  A query is the
// multi-source cheapest-cost program
//   cost(S,Y) :- src(S), edge(S,Y).        (stratum 0)
//   cost(S,Z) :- cost(S,Y), edge(Y,Z).     (stratum 0)
// optionally followed by a target filter
//   ans(S,Z) :- cost(S,Z), tgt(Z).         (stratum 1)
// Active queries occupy slots; their relations live in one fused relation
// cost(slot, S, Z).  Every sk_server_step is one semi-naive round for all
// active queries together: newly admitted queries contribute their bootstrap
// round (src ⋈ edge), running ones their delta round (Δcost ⋈ edge), and
// the union is sorted, min-collapsed and merged in one pass.  A query whose
// delta becomes empty is retired: its rows move to an answer buffer, its slot
// and memory are freed, and waiting queries are admitted while slots and the
// HBM budget allow.  Each query keeps its own round count and Bellman-Ford
// cap (crates/strata-eval/src/seminaive.rs:94-104), so an answer is the same
// alone and in any mix.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "sk_engine.h"
#include "strata_k.h"

using namespace sk;

namespace {

using Clock = std::chrono::steady_clock;

struct Query {
  u64 ticket = 0;
  std::vector<u32> sources, targets;
  Clock::time_point arrival;
  int state = 0;  // 0 waiting, 1 active, 2 done, 3 error
  int slot = -1;
  bool ran_bootstrap = false;
  bool had_delta = false;  // non-empty delta after the last round
  u64 iters = 0;
  u64 universe = 0;
  u64 rows = 0;  // current full rows (admission accounting)
  EvalErr err;
  double latency_ms = 0;
  // answer (device, sorted by (S, Z))
  u64 n = 0;
  DBuf<u128> keys;
  DBuf<i64> w;
};

}  // namespace

struct sk_server {
  sk_ctx* ctx = nullptr;
  u64 n_edges = 0;
  u64 n_nodes = 0;  // node ids are < n_nodes
  int bits = 1;
  DBuf<u64> ekeys;
  DBuf<i64> ew;
  DBuf<u64> ecsr;
  std::vector<u8> edge_node;  // host bitmap (byte per node) of ids used by the EDB
  u64 edge_universe = 0;
  u64 budget = 0;
  u32 max_batch = 64;
  int slot_bits = 6;
  std::vector<int> slot_q;  // slot -> query index, -1 free
  std::vector<std::unique_ptr<Query>> qs;
  std::map<u64, size_t> by_ticket;
  std::deque<size_t> waiting;
  u64 next_ticket = 1;
  // fused state: cost(slot, S, Z)
  RelData full, delta;
  DBuf<DevErr> perr;
  DBuf<u8> dead_dev;
  double avg_rows_per_source = 0;
  u64 finished_sources = 0, finished_rows = 0;
  EvalErr err;
};

namespace {

int bits_for(u64 nvals) {
  int b = 1;
  while (b < 63 && (1ull << b) < nvals) b++;
  return b;
}

void table_alloc_srv(void* ctx, Table* t, u64 n);

struct SrvTable {
  Pool* pool;
  Table t;
  std::vector<DBuf<u32>> cols;
  DBuf<i64> w;
};

void table_alloc_srv(void* ctx, Table* t, u64 n) {
  SrvTable* o = (SrvTable*)ctx;
  o->cols.resize(t->ncols);
  for (int c = 0; c < t->ncols; c++) {
    o->cols[c].reset(o->pool, n ? n : 1);
    t->col[c] = o->cols[c].get();
  }
  o->w.reset(o->pool, n ? n : 1);
  t->w = o->w.get();
  t->inf = nullptr;
  t->n = n;
}

struct SrvHead {
  Pool* pool;
  DBuf<u128> k;
  DBuf<i64> w;
  u64 n = 0;
};

void head_alloc_srv(void* ctx, OutDesc* o, u64 n) {
  SrvHead* h = (SrvHead*)ctx;
  h->k.ensure(h->pool, n ? n : 1);
  h->w.ensure(h->pool, n ? n : 1);
  o->keys = h->k.get();
  o->w = h->w.get();
  o->inf = nullptr;
  h->n = n;
}

// Rows per slot of a fused relation state (slot is column 0).
std::vector<u64> slot_counts(sk_server* s, const RelData& d) {
  std::vector<u64> cnt(s->max_batch, 0);
  if (!d.n) return cnt;
  Exec x = s->ctx->c.exec();
  DBuf<u64> csr(x.pool, s->max_batch + 1);
  build_csr<u128>(x, (const u128*)d.k.ptr(), d.n, s->bits * 2, s->max_batch, csr.get());
  std::vector<u64> h(s->max_batch + 1);
  SK_HIP(hipMemcpyAsync(h.data(), csr.get(), (s->max_batch + 1) * 8, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  for (u32 p = 0; p < s->max_batch; p++) cnt[p] = h[p + 1] - h[p];
  return cnt;
}

void drop_slots(sk_server* s, RelData& d, const std::vector<u8>& dead) {
  if (!d.n) return;
  Exec x = s->ctx->c.exec();
  SK_HIP(hipMemcpyAsync(s->dead_dev.get(), dead.data(), dead.size(), hipMemcpyHostToDevice, x.stream));
  RelData nd;
  nd.k.ensure(x.pool, d.n, true);
  nd.w.reset(x.pool, d.n);
  nd.n = drop_partitions<u128>(x, (const u128*)d.k.ptr(), d.w.get(), nullptr, d.n, 3, s->bits, s->dead_dev.get(),
                               (u128*)nd.k.ptr(), nd.w.get(), nullptr);
  d = std::move(nd);
}

// Membership filter of answer rows on Z ∈ targets (ans(S,Z) :- cost(S,Z), tgt(Z)).
void filter_targets(sk_server* s, Query& q, u128* keys, i64* w, u64 n) {
  Exec x = s->ctx->c.exec();
  // host-side filter of a device range: targets are few (tens)
  std::vector<u128> hk(n);
  std::vector<i64> hw(n);
  SK_HIP(hipMemcpyAsync(hk.data(), keys, n * 16, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipMemcpyAsync(hw.data(), w, n * 8, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  u128 mask = (((u128)1) << s->bits) - 1;
  std::vector<u128> ok;
  std::vector<i64> ow;
  for (u64 i = 0; i < n; i++) {
    u32 z = (u32)(hk[i] & mask);
    if (std::binary_search(q.targets.begin(), q.targets.end(), z)) {
      ok.push_back(hk[i]);
      ow.push_back(hw[i]);
    }
  }
  q.n = ok.size();
  q.keys.reset(x.pool, q.n ? q.n : 1);
  q.w.reset(x.pool, q.n ? q.n : 1);
  if (q.n) {
    SK_HIP(hipMemcpyAsync(q.keys.get(), ok.data(), q.n * 16, hipMemcpyHostToDevice, x.stream));
    SK_HIP(hipMemcpyAsync(q.w.get(), ow.data(), q.n * 8, hipMemcpyHostToDevice, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
  }
}

void retire(sk_server* s, Query& q, u64 lo, u64 n) {
  Exec x = s->ctx->c.exec();
  if (q.targets.empty()) {
    q.n = n;
    q.keys.reset(x.pool, n ? n : 1);
    q.w.reset(x.pool, n ? n : 1);
    if (n) {
      SK_HIP(hipMemcpyAsync(q.keys.get(), (const u128*)s->full.k.ptr() + lo, n * 16, hipMemcpyDeviceToDevice, x.stream));
      SK_HIP(hipMemcpyAsync(q.w.get(), s->full.w.get() + lo, n * 8, hipMemcpyDeviceToDevice, x.stream));
    }
  } else {
    filter_targets(s, q, (u128*)s->full.k.ptr() + lo, s->full.w.get() + lo, n);
  }
  SK_HIP(hipStreamSynchronize(x.stream));
  q.state = 2;
  q.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.arrival).count();
  s->finished_sources += q.sources.size();
  s->finished_rows += n;
}

u64 estimate_bytes(sk_server* s, const Query& q) {
  double per_src = s->finished_sources ? (double)s->finished_rows / s->finished_sources : (double)s->n_nodes / 4;
  // full + delta + sorted candidates, 24 bytes a row
  return (u64)(q.sources.size() * per_src * 24.0 * 3.0);
}

}  // namespace

extern "C" {

sk_status sk_server_create(sk_ctx* ctx, const uint32_t* src, const uint32_t* dst, const int64_t* w, uint64_t n_edges,
                           uint64_t query_budget_bytes, uint32_t max_batch, sk_server** out) {
  if (!ctx || !out || (n_edges && (!src || !dst || !w)) || max_batch == 0) return SK_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto s = std::make_unique<sk_server>();
  s->ctx = ctx;
  s->max_batch = max_batch;
  s->slot_q.assign(max_batch, -1);
  s->budget = query_budget_bytes;
  EvalErr e;
  try {
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    Exec x = c.exec();
    u32 maxv = 0;
    for (u64 i = 0; i < n_edges; i++) maxv = std::max(maxv, std::max(src[i], dst[i]));
    s->n_nodes = (u64)maxv + 1;
    s->bits = bits_for(std::max<u64>(s->n_nodes, max_batch));
    s->slot_bits = bits_for(max_batch);
    s->edge_node.assign(s->n_nodes, 0);
    for (u64 i = 0; i < n_edges; i++) s->edge_node[src[i]] = s->edge_node[dst[i]] = 1;
    for (u8 b : s->edge_node) s->edge_universe += b;
    // resident EDB: pack, sort by (src, dst), min per pair, CSR over src
    DBuf<u32> a(c.pool.get(), n_edges ? n_edges : 1), b(c.pool.get(), n_edges ? n_edges : 1);
    DBuf<u64> k1(c.pool.get(), n_edges ? n_edges : 1), k2(c.pool.get(), n_edges ? n_edges : 1);
    DBuf<i64> w1(c.pool.get(), n_edges ? n_edges : 1), w2(c.pool.get(), n_edges ? n_edges : 1);
    SK_HIP(hipMemcpyAsync(a.get(), src, n_edges * 4, hipMemcpyHostToDevice, c.stream));
    SK_HIP(hipMemcpyAsync(b.get(), dst, n_edges * 4, hipMemcpyHostToDevice, c.stream));
    SK_HIP(hipMemcpyAsync(w1.get(), w, n_edges * 8, hipMemcpyHostToDevice, c.stream));
    Cols cols{};
    cols.c[0] = a.get();
    cols.c[1] = b.get();
    pack_cols<u64>(x, cols, 2, s->bits, n_edges, k1.get());
    u64* sk;
    i64* sw;
    radix_sort<u64>(x, k1.get(), k2.get(), w1.get(), w2.get(), nullptr, nullptr, n_edges, 2 * s->bits, &sk, &sw,
                    nullptr);
    s->ekeys.reset(c.pool.get(), n_edges ? n_edges : 1);
    s->ew.reset(c.pool.get(), n_edges ? n_edges : 1);
    s->n_edges = unique_runs<u64>(x, sk, sw, nullptr, n_edges, s->ekeys.get(), s->ew.get(), nullptr);
    s->ecsr.reset(c.pool.get(), s->n_nodes + 1);
    build_csr<u64>(x, s->ekeys.get(), s->n_edges, s->bits, s->n_nodes, s->ecsr.get());
    s->perr.reset(c.pool.get(), max_batch);
    s->dead_dev.reset(c.pool.get(), max_batch);
    memset_async(x, s->perr.get(), 0, sizeof(DevErr) * max_batch);
    SK_HIP(hipStreamSynchronize(c.stream));
    if (s->bits * 3 > 128) throw EvalErr{SK_ERR_UNSUPPORTED_SHAPE};
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  } catch (const EvalErr& ee) {
    return (sk_status)ee.code;
  }
  if (s->budget == 0) s->budget = (u64)(ctx->c.pool->budget() * 0.6);
  *out = s.release();
  return SK_OK;
}

void sk_server_free(sk_server* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->ctx->c.stream);
  delete s;
}

sk_status sk_server_submit(sk_server* s, const uint32_t* sources, uint32_t n_sources, const uint32_t* targets,
                           uint32_t n_targets, uint64_t* ticket) {
  if (!s || !ticket || (n_sources && !sources) || (n_targets && !targets)) return SK_ERR_INVALID_ARGUMENT;
  auto q = std::make_unique<Query>();
  q->ticket = s->next_ticket++;
  q->sources.assign(sources, sources + n_sources);
  std::sort(q->sources.begin(), q->sources.end());
  q->sources.erase(std::unique(q->sources.begin(), q->sources.end()), q->sources.end());
  q->targets.assign(targets, targets + n_targets);
  std::sort(q->targets.begin(), q->targets.end());
  q->targets.erase(std::unique(q->targets.begin(), q->targets.end()), q->targets.end());
  for (u32 v : q->sources)
    if (v >= s->n_nodes) return SK_ERR_INVALID_ARGUMENT;
  for (u32 v : q->targets)
    if (v >= s->n_nodes) return SK_ERR_INVALID_ARGUMENT;
  // universe: distinct Sym values over edge ∪ src ∪ tgt (crates/strata-eval/src/naive.rs:143-155)
  u64 u = s->edge_universe;
  std::vector<u32> extra;
  for (u32 v : q->sources)
    if (!s->edge_node[v]) extra.push_back(v);
  for (u32 v : q->targets)
    if (!s->edge_node[v]) extra.push_back(v);
  std::sort(extra.begin(), extra.end());
  u += std::unique(extra.begin(), extra.end()) - extra.begin();
  q->universe = u;
  q->arrival = Clock::now();
  *ticket = q->ticket;
  s->by_ticket[q->ticket] = s->qs.size();
  s->waiting.push_back(s->qs.size());
  s->qs.push_back(std::move(q));
  return SK_OK;
}

sk_status sk_server_step(sk_server* s, uint32_t* n_active, uint32_t* n_waiting) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  Ctx& c = s->ctx->c;
  (void)hipSetDevice(c.device);
  try {
    Exec x = c.exec();
    // ---- admission: free slots, within the HBM budget
    u64 in_use = (s->full.n + s->delta.n) * 24;
    std::vector<size_t> admitted;
    while (!s->waiting.empty()) {
      int slot = -1;
      for (u32 i = 0; i < s->max_batch; i++)
        if (s->slot_q[i] < 0) {
          slot = (int)i;
          break;
        }
      if (slot < 0) break;
      Query& q = *s->qs[s->waiting.front()];
      u64 est = estimate_bytes(s, q);
      bool idle = s->full.n == 0 && s->delta.n == 0 && admitted.empty();
      if (!idle && in_use + est > s->budget) break;
      in_use += est;
      q.slot = slot;
      q.state = 1;
      s->slot_q[slot] = (int)s->waiting.front();
      admitted.push_back(s->waiting.front());
      s->waiting.pop_front();
    }
    // ---- candidates of this round
    std::vector<SrvHead> heads;
    if (!admitted.empty()) {
      // bootstrap of the new queries: cost(S,Y) :- src(S), edge(S,Y)
      std::vector<u32> qc, sc;
      for (size_t qi : admitted) {
        Query& q = *s->qs[qi];
        q.ran_bootstrap = true;
        for (u32 v : q.sources) {
          qc.push_back((u32)q.slot);
          sc.push_back(v);
        }
      }
      SrvTable t;
      t.pool = x.pool;
      t.t.ncols = 2;
      table_alloc_srv(&t, &t.t, qc.size());
      SK_HIP(hipMemcpyAsync(t.t.col[0], qc.data(), qc.size() * 4, hipMemcpyHostToDevice, x.stream));
      SK_HIP(hipMemcpyAsync(t.t.col[1], sc.data(), sc.size() * 4, hipMemcpyHostToDevice, x.stream));
      memset_async(x, t.t.w, 0, qc.size() * 8);  // src facts carry weight 0 (otimes_id)
      t.t.part_col = 0;
      t.t.perr = s->perr.get();
      ProbeDesc pd;
      pd.arity = 2;
      pd.bits = s->bits;
      pd.sn = s->n_edges;
      pd.nprobe = 1;
      pd.psrc[0] = 1;
      OutDesc o;
      o.ncols = 3;
      o.bits = s->bits;
      o.kind[0] = 0; o.idx[0] = 0;
      o.kind[1] = 0; o.idx[1] = 1;
      o.kind[2] = 1; o.idx[2] = 1;
      heads.emplace_back();
      heads.back().pool = x.pool;
      join<u64>(x, t.t, pd, s->ekeys.get(), s->ew.get(), nullptr, s->ecsr.get(), &o, true, true, true, nullptr,
                nullptr, head_alloc_srv, &heads.back());
    }
    if (s->delta.n) {
      // delta round: cost(S,Z) :- cost(S,Y), edge(Y,Z) with cost from the delta
      SrvTable t;
      t.pool = x.pool;
      t.t.ncols = 3;
      table_alloc_srv(&t, &t.t, s->delta.n);
      u32* cp[3] = {t.t.col[0], t.t.col[1], t.t.col[2]};
      unpack_cols<u128>(x, (const u128*)s->delta.k.ptr(), s->delta.n, 3, s->bits, cp);
      SK_HIP(hipMemcpyAsync(t.t.w, s->delta.w.get(), s->delta.n * 8, hipMemcpyDeviceToDevice, x.stream));
      t.t.part_col = 0;
      t.t.perr = s->perr.get();
      ProbeDesc pd;
      pd.arity = 2;
      pd.bits = s->bits;
      pd.sn = s->n_edges;
      pd.nprobe = 1;
      pd.psrc[0] = 2;
      OutDesc o;
      o.ncols = 3;
      o.bits = s->bits;
      o.kind[0] = 0; o.idx[0] = 0;
      o.kind[1] = 0; o.idx[1] = 1;
      o.kind[2] = 1; o.idx[2] = 1;
      heads.emplace_back();
      heads.back().pool = x.pool;
      join<u64>(x, t.t, pd, s->ekeys.get(), s->ew.get(), nullptr, s->ecsr.get(), &o, true, true, true, nullptr,
                nullptr, head_alloc_srv, &heads.back());
    }
    u64 total = 0;
    for (auto& h : heads) total += h.n;
    // ---- ⊕-combine (sort, min per key, merge with full)
    RelData fresh;
    fresh.n = 0;
    if (total) {
      DBuf<u128> ck(x.pool, total), ck2(x.pool, total);
      DBuf<i64> cw(x.pool, total), cw2(x.pool, total);
      u64 off = 0;
      for (auto& h : heads) {
        if (!h.n) continue;
        SK_HIP(hipMemcpyAsync(ck.get() + off, h.k.get(), h.n * 16, hipMemcpyDeviceToDevice, x.stream));
        SK_HIP(hipMemcpyAsync(cw.get() + off, h.w.get(), h.n * 8, hipMemcpyDeviceToDevice, x.stream));
        off += h.n;
      }
      heads.clear();
      u128* sk;
      i64* sw;
      int kb = 2 * s->bits + s->slot_bits;
      radix_sort<u128>(x, ck.get(), ck2.get(), cw.get(), cw2.get(), nullptr, nullptr, total, kb, &sk, &sw, nullptr);
      DBuf<u128> uk(x.pool, total);
      DBuf<i64> uw(x.pool, total);
      u64 nu = unique_runs<u128>(x, sk, sw, nullptr, total, uk.get(), uw.get(), nullptr);
      if (s->full.n == 0) {
        s->full.n = nu;
        s->full.k.ensure(x.pool, nu, true);
        s->full.w.reset(x.pool, nu);
        SK_HIP(hipMemcpyAsync(s->full.k.ptr(), uk.get(), nu * 16, hipMemcpyDeviceToDevice, x.stream));
        SK_HIP(hipMemcpyAsync(s->full.w.get(), uw.get(), nu * 8, hipMemcpyDeviceToDevice, x.stream));
        fresh.n = nu;
        fresh.k.wide = true;
        fresh.k.b = std::move(uk);
        fresh.w = std::move(uw);
      } else {
        DBuf<u128> u_k, f_k;
        DBuf<i64> u_w, f_w;
        DBuf<u8> u_f, f_f;
        u64 n_union = 0, n_fresh = 0;
        merge_diff<u128>(x, uk.get(), uw.get(), nullptr, nu, (const u128*)s->full.k.ptr(), s->full.w.get(), nullptr,
                         s->full.n, &u_k, &u_w, &u_f, &n_union, &f_k, &f_w, &f_f, &n_fresh, true, false);
        if (n_fresh) {
          RelData nf;
          nf.n = n_union;
          nf.k.wide = true;
          nf.k.b = std::move(u_k);
          nf.w = std::move(u_w);
          s->full = std::move(nf);
          fresh.n = n_fresh;
          fresh.k.wide = true;
          fresh.k.b = std::move(f_k);
          fresh.w = std::move(f_w);
        }
      }
    }
    s->delta = std::move(fresh);
    // ---- per-query bookkeeping: overflow, round count, cap, retirement
    std::vector<DevErr> herr(s->max_batch);
    SK_HIP(hipMemcpyAsync(herr.data(), s->perr.get(), s->max_batch * sizeof(DevErr), hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    bool clear_err = false;
    std::vector<u64> dcnt = slot_counts(s, s->delta);
    std::vector<u64> fcnt = slot_counts(s, s->full);
    std::vector<u8> dead(s->max_batch, 0);
    bool any_dead = false;
    u64 lo = 0;
    std::vector<u64> start(s->max_batch, 0);
    for (u32 p = 0; p < s->max_batch; p++) {
      start[p] = lo;
      lo += fcnt[p];
    }
    for (u32 p = 0; p < s->max_batch; p++) {
      int qi = s->slot_q[p];
      if (qi < 0) continue;
      Query& q = *s->qs[qi];
      bool newly = std::find(admitted.begin(), admitted.end(), (size_t)qi) != admitted.end();
      if (herr[p].code) {
        clear_err = true;
        q.err.code = SK_ERR_OVERFLOW;
        q.err.lhs = herr[p].lhs;
        q.err.rhs = herr[p].rhs;
        q.state = 3;
      }
      if (q.state == 1 && !newly && q.had_delta) {
        q.iters++;
        if (q.iters > q.universe + 2) {
          q.err.code = SK_ERR_NEGATIVE_WEIGHT_CYCLE;
          q.state = 3;
        }
      }
      q.had_delta = dcnt[p] > 0;
      q.rows = fcnt[p];
      if (q.state == 1 && !q.had_delta) retire(s, q, start[p], fcnt[p]);
      if (q.state != 1) {
        if (q.state == 3) q.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.arrival).count();
        dead[p] = 1;
        any_dead = true;
        s->slot_q[p] = -1;
      }
    }
    if (clear_err) memset_async(x, s->perr.get(), 0, sizeof(DevErr) * s->max_batch);
    if (any_dead) {
      drop_slots(s, s->full, dead);
      drop_slots(s, s->delta, dead);
    }
    SK_HIP(hipStreamSynchronize(x.stream));
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  }
  u32 na = 0;
  for (int q : s->slot_q) na += q >= 0;
  if (n_active) *n_active = na;
  if (n_waiting) *n_waiting = (uint32_t)s->waiting.size();
  return SK_OK;
}

sk_status sk_server_poll(sk_server* s, uint64_t ticket, uint64_t* n_rows, uint64_t* iterations, double* latency_ms) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  if (q.state == 3) return (sk_status)q.err.code;
  if (q.state != 2) return SK_ERR_NOT_READY;
  if (n_rows) *n_rows = q.n;
  if (iterations) *iterations = q.iters;
  if (latency_ms) *latency_ms = q.latency_ms;
  return SK_OK;
}

sk_status sk_server_take(sk_server* s, uint64_t ticket, uint32_t* src, uint32_t* dst, int64_t* w) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  if (q.state != 2) return q.state == 3 ? (sk_status)q.err.code : SK_ERR_NOT_READY;
  try {
    Exec x = s->ctx->c.exec();
    if (q.n) {
      DBuf<u32> c0(x.pool, q.n), c1(x.pool, q.n), c2(x.pool, q.n);
      u32* cp[3] = {c0.get(), c1.get(), c2.get()};
      unpack_cols<u128>(x, q.keys.get(), q.n, 3, s->bits, cp);
      if (src) SK_HIP(hipMemcpyAsync(src, c1.get(), q.n * 4, hipMemcpyDeviceToHost, x.stream));
      if (dst) SK_HIP(hipMemcpyAsync(dst, c2.get(), q.n * 4, hipMemcpyDeviceToHost, x.stream));
      if (w) SK_HIP(hipMemcpyAsync(w, q.w.get(), q.n * 8, hipMemcpyDeviceToHost, x.stream));
      SK_HIP(hipStreamSynchronize(x.stream));
    }
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  }
  q.keys.release();
  q.w.release();
  return SK_OK;
}

sk_status sk_server_release(sk_server* s, uint64_t ticket) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  q.keys.release();
  q.w.release();
  return SK_OK;
}

}  // extern "C"
