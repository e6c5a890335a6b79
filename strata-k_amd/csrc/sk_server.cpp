// sk_server.cpp — resident-EDB query serving with continuous batching.
//
// The EDB edge(node, node) (Trop) stays on the device, sorted by (src, dst)
// with the minimum weight per pair and a CSR over src.  A query is the
// multi-source cheapest-cost program
//   cost(S,Y) :- src(S), edge(S,Y).        (stratum 0)
//   cost(S,Z) :- cost(S,Y), edge(Y,Z).     (stratum 0)
// optionally followed by a target filter
//   ans(S,Z) :- cost(S,Z), tgt(Z).         (stratum 1)
//
// State.  Every source of an admitted query gets an *instance*: one row of a
// resident cell array, cost(instance, Z) for every node Z, holding the
// order-preserving code of the least weight found so far (sk_cells.h).  The
// number of instance rows is fixed at creation from the byte budget, and
// queries are admitted while slots and free rows allow; a row is reset when
// it is handed out, so nothing is allocated per query on the device except
// its answer.
//
// Rounds.  Every sk_server_step is one semi-naive round for all active
// queries together: newly admitted queries contribute their bootstrap
// (src ⋈ edge, weight 0 ⊗ w), running ones their delta round (Δcost ⋈ edge).
// Both are joins whose emitting kernel ⊕-applies each derivation to its
// instance's cell and marks strict improvements; one compaction of the marks
// gives every query's next delta.  Rounds are Jacobi: the joins read the
// previous delta list, never the cells.  Each query keeps its own round
// count, Bellman-Ford cap (crates/strata-eval/src/seminaive.rs:94-104) and
// overflow record (charged per instance), so an answer is the same alone and
// in any mix.  A query whose delta is empty retires: its rows are compacted
// (or, with a target filter, only the target cells are read) into its answer,
// and its rows and slot are freed.
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
  std::vector<u32> inst;  // instance row per source (sources sorted)
  bool had_delta = false;  // non-empty delta after the last round
  u64 iters = 0;
  u64 universe = 0;
  EvalErr err;
  double latency_ms = 0;
  // answer: rows sorted by (S, Z); S is expanded from per-source counts
  u64 n = 0;
  std::vector<u64> per_source;
  DBuf<u32> z;
  DBuf<i64> w;
};

int bits_for(u64 nvals) {
  int b = 1;
  while (b < 63 && (1ull << b) < nvals) b++;
  return b;
}

}  // namespace

struct sk_server {
  sk_ctx* ctx = nullptr;
  u64 n_edges = 0;
  u64 n_nodes = 0;  // node ids are < n_nodes
  int bits = 1;     // packed edge key width per column
  DBuf<u64> ekeys;
  DBuf<i64> ew;
  DBuf<u64> ecsr;
  std::vector<u8> edge_node;  // host bitmap (byte per node) of ids used by the EDB
  u64 edge_universe = 0;
  u32 max_batch = 64;
  std::vector<int> slot_q;  // slot -> query index, -1 free
  std::vector<std::unique_ptr<Query>> qs;
  std::map<u64, size_t> by_ticket;
  std::deque<size_t> waiting;
  u64 next_ticket = 1;
  // instance cells: cost(instance, Z)
  int zbits = 5, ibits = 1, kbits = 1;
  u32 inst_cap = 0;
  std::vector<int> inst_owner;  // instance -> query index, -1 free
  DBuf<u64> cell;
  DBuf<u32> maxf, touched, bad;
  CellDesc cd;
  DBuf<DevErr> perr;     // per instance
  DBuf<u8> dead_dev;     // per instance
  // delta list: (instance, Z) keys packed with kbits, in cell order
  u64 dn = 0;
  DBuf<u64> dkeys;
  DBuf<i64> dw;
};

namespace {

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

// The head of both rules, cost(instance, Z), goes to the cells: nothing to allocate.
void head_none(void*, OutDesc* o, u64) {
  o->keys = nullptr;
  o->w = nullptr;
  o->inf = nullptr;
}

// One join round: table rows (instance, node, acc) ⋈ edge(node, Z) ⊕-applied
// to cost(instance, Z).
void apply_join(sk_server* s, const Exec& x, SrvTable& t) {
  t.t.part_col = 0;  // an overflow is charged to the row's instance
  t.t.perr = s->perr.get();
  ProbeDesc pd;
  pd.arity = 2;
  pd.bits = s->bits;
  pd.sn = s->n_edges;
  pd.nprobe = 1;
  pd.psrc[0] = 1;
  OutDesc o;
  o.ncols = 2;
  o.bits = s->kbits;
  o.kind[0] = 0;
  o.idx[0] = 0;
  o.kind[1] = 1;
  o.idx[1] = 1;
  o.cd = s->cd;
  join<u64>(x, t.t, pd, s->ekeys.get(), s->ew.get(), nullptr, s->ecsr.get(), &o, true, false, true, nullptr, nullptr,
            head_none, nullptr);
}

void reset_rows(sk_server* s, const Exec& x, const std::vector<u32>& inst) {
  const u64 row = 1ull << s->zbits;
  for (u32 i : inst) {
    memset_async(x, s->cell.get() + (u64)i * row, 0xFF, row * 8);
    memset_async(x, s->maxf.get() + (u64)i * row / 32, 0, row / 8);
  }
}

void release_rows(sk_server* s, Query& q) {
  for (u32 i : q.inst) s->inst_owner[i] = -1;
  q.inst.clear();
}

// Answer of a finished query: its rows in source order (sorted by (S, Z)).
void retire(sk_server* s, Query& q) {
  Exec x = s->ctx->c.exec();
  const u64 row = 1ull << s->zbits;
  q.per_source.assign(q.sources.size(), 0);
  if (q.targets.empty()) {
    std::vector<DBuf<u64>> keys(q.inst.size());
    std::vector<DBuf<i64>> ws(q.inst.size());
    DBuf<u8> none;
    u64 total = 0;
    for (size_t j = 0; j < q.inst.size(); j++) {
      q.per_source[j] =
          cells_compact<u64>(x, s->cd, row, false, s->kbits, &keys[j], &ws[j], &none, false, (u64)q.inst[j] * row);
      total += q.per_source[j];
    }
    q.n = total;
    q.z.reset(x.pool, total ? total : 1);
    q.w.reset(x.pool, total ? total : 1);
    DBuf<u32> scratch(x.pool, total ? total : 1);
    u64 off = 0;
    for (size_t j = 0; j < q.inst.size(); j++) {
      u64 m = q.per_source[j];
      if (!m) continue;
      u32* cp[2] = {scratch.get() + off, q.z.get() + off};
      unpack_cols<u64>(x, keys[j].get(), m, 2, s->kbits, cp);
      SK_HIP(hipMemcpyAsync(q.w.get() + off, ws[j].get(), m * 8, hipMemcpyDeviceToDevice, x.stream));
      off += m;
    }
  } else {
    // ans(S,Z) :- cost(S,Z), tgt(Z): read only the target cells
    const u64 nt = q.targets.size();
    std::vector<u64> idx;
    idx.reserve(q.inst.size() * nt);
    for (u32 i : q.inst)
      for (u32 t : q.targets) idx.push_back((u64)i * row + t);
    DBuf<u64> didx(x.pool, idx.size());
    DBuf<i64> gw(x.pool, idx.size());
    DBuf<u8> gp(x.pool, idx.size());
    SK_HIP(hipMemcpyAsync(didx.get(), idx.data(), idx.size() * 8, hipMemcpyHostToDevice, x.stream));
    cells_gather(x, s->cd, didx.get(), idx.size(), gw.get(), gp.get());
    std::vector<i64> hw(idx.size());
    std::vector<u8> hp(idx.size());
    SK_HIP(hipMemcpyAsync(hw.data(), gw.get(), idx.size() * 8, hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipMemcpyAsync(hp.data(), gp.get(), idx.size(), hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    std::vector<u32> oz;
    std::vector<i64> ow;
    for (size_t j = 0; j < q.inst.size(); j++)
      for (u64 k = 0; k < nt; k++)
        if (hp[j * nt + k]) {
          oz.push_back(q.targets[k]);
          ow.push_back(hw[j * nt + k]);
          q.per_source[j]++;
        }
    q.n = oz.size();
    q.z.reset(x.pool, q.n ? q.n : 1);
    q.w.reset(x.pool, q.n ? q.n : 1);
    if (q.n) {
      SK_HIP(hipMemcpyAsync(q.z.get(), oz.data(), q.n * 4, hipMemcpyHostToDevice, x.stream));
      SK_HIP(hipMemcpyAsync(q.w.get(), ow.data(), q.n * 8, hipMemcpyHostToDevice, x.stream));
    }
  }
  SK_HIP(hipStreamSynchronize(x.stream));
  q.state = 2;
  q.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.arrival).count();
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
  try {
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    Exec x = c.exec();
    u32 maxv = 0;
    for (u64 i = 0; i < n_edges; i++) maxv = std::max(maxv, std::max(src[i], dst[i]));
    s->n_nodes = (u64)maxv + 1;
    s->bits = bits_for(s->n_nodes);
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
    SK_HIP(hipStreamSynchronize(c.stream));
    a.release(); b.release(); k1.release(); k2.release(); w1.release(); w2.release();
    // instance rows from the byte budget (default: half the pool's budget)
    s->zbits = std::max(s->bits, 5);
    const u64 row = 1ull << s->zbits;
    const u64 row_bytes = row * 8 + row / 8 * 2;
    const u64 budget = query_budget_bytes ? query_budget_bytes : c.pool->budget() / 2;
    u64 cap = budget / row_bytes;
    cap = std::min<u64>(cap, 4096);
    if (cap == 0) throw BudgetError{row_bytes};
    s->inst_cap = (u32)cap;
    s->ibits = bits_for(std::max<u64>(cap, 2));
    s->kbits = std::max(s->ibits, s->zbits);
    if (2 * s->kbits > 64) throw EvalErr{SK_ERR_UNSUPPORTED_SHAPE};
    s->inst_owner.assign(cap, -1);
    const u64 ncells = cap * row;
    s->cell.reset(c.pool.get(), ncells);
    s->maxf.reset(c.pool.get(), ncells / 32);
    s->touched.reset(c.pool.get(), ncells / 32);
    s->bad.reset(c.pool.get(), 1);
    memset_async(x, s->cell.get(), 0xFF, ncells * 8);
    memset_async(x, s->maxf.get(), 0, ncells / 8);
    memset_async(x, s->touched.get(), 0, ncells / 8);
    memset_async(x, s->bad.get(), 0, 4);
    CellDesc& cd = s->cd;
    cd.on = 1;
    cd.trop = 1;
    cd.ncols = 2;
    cd.ncodes = std::max<u64>(s->n_nodes, cap);
    cd.ebits[0] = s->ibits;
    cd.ebits[1] = s->zbits;
    cd.shift[0] = s->zbits;
    cd.shift[1] = 0;
    cd.cell = s->cell.get();
    cd.maxf = s->maxf.get();
    cd.touched = s->touched.get();
    cd.bad = s->bad.get();
    s->perr.reset(c.pool.get(), cap);
    s->dead_dev.reset(c.pool.get(), cap);
    memset_async(x, s->perr.get(), 0, sizeof(DevErr) * cap);
    SK_HIP(hipStreamSynchronize(c.stream));
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  } catch (const EvalErr& ee) {
    return (sk_status)ee.code;
  }
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
  // a query needs one instance row per source: larger ones can never run
  if (q->sources.size() > s->inst_cap) return SK_ERR_OUT_OF_BUDGET;
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
    // ---- admission: a free slot and a free instance row per source
    u64 free_rows = 0;
    for (int o : s->inst_owner) free_rows += o < 0;
    std::vector<size_t> admitted;
    while (!s->waiting.empty()) {
      int slot = -1;
      for (u32 i = 0; i < s->max_batch; i++)
        if (s->slot_q[i] < 0) {
          slot = (int)i;
          break;
        }
      if (slot < 0) break;
      const size_t qi = s->waiting.front();
      Query& q = *s->qs[qi];
      if (q.sources.size() > free_rows) break;  // in arrival order: wait for rows to free up
      q.inst.clear();
      for (u32 i = 0; i < s->inst_cap && q.inst.size() < q.sources.size(); i++)
        if (s->inst_owner[i] < 0) {
          s->inst_owner[i] = (int)qi;
          q.inst.push_back(i);
        }
      free_rows -= q.sources.size();
      reset_rows(s, x, q.inst);
      q.slot = slot;
      q.state = 1;
      s->slot_q[slot] = (int)qi;
      admitted.push_back(qi);
      s->waiting.pop_front();
    }
    // ---- this round's derivations, applied to the cells
    if (!admitted.empty()) {
      // bootstrap: cost(S,Y) :- src(S), edge(S,Y); the src fact carries weight 0
      std::vector<u32> ic, sc;
      for (size_t qi : admitted) {
        Query& q = *s->qs[qi];
        for (size_t j = 0; j < q.sources.size(); j++) {
          ic.push_back(q.inst[j]);
          sc.push_back(q.sources[j]);
        }
      }
      SrvTable t;
      t.pool = x.pool;
      t.t.ncols = 2;
      table_alloc_srv(&t, &t.t, ic.size());
      SK_HIP(hipMemcpyAsync(t.t.col[0], ic.data(), ic.size() * 4, hipMemcpyHostToDevice, x.stream));
      SK_HIP(hipMemcpyAsync(t.t.col[1], sc.data(), sc.size() * 4, hipMemcpyHostToDevice, x.stream));
      memset_async(x, t.t.w, 0, ic.size() * 8);
      apply_join(s, x, t);
    }
    if (s->dn) {
      // delta round: cost(S,Z) :- cost(S,Y), edge(Y,Z) over last round's delta
      SrvTable t;
      t.pool = x.pool;
      t.t.ncols = 2;
      table_alloc_srv(&t, &t.t, s->dn);
      u32* cp[2] = {t.t.col[0], t.t.col[1]};
      unpack_cols<u64>(x, s->dkeys.get(), s->dn, 2, s->kbits, cp);
      SK_HIP(hipMemcpyAsync(t.t.w, s->dw.get(), s->dn * 8, hipMemcpyDeviceToDevice, x.stream));
      apply_join(s, x, t);
    }
    // ---- next delta: the improved cells (instance-major, so grouped by instance)
    u32 hi = 0;
    for (u32 i = 0; i < s->inst_cap; i++)
      if (s->inst_owner[i] >= 0) hi = i + 1;
    {
      DBuf<u64> nk;
      DBuf<i64> nw;
      DBuf<u8> nf;
      const u64 row = 1ull << s->zbits;
      s->dn = hi ? cells_compact<u64>(x, s->cd, (u64)hi * row, true, s->kbits, &nk, &nw, &nf, false, 0) : 0;
      s->dkeys = std::move(nk);
      s->dw = std::move(nw);
    }
    // ---- per-query bookkeeping: overflow, round count, cap, retirement
    std::vector<DevErr> herr(s->inst_cap);
    SK_HIP(hipMemcpyAsync(herr.data(), s->perr.get(), s->inst_cap * sizeof(DevErr), hipMemcpyDeviceToHost, x.stream));
    std::vector<u64> dcnt(s->inst_cap, 0);
    if (s->dn) {
      DBuf<u64> csr(x.pool, s->inst_cap + 1);
      build_csr<u64>(x, s->dkeys.get(), s->dn, s->kbits, s->inst_cap, csr.get());
      std::vector<u64> h(s->inst_cap + 1);
      SK_HIP(hipMemcpyAsync(h.data(), csr.get(), (s->inst_cap + 1) * 8, hipMemcpyDeviceToHost, x.stream));
      SK_HIP(hipStreamSynchronize(x.stream));
      for (u32 i = 0; i < s->inst_cap; i++) dcnt[i] = h[i + 1] - h[i];
    }
    SK_HIP(hipStreamSynchronize(x.stream));
    bool clear_err = false, any_dead = false;
    std::vector<u8> dead(s->inst_cap, 0);
    for (u32 p = 0; p < s->max_batch; p++) {
      int qi = s->slot_q[p];
      if (qi < 0) continue;
      Query& q = *s->qs[qi];
      const bool newly = std::find(admitted.begin(), admitted.end(), (size_t)qi) != admitted.end();
      u64 qd = 0;
      for (u32 i : q.inst) {
        qd += dcnt[i];
        if (herr[i].code && q.err.code == 0) {
          q.err.code = SK_ERR_OVERFLOW;
          q.err.lhs = herr[i].lhs;
          q.err.rhs = herr[i].rhs;
          q.state = 3;
        }
        if (herr[i].code) clear_err = true;
      }
      if (q.state == 1 && !newly && q.had_delta) {
        q.iters++;
        if (q.iters > q.universe + 2) {
          q.err.code = SK_ERR_NEGATIVE_WEIGHT_CYCLE;
          q.state = 3;
        }
      }
      q.had_delta = qd > 0;
      if (q.state == 1 && !q.had_delta) retire(s, q);
      if (q.state != 1) {
        if (q.state == 3) {
          q.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.arrival).count();
          for (u32 i : q.inst) dead[i] = 1;
          any_dead = any_dead || qd > 0;
        }
        release_rows(s, q);
        s->slot_q[p] = -1;
      }
    }
    if (clear_err) memset_async(x, s->perr.get(), 0, sizeof(DevErr) * s->inst_cap);
    if (any_dead) {
      // a failed query's delta rows must not feed the next round
      SK_HIP(hipMemcpyAsync(s->dead_dev.get(), dead.data(), s->inst_cap, hipMemcpyHostToDevice, x.stream));
      DBuf<u64> nk(x.pool, s->dn);
      DBuf<i64> nw(x.pool, s->dn);
      s->dn = drop_partitions<u64>(x, s->dkeys.get(), s->dw.get(), nullptr, s->dn, 2, s->kbits, s->dead_dev.get(),
                                   nk.get(), nw.get(), nullptr);
      s->dkeys = std::move(nk);
      s->dw = std::move(nw);
    }
    u32 bad = 0;
    SK_HIP(hipMemcpyAsync(&bad, s->bad.get(), 4, hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    if (bad) return SK_ERR_DEVICE;
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
      if (dst) SK_HIP(hipMemcpyAsync(dst, q.z.get(), q.n * 4, hipMemcpyDeviceToHost, x.stream));
      if (w) SK_HIP(hipMemcpyAsync(w, q.w.get(), q.n * 8, hipMemcpyDeviceToHost, x.stream));
      if (src) {
        u64 off = 0;
        for (size_t j = 0; j < q.sources.size(); j++) {
          std::fill(src + off, src + off + q.per_source[j], q.sources[j]);
          off += q.per_source[j];
        }
      }
      SK_HIP(hipStreamSynchronize(x.stream));
    }
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  }
  q.z.release();
  q.w.release();
  return SK_OK;
}

sk_status sk_server_release(sk_server* s, uint64_t ticket) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  q.z.release();
  q.w.release();
  return SK_OK;
}

}  // extern "C"
