// sk_server.cpp — resident-EDB query serving with continuous batching.
//
// The EDB edge(node, node) (Trop) stays on the device, sorted by (src, dst)
// with the minimum weight per pair, with a CSR over src (push joins) and its
// transpose, a CSR over dst with the pull round's chunk list.  A query is
// the multi-source cheapest-cost program
//   cost(S,Y) :- src(S), edge(S,Y).        (stratum 0)
//   cost(S,Z) :- cost(S,Y), edge(Y,Z).     (stratum 0)
// optionally followed by a target filter
//   ans(S,Z) :- cost(S,Z), tgt(Z).         (stratum 1)
//
// Instances.  cost(S,·) depends on S alone (a Jacobi round never mixes
// sources), so every source of a query runs as its own *instance*: a lane of
// a lane group (sk_serve.h), 64 instances side by side over the node rows.
// The number of groups is fixed at creation from the byte budget; lanes are
// handed out per source, lowest first, to the running queries in arrival
// order and then to waiting ones, so a query with more sources than there
// are lanes runs in pieces as lanes free up, with the same answer.  A lane is
// absent everywhere when it is handed out: finishing or failing instances
// reset their lanes.
//
// Rounds.  Every sk_server_step is one semi-naive round for all instances
// together: new instances contribute their bootstrap (src ⋈ edge, weight 0 ⊗
// w), running ones their delta round (Δcost ⋈ edge), both ⊕-applied to the
// cells with strict improvements marked.  A group whose delta would cost the
// push join at least |E| outputs pulls instead: one pass over the in-edge
// index serves all 64 lanes (the dense engine's pull kernel, on the group's
// cells).  One touched compaction over the groups in use gives every
// instance's next delta.  Rounds are Jacobi: the joins read the previous
// delta, never the cells.  Each instance keeps its own round count; a
// query's count is the largest of its instances' (equal to the lock-step
// count, since an instance's delta, once empty, stays empty), its
// Bellman-Ford cap (crates/strata-eval/src/seminaive.rs:94-104) applies per
// instance, and overflows are charged per instance.  An instance whose delta
// is empty is done: its lane's cells (or, with a target filter, its target
// cells) become the answer part of its source.  A query is done when every
// source's part is; parts can also be taken one by one, in source order,
// while the rest still run.
//
// Cells.  When every edge weight lies in [0, 2^26) the cells start narrow
// (u32 weights) with packed deltas, the form the pull reads; a sum reaching
// 2^26 raises the widen flag, and the cells are widened in place to u64
// codes (their buffer is sized for that) and the round is applied again by
// push joins (⊕ is idempotent).  Other EDBs run wide from the start, push
// only.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "sk_engine.h"
#include "sk_join.h"
#include "sk_serve.h"
#include "strata_k.h"

using namespace sk;

namespace {

using Clock = std::chrono::steady_clock;

// One source's answer: rows (Z, w) sorted by Z.
struct Part {
  u64 n = 0;
  bool ready = false, taken = false;
  DBuf<u32> z;  // unfiltered answers stay on the device until taken
  DBuf<i64> w;
  std::vector<u32> hz;  // filtered answers (at most |targets| rows) on the host
  std::vector<i64> hw;
};

struct Query {
  u64 ticket = 0;
  std::vector<u32> sources, targets;
  Clock::time_point arrival;
  int state = 0;  // 0 waiting, 1 active, 2 done, 3 error
  int slot = -1;
  u32 next = 0;       // sources[next..] have no instance yet
  u32 finished = 0;   // sources whose part is ready
  u32 next_take = 0;  // the next part to hand out in source order
  u64 held = 0;       // device bytes of ready parts not yet taken
  u64 iters = 0;
  u64 universe = 0;
  EvalErr err;
  double latency_ms = 0;
  u64 n = 0;  // rows over all parts
  std::vector<Part> parts;
};

// State of one lane (instance slot).
struct Lane {
  int q = -1;      // owning query, -1 free
  u32 src = 0;     // index into the query's sources
  u64 rounds = 0;  // delta rounds run
  bool fresh = false;      // bootstrap in the current step
  bool had_delta = false;  // non-empty delta after the last round
  bool poisoned = false;   // a failed instance's lane, freed after one more round
};

int bits_for(u64 nvals) {
  int b = 1;
  while (b < 63 && (1ull << b) < nvals) b++;
  return b;
}

constexpr u64 kTableSlice = 1ull << 27;  // push-join table rows per launch sequence

}  // namespace

struct sk_server {
  sk_ctx* ctx = nullptr;
  u64 n_edges = 0;
  u64 n_nodes = 0;  // node ids are < n_nodes
  int bits = 1;     // packed edge key width per column
  DBuf<u64> ekeys;  // out-edges (src, dst), sorted
  DBuf<i64> ew;
  DBuf<u64> ecsr;
  DBuf<u32> iadj;  // in-edges sorted by (dst, src): the src of each
  DBuf<i64> iw;
  DBuf<PullChunk> ichunks;
  u64 ichunks_n = 0;
  u64 irows = 0;  // nodes with an in-edge
  std::vector<u8> edge_node;  // host bitmap (byte per node) of ids used by the EDB
  u64 edge_universe = 0;
  u32 max_batch = 64;
  std::vector<int> slot_q;  // slot -> query index, -1 free
  std::vector<std::unique_ptr<Query>> qs;
  std::map<u64, size_t> by_ticket;
  std::deque<size_t> waiting;
  std::vector<size_t> active;  // admitted queries in arrival order
  u64 next_ticket = 1;
  u64 hold = 0;  // lanes are withheld from a query holding more untaken part bytes
  // lane groups
  int zbits = 8;  // rows per group = 2^zbits >= max(n_nodes, 256)
  int gbits = 1;
  u32 ngroups = 0;
  u64 gcells = 0;   // cells per group = 2^(zbits + 6)
  bool narrow = false;
  DBuf<u64> cells;  // ngroups * gcells words; narrow cells are its first half as u32
  DBuf<u32> maxf, touched, widen, bad;
  DBuf<u64> rowrec8;
  CellDesc cd;  // three columns (group, node, lane) over all groups
  std::vector<Lane> lanes;
  DBuf<DevErr> perr;  // per lane
  // the last round's delta, over groups [0, dhi)
  u64 dn = 0;
  u32 dhi = 0;
  DBuf<u32> dpack;  // narrow
  DBuf<i64> dw;     // wide
  DBuf<u8> dlane;   // wide
  bool dpacked = false;  // the delta is in packed form (dpack) rather than (dw, dlane)
  std::vector<u64> gcost, gstart;
  // counters
  u64 steps = 0, pulls = 0, pushes = 0;
  u64 pull_cand = 0, push_out = 0;  // delta rounds' (edge, delta cell) pairs by pull and by push
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

// The cells of group g as the two-column view lane | node << 6.
CellDesc group_cd(const sk_server* s, u32 g) {
  CellDesc c = s->cd;
  c.ncols = 2;
  c.ebits[0] = 6;
  c.shift[0] = 0;
  c.ebits[1] = s->zbits;
  c.shift[1] = 6;
  c.ebits[2] = 0;
  c.shift[2] = 0;
  const u64 off = (u64)g * s->gcells;
  if (s->narrow) c.cell32 = (u32*)s->cells.get() + off;
  else c.cell = s->cells.get() + off;
  c.maxf = s->maxf.get() + off / 32;
  c.touched = s->touched.get() + off / 32;
  c.rowrec8 = nullptr;
  c.dpack = nullptr;
  return c;
}

// A join round: table rows (group, node, lane, instance; acc) ⋈ edge(node, Z)
// ⊕-applied to cost(instance, Z) (the instance column carries overflows).
void apply_join(sk_server* s, const Exec& x, Table& t) {
  if (!t.n) return;
  t.part_col = 3;
  t.perr = s->perr.get();
  ProbeDesc pd;
  pd.arity = 2;
  pd.bits = s->bits;
  pd.sn = s->n_edges;
  pd.nprobe = 1;
  pd.psrc[0] = 1;
  OutDesc o;
  o.ncols = 3;  // cells only: no packed key is written
  o.kind[0] = 0;
  o.idx[0] = 0;
  o.kind[1] = 1;
  o.idx[1] = 1;
  o.kind[2] = 0;
  o.idx[2] = 2;
  o.cd = s->cd;
  join<u64>(x, t, pd, s->ekeys.get(), s->ew.get(), nullptr, s->ecsr.get(), &o, true, false, true, nullptr, nullptr,
            head_none, nullptr);
}

// Bootstrap rows of the instances given lanes in this step.
void bootstrap(sk_server* s, const Exec& x, const std::vector<u32>& fresh) {
  if (fresh.empty()) return;
  const u64 n = fresh.size();
  std::vector<u32> hg(n), hs(n), hl(n), hi(n);
  for (u64 i = 0; i < n; i++) {
    const u32 ln = fresh[i];
    const Lane& L = s->lanes[ln];
    hg[i] = ln / 64;
    hl[i] = ln % 64;
    hi[i] = ln;
    hs[i] = s->qs[L.q]->sources[L.src];
  }
  SrvTable t;
  t.pool = x.pool;
  t.t.ncols = 4;
  table_alloc_srv(&t, &t.t, n);
  SK_HIP(hipMemcpyAsync(t.t.col[0], hg.data(), n * 4, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipMemcpyAsync(t.t.col[1], hs.data(), n * 4, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipMemcpyAsync(t.t.col[2], hl.data(), n * 4, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipMemcpyAsync(t.t.col[3], hi.data(), n * 4, hipMemcpyHostToDevice, x.stream));
  memset_async(x, t.t.w, 0, n * 8);
  apply_join(s, x, t.t);
  s->pushes++;
  SK_HIP(hipStreamSynchronize(x.stream));  // the host vectors above must outlive the copies
}

// The delta round of the groups listed, as push joins over table slices.
void push_groups(sk_server* s, const Exec& x, const std::vector<u32>& gs) {
  std::vector<std::pair<u32, std::pair<u64, u64>>> ranges;  // group, [p0, p1)
  for (u32 g : gs) {
    const u64 p0 = s->gstart[g], p1 = g + 1 < s->dhi ? s->gstart[g + 1] : s->dn;
    for (u64 a = p0; a < p1; a += kTableSlice) ranges.push_back({g, {a, std::min(p1, a + kTableSlice)}});
  }
  size_t i = 0;
  while (i < ranges.size()) {
    u64 tot = 0;
    size_t j = i;
    while (j < ranges.size() && (j == i || tot + (ranges[j].second.second - ranges[j].second.first) <= kTableSlice)) {
      tot += ranges[j].second.second - ranges[j].second.first;
      j++;
    }
    SrvTable t;
    t.pool = x.pool;
    t.t.ncols = 4;
    table_alloc_srv(&t, &t.t, tot);
    u64 at = 0;
    for (size_t k = i; k < j; k++) {
      const u32 g = ranges[k].first;
      const u64 p0 = ranges[k].second.first, p1 = ranges[k].second.second;
      srv_table(x, s->rowrec8.get(), s->zbits, g, s->dpacked ? s->dpack.get() : nullptr,
                s->dpacked ? nullptr : s->dw.get(), s->dpacked ? nullptr : s->dlane.get(), p0, p1,
                t.t.col[0] + at, t.t.col[2] + at, t.t.col[1] + at, t.t.col[3] + at, t.t.w + at);
      at += p1 - p0;
    }
    apply_join(s, x, t.t);
    s->pushes++;
    i = j;
  }
}

}  // namespace

namespace {

// Group g's delta round as a pull over the in-edge index (narrow cells,
// packed delta): all 64 lanes in one pass.
void pull_group(sk_server* s, const Exec& x, u32 g) {
  PullDesc pd;
  pd.nchunks = s->ichunks_n;
  pd.chunks = s->ichunks.get();
  pd.adj = s->iadj.get();
  pd.w = s->iw.get();
  pd.dpack = s->dpack.get();
  pd.rowrec8 = s->rowrec8.get() + ((u64)g << s->zbits);
  pd.lbits = 6;
  pd.delta_first = 1;
  pd.nrows = s->n_nodes;
  // algorithmic bytes: every in-edge (source, weight, the source row's
  // record) once, every candidate's packed delta word once, every
  // destination row's 64 narrow cells once
  const double bytes = (double)s->n_edges * (4.0 + 8.0 + 8.0) + (double)s->gcost[g] * 4.0 + (double)s->irows * 256.0;
  pull_apply(x, pd, group_cd(s, g), true, bytes, s->gcost[g]);
  s->pulls++;
}

void widen_cells(sk_server* s, const Exec& x) {
  srv_widen_in_place(x, s->cells.get(), (u64)s->ngroups * s->gcells);
  s->narrow = false;
  s->cd.narrow = 0;
  s->cd.cell32 = nullptr;
  s->cd.cell = s->cells.get();
  s->cd.widen = nullptr;
  s->cd.narrow_lim = kCellTop32;
  memset_async(x, s->widen.get(), 0, 4);
}

// A group pulls when its push join would emit at least this many rows:
// measured at RMAT scale 24 (263M edges), a pull costs ~5 ms plus ~3.7 ps
// per (edge, delta cell) pair and a push ~60 ps per output row, even at
// about a third of |E|.
#ifndef SK_SRV_PULL_PCT
#define SK_SRV_PULL_PCT 30
#endif
u64 pull_limit(u64 edges) { return edges / 100 * SK_SRV_PULL_PCT + edges % 100 * SK_SRV_PULL_PCT / 100; }

void fail(Query& q, int code, i64 lhs = 0, i64 rhs = 0) {
  if (q.state != 1) return;
  q.state = 3;
  q.err.code = code;
  q.err.lhs = lhs;
  q.err.rhs = rhs;
  q.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.arrival).count();
  for (Part& p : q.parts) {
    p.z.release();
    p.w.release();
  }
  q.held = 0;
}

// Copy part p of q to the host arrays at row `off` and release it.
void copy_part(const Exec& x, Query& q, u32 j, u32* src, u32* dst, int64_t* w, u64 off) {
  Part& p = q.parts[j];
  if (p.n) {
    if (q.targets.empty()) {
      if (dst) SK_HIP(hipMemcpyAsync(dst + off, p.z.get(), p.n * 4, hipMemcpyDeviceToHost, x.stream));
      if (w) SK_HIP(hipMemcpyAsync(w + off, p.w.get(), p.n * 8, hipMemcpyDeviceToHost, x.stream));
    } else {
      if (dst) std::memcpy(dst + off, p.hz.data(), p.n * 4);
      if (w) std::memcpy(w + off, p.hw.data(), p.n * 8);
    }
    if (src) std::fill(src + off, src + off + p.n, q.sources[j]);
  }
}

void drop_part(Query& q, u32 j) {
  Part& p = q.parts[j];
  if (q.targets.empty() && p.ready && !p.taken) q.held -= std::min<u64>(q.held, p.n * 12);
  p.taken = true;
  p.z.release();
  p.w.release();
  p.hz = {};
  p.hw = {};
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
    bool packable = true;
    for (u64 i = 0; i < n_edges; i++) {
      maxv = std::max(maxv, std::max(src[i], dst[i]));
      if (w[i] < 0 || (u64)w[i] >= kPackLim) packable = false;
    }
    s->n_nodes = (u64)maxv + 1;
    s->bits = bits_for(s->n_nodes);
    s->edge_node.assign(s->n_nodes, 0);
    for (u64 i = 0; i < n_edges; i++) s->edge_node[src[i]] = s->edge_node[dst[i]] = 1;
    for (u8 b : s->edge_node) s->edge_universe += b;
    Pool* pool = c.pool.get();
    const u64 m0 = n_edges ? n_edges : 1;
    // resident EDB: pack, sort by (src, dst), min per pair, CSR over src
    {
      DBuf<u32> a(pool, m0), b(pool, m0);
      DBuf<u64> k1(pool, m0), k2(pool, m0);
      DBuf<i64> w1(pool, m0), w2(pool, m0);
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
      s->ekeys.reset(pool, m0);
      s->ew.reset(pool, m0);
      s->n_edges = unique_runs<u64>(x, sk, sw, nullptr, n_edges, s->ekeys.get(), s->ew.get(), nullptr);
      s->ecsr.reset(pool, s->n_nodes + 1);
      build_csr<u64>(x, s->ekeys.get(), s->n_edges, s->bits, s->n_nodes, s->ecsr.get());
      SK_HIP(hipStreamSynchronize(c.stream));
    }
    // its transpose for pull rounds: sorted by (dst, src), CSR over dst
    {
      const u64 m = s->n_edges, mm = m ? m : 1;
      DBuf<u32> ca(pool, mm), cb(pool, mm);
      u32* cp[2] = {ca.get(), cb.get()};
      unpack_cols<u64>(x, s->ekeys.get(), m, 2, s->bits, cp);
      Cols cols{};
      cols.c[0] = cb.get();
      cols.c[1] = ca.get();
      DBuf<u64> k1(pool, mm), k2(pool, mm);
      DBuf<i64> w1(pool, mm), w2(pool, mm);
      pack_cols<u64>(x, cols, 2, s->bits, m, k1.get());
      if (m) SK_HIP(hipMemcpyAsync(w1.get(), s->ew.get(), m * 8, hipMemcpyDeviceToDevice, c.stream));
      u64* sk;
      i64* sw;
      radix_sort<u64>(x, k1.get(), k2.get(), w1.get(), w2.get(), nullptr, nullptr, m, 2 * s->bits, &sk, &sw, nullptr);
      DBuf<u64> icsr(pool, s->n_nodes + 1);
      build_csr<u64>(x, sk, m, s->bits, s->n_nodes, icsr.get());
      s->iadj.reset(pool, mm);
      u32* ip[2] = {ca.get(), s->iadj.get()};
      unpack_cols<u64>(x, sk, m, 2, s->bits, ip);
      s->iw.reset(pool, mm);
      if (m) SK_HIP(hipMemcpyAsync(s->iw.get(), sw, m * 8, hipMemcpyDeviceToDevice, c.stream));
      s->ichunks_n = pull_chunks(x, icsr.get(), s->n_nodes, &s->ichunks, &s->irows);
      SK_HIP(hipStreamSynchronize(c.stream));
    }
    // lane groups from the byte budget (default: half the pool's budget);
    // a group's cells are sized for wide codes, narrow cells use half
    s->zbits = std::max(s->bits, 8);  // blocks of 256 rows stay within a group
    s->gcells = 1ull << (s->zbits + 6);
    const u64 rows = 1ull << s->zbits;
    const u64 per_group = s->gcells * 8 + s->gcells / 4 + rows * 8 + s->gcells * 4;
    const u64 budget = query_budget_bytes ? query_budget_bytes : c.pool->budget() / 2;
    u64 ng = std::min<u64>(budget / per_group, 64);
    if (ng == 0) throw BudgetError{per_group};
    for (;;) {
      try {
        s->cells.reset(pool, ng * s->gcells);
        s->maxf.reset(pool, ng * s->gcells / 32);
        s->touched.reset(pool, ng * s->gcells / 32);
        s->rowrec8.reset(pool, ng * rows);
        break;
      } catch (const BudgetError&) {
        s->cells.release();
        s->maxf.release();
        s->touched.release();
        s->rowrec8.release();
        if (ng <= 1) throw;
        ng--;
      }
    }
    s->ngroups = (u32)ng;
    s->gbits = bits_for(std::max<u64>(ng, 2));
    s->narrow = packable;
    s->widen.reset(pool, 1);
    s->bad.reset(pool, 1);
    memset_async(x, s->cells.get(), 0xFF, ng * s->gcells * 8);
    memset_async(x, s->maxf.get(), 0, ng * s->gcells / 8);
    memset_async(x, s->touched.get(), 0, ng * s->gcells / 8);
    memset_async(x, s->widen.get(), 0, 4);
    memset_async(x, s->bad.get(), 0, 4);
    CellDesc& cd = s->cd;
    cd.on = 1;
    cd.trop = 1;
    cd.ncols = 3;
    cd.ncodes = std::max<u64>(s->n_nodes, std::max<u64>(ng, 64));
    cd.ebits[0] = s->gbits;
    cd.shift[0] = s->zbits + 6;
    cd.ebits[1] = s->zbits;
    cd.shift[1] = 6;
    cd.ebits[2] = 6;
    cd.shift[2] = 0;
    if (s->narrow) {
      cd.narrow = 1;
      cd.cell32 = (u32*)s->cells.get();
      cd.widen = s->widen.get();
      cd.narrow_lim = (u32)kPackLim;
    } else {
      cd.cell = s->cells.get();
    }
    cd.maxf = s->maxf.get();
    cd.touched = s->touched.get();
    cd.bad = s->bad.get();
    cd.rowrec8 = s->rowrec8.get();
    cd.rlbits = 6;
    s->lanes.assign(ng * 64, Lane{});
    s->perr.reset(pool, ng * 64);
    memset_async(x, s->perr.get(), 0, sizeof(DevErr) * ng * 64);
    s->gcost.assign(ng, 0);
    s->gstart.assign(ng, 0);
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
  q->parts.resize(q->sources.size());
  q->ticket = s->next_ticket++;
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
    s->steps++;
    const u32 L = s->ngroups * 64;
    // ---- lanes: running queries first (arrival order), then waiting ones
    std::vector<u32> freel;
    for (u32 i = 0; i < L; i++)
      if (s->lanes[i].q < 0 && !s->lanes[i].poisoned) freel.push_back(i);
    size_t fi = 0;
    std::vector<u32> fresh;
    auto give = [&](size_t qi) {
      Query& q = *s->qs[qi];
      if (s->hold && q.held > s->hold) return;
      while (q.next < q.sources.size() && fi < freel.size()) {
        const u32 ln = freel[fi++];
        Lane& Ln = s->lanes[ln];
        Ln = Lane{};
        Ln.q = (int)qi;
        Ln.src = q.next++;
        Ln.fresh = true;
        fresh.push_back(ln);
      }
    };
    for (size_t qi : s->active)
      if (s->qs[qi]->state == 1) give(qi);
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
      if (!q.sources.empty() && fi >= freel.size()) break;  // in arrival order: wait for lanes
      s->waiting.pop_front();
      q.slot = slot;
      q.state = 1;
      s->slot_q[slot] = (int)qi;
      s->active.push_back(qi);
      give(qi);
    }
    // ---- the round: bootstraps, then each group's delta by pull or push
    auto round = [&](bool allow_pull) {
      bootstrap(s, x, fresh);
      std::vector<u32> pushg;
      for (u32 g = 0; g < s->dhi; g++) {
        if (!s->gcost[g]) continue;  // no delta, or no out-edges from it
        if (allow_pull && s->narrow && s->dpacked && s->gcost[g] >= pull_limit(s->n_edges)) {
          pull_group(s, x, g);
          s->pull_cand += s->gcost[g];
        } else {
          pushg.push_back(g);
          s->push_out += s->gcost[g];
        }
      }
      push_groups(s, x, pushg);
    };
    round(true);
    if (s->narrow) {
      u32 wf = 0;
      SK_HIP(hipMemcpyAsync(&wf, s->widen.get(), 4, hipMemcpyDeviceToHost, x.stream));
      SK_HIP(hipStreamSynchronize(x.stream));
      if (wf) {
        // a sum left the narrow range: wide cells, and the round again
        widen_cells(s, x);
        round(false);
      }
    }
    // ---- next delta: the touched cells of the groups in use
    u32 hi = 0;
    for (u32 i = 0; i < L; i++)
      if (s->lanes[i].q >= 0 || s->lanes[i].poisoned) hi = i / 64 + 1;
    const u64 rows = 1ull << s->zbits;
    std::vector<u64> hmask;
    {
      DBuf<u32> np;
      DBuf<i64> nw;
      DBuf<u8> nl;
      u64 n = 0;
      // with it, per group: the lanes holding a delta cell, the push cost
      // (delta cells times their node's out-degree) and the first position
      DBuf<u64> lm, gc, gs;
      if (hi) {
        lm.reset(x.pool, hi);
        gc.reset(x.pool, hi);
        gs.reset(x.pool, hi);
        memset_async(x, lm.get(), 0, (u64)hi * 8);
        memset_async(x, gc.get(), 0, (u64)hi * 8);
        CellDesc cc = s->cd;
        cc.rnrows = (u64)hi * rows;
        cc.pc_csr = s->ecsr.get();
        cc.pc_nv = s->n_nodes;
        cc.grp_shift = s->zbits;
        cc.grp_mask = lm.get();
        cc.grp_cost = gc.get();
        cc.grp_start = gs.get();
        n = s->narrow ? cells_compact_rows(x, cc, (u64)hi * s->gcells, &np, nullptr, nullptr)
                      : cells_compact_rows(x, cc, (u64)hi * s->gcells, nullptr, &nw, &nl);
      }
      s->dpack = std::move(np);
      s->dw = std::move(nw);
      s->dlane = std::move(nl);
      s->dpacked = s->narrow;
      s->dn = n;
      s->dhi = hi;
      std::fill(s->gcost.begin(), s->gcost.end(), 0);
      std::fill(s->gstart.begin(), s->gstart.end(), 0);
      hmask.assign(hi, 0);
      if (hi) {
        SK_HIP(hipMemcpyAsync(hmask.data(), lm.get(), (u64)hi * 8, hipMemcpyDeviceToHost, x.stream));
        SK_HIP(hipMemcpyAsync(s->gcost.data(), gc.get(), (u64)hi * 8, hipMemcpyDeviceToHost, x.stream));
        SK_HIP(hipMemcpyAsync(s->gstart.data(), gs.get(), (u64)hi * 8, hipMemcpyDeviceToHost, x.stream));
        SK_HIP(hipStreamSynchronize(x.stream));
      }
    }
    std::vector<DevErr> herr;
    u32 bad = 0;
    {
      if (!s->narrow && hi) {
        herr.resize((u64)hi * 64);
        SK_HIP(hipMemcpyAsync(herr.data(), s->perr.get(), (u64)hi * 64 * sizeof(DevErr), hipMemcpyDeviceToHost,
                              x.stream));
      }
      SK_HIP(hipMemcpyAsync(&bad, s->bad.get(), 4, hipMemcpyDeviceToHost, x.stream));
      SK_HIP(hipStreamSynchronize(x.stream));
    }
    if (bad) return SK_ERR_DEVICE;
    // ---- per instance: overflow, round count and cap
    bool clear_err = false;
    for (u32 i = 0; i < hi * 64; i++) {
      Lane& Ln = s->lanes[i];
      if (!herr.empty() && herr[i].code) clear_err = true;
      if (Ln.q < 0) continue;
      Query& q = *s->qs[Ln.q];
      if (!herr.empty() && herr[i].code) fail(q, SK_ERR_OVERFLOW, herr[i].lhs, herr[i].rhs);
      if (q.state != 1) continue;
      if (!Ln.fresh && Ln.had_delta) {
        Ln.rounds++;
        if (Ln.rounds > q.universe + 2) fail(q, SK_ERR_NEGATIVE_WEIGHT_CYCLE);
      }
      q.iters = std::max(q.iters, Ln.rounds);
    }
    if (clear_err) memset_async(x, s->perr.get(), 0, sizeof(DevErr) * (u64)hi * 64);
    // ---- finished instances become answer parts; failed ones release their lanes
    std::vector<u64> take(hi, 0), clear(hi, 0), poison(hi, 0);
    std::vector<u32> filt;
    for (u32 i = 0; i < hi * 64; i++) {
      Lane& Ln = s->lanes[i];
      const u32 g = i / 64;
      const u64 bit = 1ull << (i % 64);
      const bool d = (hmask[g] >> (i % 64)) & 1;
      if (Ln.q < 0) {
        if (Ln.poisoned && !d) clear[g] |= bit;  // its last delta died out
        continue;
      }
      Query& q = *s->qs[Ln.q];
      if (q.state != 1) {
        if (d) poison[g] |= bit;
        else clear[g] |= bit;
        continue;
      }
      Ln.had_delta = d;
      Ln.fresh = false;
      if (d) continue;
      if (q.targets.empty()) take[g] |= bit;
      else filt.push_back(i);
      clear[g] |= bit;
    }
    if (!filt.empty()) {
      // ans(S,Z) :- cost(S,Z), tgt(Z): read only the target cells
      std::vector<u64> idx;
      for (u32 i : filt) {
        const Query& q = *s->qs[s->lanes[i].q];
        const u64 base = (u64)(i / 64) * s->gcells + (i % 64);
        for (u32 t : q.targets) idx.push_back(base + ((u64)t << 6));
      }
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
      u64 k = 0;
      for (u32 i : filt) {
        Lane& Ln = s->lanes[i];
        Query& q = *s->qs[Ln.q];
        Part& p = q.parts[Ln.src];
        for (u32 t : q.targets) {
          if (hp[k]) {
            p.hz.push_back(t);
            p.hw.push_back(hw[k]);
          }
          k++;
        }
        p.n = p.hz.size();
        p.ready = true;
        q.finished++;
        q.n += p.n;
      }
    }
    for (u32 g = 0; g < hi; g++) {
      if (poison[g]) {
        DBuf<u32> oz[64];
        DBuf<i64> ow[64];
        u64 cnt[64];
        srv_lanes(x, group_cd(s, g), s->n_nodes, 0, poison[g], true, oz, ow, cnt);
      }
      if (!(take[g] | clear[g])) continue;
      DBuf<u32> oz[64];
      DBuf<i64> ow[64];
      u64 cnt[64];
      u64 tk = take[g];
      try {
        srv_lanes(x, group_cd(s, g), s->n_nodes, tk, clear[g], false, oz, ow, cnt);
      } catch (const BudgetError&) {
        // no room for these answers: their queries fail, the lanes are reset
        for (int l = 0; l < 64; l++)
          if ((tk >> l) & 1) fail(*s->qs[s->lanes[g * 64 + l].q], SK_ERR_OUT_OF_BUDGET);
        tk = 0;
        srv_lanes(x, group_cd(s, g), s->n_nodes, 0, clear[g], false, oz, ow, cnt);
      }
      for (int l = 0; l < 64; l++) {
        if (!((tk >> l) & 1)) continue;
        Lane& Ln = s->lanes[g * 64 + l];
        Query& q = *s->qs[Ln.q];
        Part& p = q.parts[Ln.src];
        p.n = cnt[l];
        p.z = std::move(oz[l]);
        p.w = std::move(ow[l]);
        p.ready = true;
        q.finished++;
        q.n += p.n;
        q.held += p.n * 12;
      }
    }
    for (u32 i = 0; i < hi * 64; i++) {
      const u64 bit = 1ull << (i % 64);
      Lane& Ln = s->lanes[i];
      if (poison[i / 64] & bit) {
        Ln = Lane{};
        Ln.poisoned = true;
      } else if (clear[i / 64] & bit) {
        Ln = Lane{};
      }
    }
    // ---- queries: done when every part is ready; finished ones free their slot
    std::vector<size_t> still;
    for (size_t qi : s->active) {
      Query& q = *s->qs[qi];
      if (q.state == 1 && q.finished == q.sources.size()) {
        q.state = 2;
        q.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.arrival).count();
      }
      if (q.state == 1) {
        still.push_back(qi);
        continue;
      }
      if (q.slot >= 0) s->slot_q[q.slot] = -1;
      q.slot = -1;
    }
    s->active.swap(still);
    SK_HIP(hipStreamSynchronize(x.stream));
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  }
  if (n_active) *n_active = (uint32_t)s->active.size();
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
    u64 off = 0;
    for (u32 j = q.next_take; j < q.sources.size(); j++) {
      copy_part(x, q, j, src, dst, w, off);
      off += q.parts[j].n;
    }
    SK_HIP(hipStreamSynchronize(x.stream));
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  }
  for (u32 j = q.next_take; j < q.sources.size(); j++) drop_part(q, j);
  q.next_take = (u32)q.sources.size();
  return SK_OK;
}

sk_status sk_server_release(sk_server* s, uint64_t ticket) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  for (u32 j = 0; j < q.sources.size(); j++) drop_part(q, j);
  q.next_take = (u32)q.sources.size();
  return SK_OK;
}

sk_status sk_server_poll_source(sk_server* s, uint64_t ticket, uint32_t* index, uint32_t* source, uint64_t* n_rows) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  if (q.state == 3) return (sk_status)q.err.code;
  if (q.next_take >= q.sources.size()) {
    if (q.state != 2) return SK_ERR_NOT_READY;  // an empty source set: done at admission
    if (index) *index = (uint32_t)q.sources.size();
    if (n_rows) *n_rows = 0;
    return SK_OK;
  }
  const Part& p = q.parts[q.next_take];
  if (!p.ready) return SK_ERR_NOT_READY;
  if (index) *index = q.next_take;
  if (source) *source = q.sources[q.next_take];
  if (n_rows) *n_rows = p.n;
  return SK_OK;
}

sk_status sk_server_take_source(sk_server* s, uint64_t ticket, uint32_t* dst, int64_t* w) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return SK_ERR_INVALID_ARGUMENT;
  Query& q = *s->qs[it->second];
  if (q.state == 3) return (sk_status)q.err.code;
  if (q.next_take >= q.sources.size()) return SK_ERR_INVALID_ARGUMENT;
  if (!q.parts[q.next_take].ready) return SK_ERR_NOT_READY;
  try {
    Exec x = s->ctx->c.exec();
    copy_part(x, q, q.next_take, nullptr, dst, w, 0);
    SK_HIP(hipStreamSynchronize(x.stream));
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  }
  drop_part(q, q.next_take);
  q.next_take++;
  return SK_OK;
}

sk_status sk_server_set_hold(sk_server* s, uint64_t bytes) {
  if (!s) return SK_ERR_INVALID_ARGUMENT;
  s->hold = bytes;
  return SK_OK;
}

sk_status sk_server_stats(const sk_server* s, uint64_t* out, uint32_t n) {
  if (!s || (n && !out)) return SK_ERR_INVALID_ARGUMENT;
  const uint64_t v[9] = {s->steps,    s->pulls,      s->pushes,   (u64)s->ngroups * 64, s->narrow ? 1ull : 0ull,
                         s->ngroups, s->n_edges, s->pull_cand, s->push_out};
  for (uint32_t i = 0; i < n && i < 9; i++) out[i] = v[i];
  return SK_OK;
}

}  // extern "C"
