// sk_batch.cpp — many independent programs evaluated together.
//
// Programs whose Core-IR is identical up to predicate names form a group.  A
// group is evaluated as ONE fused program: every atom gets a leading
// program-id column bound to a shared variable, so the fused relations are
// the disjoint union of the programs' relations keyed by program id, every
// round is one set of launches over all programs of the group, and the
// rounds end when the last program converges.  Per program the engine keeps
// the reference's own bookkeeping: its round count (rounds in which its delta
// was non-empty), its Bellman-Ford cap (its own universe + 2) and its errors;
// a program that fails stops contributing without disturbing the others.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "sk_engine.h"
#include "strata_k.h"

using namespace sk;

namespace {

struct BProgram {
  Program prog;  // own copy
  // the caller's fact buffers, borrowed until sk_batch_load returns
  std::vector<std::string> edb_preds;
  std::vector<u64> edb_n;
  std::vector<std::vector<const u32*>> edb_cols;
  std::vector<const i64*> edb_w;
  int group = -1;
  int member = -1;  // index within the group
  EvalErr err;
  std::vector<u64> iters;
  u64 universe = 0;
};

// Structural signature of a program up to predicate names.
std::string signature(const Program& p) {
  std::string s = "S" + std::to_string(p.num_strata) + ";";
  for (auto& d : p.preds) s += "p" + std::to_string(d.arity) + "," + std::to_string(d.sem) + "," + std::to_string(d.stratum) + ";";
  auto atom = [&](const Atom& a) {
    std::string t = (a.neg ? "!" : "") + std::to_string(a.rel) + "(";
    for (auto& x : a.args) t += std::to_string(x.kind) + ":" + std::to_string(x.a) + ":" + std::to_string(x.b) + ",";
    return t + ")";
  };
  for (auto& r : p.rules) {
    s += "r" + std::to_string(r.stratum) + "," + std::to_string(r.var_count) + ":" + atom(r.head) + ":-";
    for (auto& b : r.body) s += atom(b);
    s += ";";
  }
  return s;
}

}  // namespace

struct sk_batch {
  sk_ctx* ctx = nullptr;
  std::vector<BProgram> progs;
  // Groups run concurrently, each from its own host thread on its own
  // context (stream, pool, error record): their kernels overlap on the
  // device.  The context outlives the engine that uses it.
  struct Group {
    std::vector<int> members;
    std::unique_ptr<Ctx> gc;
    std::unique_ptr<Engine> eng;
    Program fused;
    EvalErr err;  // group-wide error (shape, device, budget)
  };
  std::vector<Group> groups;
  bool loaded = false;
  bool graph = true;  // device-driven rounds where a stratum allows them
  double exec_ms = 0;
  EvalErr err;
  // per-kernel records of the last execute, summed over the groups (when
  // the parent context's profiling is on)
  std::map<std::string, KRecord> kernels;
};

namespace {

// The fused program: slot `var_count` of every rule is the program id and
// becomes column 0 of every atom; arities grow by one.
Program fuse(const Program& p) {
  Program f = p;
  for (auto& d : f.preds) d.arity += 1;
  for (auto& r : f.rules) {
    u32 pid = r.var_count;
    r.var_count += 1;
    auto add = [&](Atom& a) { a.args.insert(a.args.begin(), Term{SK_T_VAR, pid, 0}); };
    add(r.head);
    for (auto& b : r.body) add(b);
  }
  return f;
}

// Runs f(group index) for every group, groups spread over up to 8 host
// threads (one group per thread at a time); each thread makes the device
// current first.
template <class F>
void for_groups(sk_batch* b, F f) {
  const size_t ng = b->groups.size();
  const int dev = b->ctx->c.device;
  const size_t nt = std::min<size_t>(ng, 8);
  if (nt <= 1) {
    for (size_t i = 0; i < ng; i++) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> ts;
  for (size_t t = 0; t < nt; t++)
    ts.emplace_back([&] {
      (void)hipSetDevice(dev);
      for (size_t i; (i = next.fetch_add(1)) < ng;) f(i);
    });
  for (auto& t : ts) t.join();
}

// Runs f(m) for m in [0, n) on up to 8 host threads.
template <class F>
void parallel_members(size_t n, F f) {
  const size_t nt = std::min<size_t>(n, 8);
  if (nt <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> ts;
  for (size_t t = 0; t < nt; t++)
    ts.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto& t : ts) t.join();
}

// Group-wide failure of a load or execute step.
template <class F>
void guarded_group(sk_batch::Group& g, F f) {
  try {
    f();
  } catch (const EvalErr& e) {
    g.err = e;
  } catch (const DeviceError& d) {
    g.err.code = SK_ERR_DEVICE;
    g.err.msg = d.where;
  } catch (const BudgetError&) {
    g.err.code = SK_ERR_OUT_OF_BUDGET;
  } catch (const std::exception& e) {
    g.err.code = SK_ERR_DEVICE;
    g.err.msg = e.what();
  }
}

}  // namespace

extern "C" {

sk_status sk_batch_create(sk_ctx* ctx, sk_batch** out) {
  if (!ctx || !out) return SK_ERR_INVALID_ARGUMENT;
  *out = new sk_batch();
  (*out)->ctx = ctx;
  return SK_OK;
}

sk_status sk_batch_set_graph(sk_batch* b, int enabled) {
  if (!b || b->loaded) return SK_ERR_INVALID_ARGUMENT;
  b->graph = enabled != 0;
  return SK_OK;
}

void sk_batch_free(sk_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->c.device);
  for (auto& g : b->groups) {
    g.eng.reset();
    if (g.gc) g.gc->release();
  }
  delete b;
}

sk_status sk_batch_add(sk_batch* b, const sk_program* prog, uint32_t n_edb, const char* const* edb_preds,
                       const uint64_t* edb_n, const uint32_t* const* const* edb_cols, const int64_t* const* edb_w,
                       uint32_t* out_index) {
  if (!b || !prog || b->loaded) return SK_ERR_INVALID_ARGUMENT;
  BProgram bp;
  bp.prog = Program::from_c(prog);
  for (uint32_t i = 0; i < n_edb; i++) {
    int r = bp.prog.rel_of(edb_preds[i]);
    if (r < 0) return SK_ERR_UNKNOWN_PRED;
    u32 ar = bp.prog.preds[r].arity;
    if (edb_n[i] && !edb_cols[i]) return SK_ERR_INVALID_ARGUMENT;
    // weights exactly for Trop predicates, as a plain run's staging requires
    // (Engine::stage_sym): one group shares one staged layout per relation
    const bool weighted = edb_w && edb_w[i];
    if (weighted != (bp.prog.preds[r].sem == SK_TROP)) return SK_ERR_SEMIRING_MISMATCH;
    bp.edb_preds.push_back(edb_preds[i]);
    bp.edb_n.push_back(edb_n[i]);
    std::vector<const u32*> cols(ar);
    for (u32 j = 0; j < ar; j++) {
      cols[j] = edb_n[i] ? edb_cols[i][j] : nullptr;
      if (edb_n[i] && !cols[j]) return SK_ERR_INVALID_ARGUMENT;
    }
    bp.edb_cols.push_back(std::move(cols));
    bp.edb_w.push_back(edb_w ? edb_w[i] : nullptr);
  }
  if (out_index) *out_index = (uint32_t)b->progs.size();
  b->progs.push_back(std::move(bp));
  return SK_OK;
}

sk_status sk_batch_load(sk_batch* b) {
  if (!b || b->loaded) return SK_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(b->ctx->c.device);
  std::map<std::string, int> gid;
  for (size_t i = 0; i < b->progs.size(); i++) {
    std::string sig = signature(b->progs[i].prog);
    auto it = gid.find(sig);
    if (it == gid.end()) {
      it = gid.emplace(sig, (int)b->groups.size()).first;
      b->groups.emplace_back();
    }
    b->progs[i].group = it->second;
    b->progs[i].member = (int)b->groups[it->second].members.size();
    b->groups[it->second].members.push_back((int)i);
  }
  // each group gets its own context, with an even share of the budget
  const u64 share = b->ctx->c.pool->budget() / std::max<size_t>(1, b->groups.size());
  for_groups(b, [&](size_t gi) {
    sk_batch::Group& g = b->groups[gi];
    const Program& p0 = b->progs[g.members[0]].prog;
    g.fused = fuse(p0);
    guarded_group(g, [&] {
      g.gc = std::make_unique<Ctx>();
      g.gc->init(b->ctx->c.device, share);
      g.eng = std::make_unique<Engine>(g.gc.get(), g.fused);
      g.eng->validate();
      g.eng->set_partitions((u32)g.members.size());
      g.eng->set_graph(b->graph);
      const size_t nm = g.members.size();
      // per program: its universe (distinct Sym values of its EDB, a bitmap
      // over its largest id) and its largest id, on up to 8 host threads
      std::vector<u32> pmax(nm, 0);
      parallel_members(nm, [&](size_t m) {
        BProgram& bp = b->progs[g.members[m]];
        u32 mx = 0;
        for (size_t e = 0; e < bp.edb_preds.size(); e++)
          for (const u32* c : bp.edb_cols[e])
            for (u64 i = 0; i < bp.edb_n[e]; i++) mx = std::max(mx, c[i]);
        std::vector<u64> seen(((u64)mx >> 6) + 1, 0);
        for (size_t e = 0; e < bp.edb_preds.size(); e++)
          for (const u32* c : bp.edb_cols[e])
            for (u64 i = 0; i < bp.edb_n[e]; i++) seen[c[i] >> 6] |= 1ull << (c[i] & 63);
        u64 distinct = 0;
        for (u64 wd : seen) distinct += (u64)__builtin_popcountll(wd);
        bp.universe = distinct;
        pmax[m] = mx;
      });
      u64 max_sym = nm ? nm - 1 : 0;  // program ids
      for (size_t m = 0; m < nm; m++) {
        g.eng->set_partition_universe((u32)m, b->progs[g.members[m]].universe);
        max_sym = std::max<u64>(max_sym, pmax[m]);
      }
      g.eng->set_max_sym_hint(max_sym);
      // stage every member's facts with its program id as column 0: sized
      // once per relation, each program filling its own row range
      std::map<std::string, u64> rows;
      std::map<std::string, bool> weighted;
      std::vector<std::vector<u64>> at(nm);
      for (size_t m = 0; m < nm; m++) {
        BProgram& bp = b->progs[g.members[m]];
        for (size_t e = 0; e < bp.edb_preds.size(); e++) {
          const std::string& pname = p0.preds[bp.prog.rel_of(bp.edb_preds[e])].name;
          at[m].push_back(rows[pname]);
          rows[pname] += bp.edb_n[e];
          weighted[pname] = bp.edb_w[e] != nullptr;
        }
      }
      // weights that do not fit a narrow cell (host threads), then every
      // relation straight to device columns, program id first
      std::vector<char> narrow(nm, 1);
      parallel_members(nm, [&](size_t m) {
        BProgram& bp = b->progs[g.members[m]];
        for (size_t e = 0; e < bp.edb_preds.size() && narrow[m]; e++)
          if (const i64* w = bp.edb_w[e])
            for (u64 i = 0; i < bp.edb_n[e]; i++)
              if ((u64)w[i] >= 0xFFFFFFFFull) {
                narrow[m] = 0;
                break;
              }
      });
      std::map<std::string, std::vector<Engine::DevSeg>> segs;
      for (size_t m = 0; m < nm; m++) {
        BProgram& bp = b->progs[g.members[m]];
        for (size_t e = 0; e < bp.edb_preds.size(); e++) {
          const std::string& pname = p0.preds[bp.prog.rel_of(bp.edb_preds[e])].name;
          segs[pname].push_back(Engine::DevSeg{(u32)m, at[m][e], bp.edb_n[e], bp.edb_cols[e].data(), bp.edb_w[e]});
        }
      }
      for (auto& kv : segs) g.eng->stage_dev(kv.first, rows[kv.first], weighted[kv.first], kv.second);
      bool all_narrow = true;
      for (char c : narrow) all_narrow = all_narrow && c;
      if (!all_narrow) g.eng->mark_not_narrow();
      for (int m : g.members) {
        BProgram& bp = b->progs[m];
        bp.edb_cols.clear();
        bp.edb_w.clear();
      }
      g.eng->load();
    });
  });
  b->loaded = true;
  return SK_OK;
}

sk_status sk_batch_execute(sk_batch* b) {
  if (!b || !b->loaded) return SK_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(b->ctx->c.device);
  const bool prof = b->ctx->c.timer.enabled;
  auto t0 = std::chrono::steady_clock::now();
  for_groups(b, [&](size_t gi) {
    sk_batch::Group& g = b->groups[gi];
    if (g.err.code || !g.eng) return;
    guarded_group(g, [&] {
      g.gc->timer.reset();
      g.gc->timer.enabled = prof;
      g.eng->execute();
      SK_HIP(hipStreamSynchronize(g.gc->stream));
    });
  });
  b->exec_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  b->kernels.clear();
  for (auto& g : b->groups) {
    if (!g.gc) continue;
    guarded_group(g, [&] { g.gc->timer.collect(); });
    g.gc->timer.enabled = false;
    for (auto& kv : g.gc->timer.kernels) {
      KRecord& r = b->kernels[kv.first];
      r.name = kv.second.name;
      r.cls = kv.second.cls;
      r.launches += kv.second.launches;
      r.ms += kv.second.ms;
      r.bytes += kv.second.bytes;
      r.items += kv.second.items;
    }
  }
  for (auto& bp : b->progs) {
    auto& g = b->groups[bp.group];
    if (g.err.code) {
      bp.err = g.err;
      continue;
    }
    bp.err = g.eng->partition_error((u32)bp.member);
    bp.iters.assign(bp.prog.num_strata, 0);
    for (u32 k = 0; k < bp.prog.num_strata; k++) bp.iters[k] = g.eng->partition_iterations((u32)bp.member, k);
  }
  return SK_OK;
}

sk_status sk_batch_status(const sk_batch* b, uint32_t index, sk_error_info* out) {
  if (!b || index >= b->progs.size()) return SK_ERR_INVALID_ARGUMENT;
  const EvalErr& e = b->progs[index].err;
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->code = e.code;
    out->slot = e.slot;
    out->lhs = e.lhs;
    out->rhs = e.rhs;
    std::snprintf(out->pred, sizeof out->pred, "%s", e.pred.c_str());
    std::snprintf(out->message, sizeof out->message, "%s", e.msg.c_str());
  }
  return (sk_status)e.code;
}

uint64_t sk_batch_iterations(const sk_batch* b, uint32_t index, uint32_t stratum) {
  if (!b || index >= b->progs.size()) return 0;
  const auto& it = b->progs[index].iters;
  return stratum < it.size() ? it[stratum] : 0;
}

int64_t sk_batch_result_size(sk_batch* b, uint32_t index, const char* pred) {
  if (!b || index >= b->progs.size() || !pred) return -1;
  BProgram& bp = b->progs[index];
  auto& g = b->groups[bp.group];
  if (bp.err.code || !g.eng) return -1;
  int r = bp.prog.rel_of(pred);
  if (r < 0) return -1;
  const std::string& fname = g.fused.preds[r].name;
  try {
    return (int64_t)g.eng->partition_size(fname, (u32)bp.member);
  } catch (const EvalErr& e) {
    bp.err = e;
  } catch (const DeviceError& d) {
    bp.err.code = SK_ERR_DEVICE;
    bp.err.msg = d.where;
  } catch (const BudgetError&) {
    bp.err.code = SK_ERR_OUT_OF_BUDGET;
    bp.err.msg = "device memory budget exceeded while reading the result";
  } catch (const std::exception& e) {
    bp.err.code = SK_ERR_DEVICE;
    bp.err.msg = e.what();
  }
  return -1;
}

sk_status sk_batch_result_read_sym(sk_batch* b, uint32_t index, const char* pred, uint32_t* const* cols,
                                   int64_t* w, uint8_t* inf) {
  if (!b || index >= b->progs.size() || !pred) return SK_ERR_INVALID_ARGUMENT;
  BProgram& bp = b->progs[index];
  auto& g = b->groups[bp.group];
  if (bp.err.code) return (sk_status)bp.err.code;
  int r = bp.prog.rel_of(pred);
  if (r < 0) return SK_ERR_UNKNOWN_PRED;
  (void)hipSetDevice(b->ctx->c.device);
  try {
    g.eng->partition_read_sym(g.fused.preds[r].name, (u32)bp.member, cols, w, inf);
  } catch (const EvalErr& e) {
    return (sk_status)e.code;
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  } catch (const std::exception&) {
    return SK_ERR_DEVICE;
  }
  return SK_OK;
}

double sk_batch_execute_ms(const sk_batch* b) { return b ? b->exec_ms : 0; }

uint32_t sk_batch_kernel_count(const sk_batch* b) { return b ? (uint32_t)b->kernels.size() : 0; }

sk_status sk_batch_kernel_record(const sk_batch* b, uint32_t i, char* name, size_t name_len, uint32_t* kernel_class,
                                 sk_kernel_stats* out) {
  if (!b || i >= b->kernels.size()) return SK_ERR_INVALID_ARGUMENT;
  auto it = b->kernels.begin();
  std::advance(it, i);
  const KRecord& r = it->second;
  if (name && name_len) std::snprintf(name, name_len, "%s", r.name.c_str());
  if (kernel_class) *kernel_class = (uint32_t)r.cls;
  if (out) {
    out->launches = r.launches;
    out->ms = r.ms;
    out->bytes = r.bytes;
    out->items = r.items;
  }
  return SK_OK;
}

}  // extern "C"
