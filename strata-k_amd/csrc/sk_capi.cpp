// sk_capi.cpp — the extern "C" boundary (include/strata_k.h).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "sk_engine.h"
#include "strata_k.h"

using namespace sk;

struct sk_run {
  sk_ctx* ctx = nullptr;
  std::unique_ptr<Engine> eng;
  EvalErr err;
  KStat stats[KC_N];
  std::vector<KRecord> kernels;
};

namespace {

void set_info(const EvalErr& e, sk_error_info* out) {
  if (!out) return;
  std::memset(out, 0, sizeof *out);
  out->code = e.code;
  out->slot = e.slot;
  out->lhs = e.lhs;
  out->rhs = e.rhs;
  std::snprintf(out->pred, sizeof out->pred, "%s", e.pred.c_str());
  std::snprintf(out->message, sizeof out->message, "%s", e.msg.c_str());
}

EvalErr device_err(const DeviceError& d) {
  EvalErr e;
  e.code = SK_ERR_DEVICE;
  e.msg = std::string(hipGetErrorString(d.err)) + " in " + d.where;
  return e;
}

EvalErr budget_err(const BudgetError& b) {
  EvalErr e;
  e.code = SK_ERR_OUT_OF_BUDGET;
  e.msg = "device memory budget exceeded (request of " + std::to_string(b.want) + " bytes)";
  return e;
}

// Run f, converting every engine exception into an sk_status.
template <class F>
sk_status guarded(EvalErr* slot, F f) {
  try {
    f();
    return SK_OK;
  } catch (const EvalErr& e) {
    *slot = e;
  } catch (const DeviceError& d) {
    *slot = device_err(d);
  } catch (const BudgetError& b) {
    *slot = budget_err(b);
  } catch (const std::bad_alloc&) {
    EvalErr e;
    e.code = SK_ERR_OUT_OF_BUDGET;
    e.msg = "host memory exhausted";
    *slot = e;
  }
  return (sk_status)slot->code;
}

}  // namespace

extern "C" {

int sk_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n > 0;
}

const char* sk_build_info(void) { return "strata-k gfx950 engine (wave64, hand-written HIP kernels)"; }

sk_status sk_ctx_create(int device, uint64_t budget, sk_ctx** out) {
  if (!out) return SK_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!sk_device_available()) return SK_ERR_NOT_BUILT;
  auto* c = new sk_ctx();
  EvalErr e;
  sk_status s = guarded(&e, [&] {
    SK_HIP(hipSetDevice(device));
    size_t freeb = 0, total = 0;
    SK_HIP(hipMemGetInfo(&freeb, &total));
    u64 b = budget ? budget : (u64)(freeb * 0.9);
    c->c.init(device, b);
  });
  if (s != SK_OK) {
    c->c.release();
    delete c;
    return s;
  }
  *out = c;
  return SK_OK;
}

void sk_ctx_destroy(sk_ctx* ctx) {
  if (!ctx) return;
  ctx->c.release();
  delete ctx;
}

sk_status sk_ctx_synchronize(sk_ctx* ctx) {
  if (!ctx) return SK_ERR_INVALID_ARGUMENT;
  if (hipStreamSynchronize(ctx->c.stream) != hipSuccess) return SK_ERR_DEVICE;
  return ctx->c.fault_h && *ctx->c.fault_h ? SK_ERR_DEVICE : SK_OK;
}

void sk_ctx_error(const sk_ctx* ctx, sk_error_info* out) {
  if (ctx) set_info(ctx->c.last, out);
}

uint64_t sk_ctx_device_bytes(const sk_ctx* ctx) { return ctx ? ctx->c.pool->held() : 0; }
uint64_t sk_ctx_device_peak(const sk_ctx* ctx) { return ctx ? ctx->c.pool->peak() : 0; }

void sk_set_profiling(sk_ctx* ctx, int enabled) {
  if (!ctx) return;
  ctx->c.timer.collect();
  ctx->c.timer.enabled = enabled != 0;
}

sk_status sk_ctx_kernel_reset(sk_ctx* ctx) {
  if (!ctx) return SK_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->c.device);
  EvalErr tmp;
  sk_status s = guarded(&tmp, [&] { ctx->c.timer.collect(); });
  ctx->c.timer.reset();
  ctx->kernels.clear();
  return s;
}

uint32_t sk_ctx_kernel_count(sk_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->c.device);
  EvalErr tmp;
  guarded(&tmp, [&] { ctx->c.timer.collect(); });
  ctx->kernels.clear();
  for (auto& kv : ctx->c.timer.kernels) ctx->kernels.push_back(kv.second);
  return (uint32_t)ctx->kernels.size();
}

sk_status sk_ctx_kernel_record(const sk_ctx* ctx, uint32_t i, char* name, size_t name_len, uint32_t* cls,
                               sk_kernel_stats* out) {
  if (!ctx || i >= ctx->kernels.size()) return SK_ERR_INVALID_ARGUMENT;
  const KRecord& k = ctx->kernels[i];
  if (name && name_len) std::snprintf(name, name_len, "%s", k.name.c_str());
  if (cls) *cls = (uint32_t)k.cls;
  if (out) {
    out->launches = k.launches;
    out->ms = k.ms;
    out->bytes = k.bytes;
    out->items = k.items;
  }
  return SK_OK;
}

sk_status sk_prepare(sk_ctx* ctx, const sk_program* prog, sk_run** out) {
  if (!ctx || !prog || !out) return SK_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* r = new sk_run();
  r->ctx = ctx;
  sk_status s = guarded(&r->err, [&] {
    r->eng = std::make_unique<Engine>(&ctx->c, Program::from_c(prog));
    r->eng->validate();
  });
  if (s != SK_OK) {
    ctx->c.last = r->err;
    delete r;
    return s;
  }
  *out = r;
  return SK_OK;
}

sk_status sk_prepare_json(sk_ctx* ctx, const char* json, size_t len, sk_run** out) {
  if (!ctx || !json || !out) return SK_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  sk_program_json* pj = nullptr;
  char err[256];
  sk_status s = sk_program_from_json(json, len, &pj, err, sizeof err);
  if (s != SK_OK) {
    ctx->c.last = EvalErr{};
    ctx->c.last.code = s;
    ctx->c.last.msg = err;
    return s;
  }
  s = sk_prepare(ctx, sk_program_json_get(pj), out);  // the engine copies what it keeps
  sk_program_json_free(pj);
  return s;
}

void sk_run_free(sk_run* run) {
  if (!run) return;
  (void)hipSetDevice(run->ctx->c.device);
  (void)hipStreamSynchronize(run->ctx->c.stream);
  delete run;
}

sk_status sk_run_add_facts(sk_run* run, const char* pred, uint64_t n, const sk_val* rows, const int64_t* w,
                           const uint8_t* inf) {
  if (!run || !pred || (n && !rows)) return SK_ERR_INVALID_ARGUMENT;
  return guarded(&run->err, [&] { run->eng->stage_rows(pred, n, rows, w, inf); });
}

sk_status sk_run_add_facts_sym(sk_run* run, const char* pred, uint64_t n, uint32_t arity, const uint32_t* const* cols,
                               const int64_t* w, const uint8_t* inf) {
  if (!run || !pred || (n && arity && !cols)) return SK_ERR_INVALID_ARGUMENT;
  return guarded(&run->err, [&] { run->eng->stage_sym(pred, n, arity, cols, w, inf); });
}

sk_status sk_run_load(sk_run* run) {
  if (!run) return SK_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(run->ctx->c.device);
  sk_status s = guarded(&run->err, [&] { run->eng->load(); });
  if (s) run->ctx->c.last = run->err;
  return s;
}

sk_status sk_run_execute(sk_run* run) {
  if (!run) return SK_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(run->ctx->c.device);
  Timer& t = run->ctx->c.timer;
  t.reset();
  sk_status s = guarded(&run->err, [&] { run->eng->execute(); });
  EvalErr tmp;
  guarded(&tmp, [&] { t.collect(); });
  for (int i = 0; i < KC_N; i++) run->stats[i] = t.stats[i];
  run->kernels.clear();
  for (auto& kv : t.kernels) run->kernels.push_back(kv.second);
  if (s) run->ctx->c.last = run->err;
  return s;
}

sk_status sk_run_reset(sk_run* run) {
  if (!run) return SK_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(run->ctx->c.device);
  return guarded(&run->err, [&] { run->eng->reset(); });
}

sk_status sk_run_set_dense(sk_run* run, int enabled) {
  if (!run) return SK_ERR_INVALID_ARGUMENT;
  run->eng->set_dense(enabled != 0);
  return SK_OK;
}

sk_status sk_run_set_pull(sk_run* run, int mode) {
  if (!run || mode < 0 || mode > 2) return SK_ERR_INVALID_ARGUMENT;
  run->eng->set_pull(mode);
  return SK_OK;
}

sk_status sk_run_set_graph(sk_run* run, int enabled) {
  if (!run) return SK_ERR_INVALID_ARGUMENT;
  run->eng->set_graph(enabled != 0);
  return SK_OK;
}

sk_status sk_run_set_pack(sk_run* run, int enabled) {
  if (!run) return SK_ERR_INVALID_ARGUMENT;
  run->eng->set_pack(enabled != 0);
  return SK_OK;
}

void sk_run_error(const sk_run* run, sk_error_info* out) {
  if (run) set_info(run->err, out);
}

double sk_run_execute_ms(const sk_run* run) { return run && run->eng ? run->eng->exec_ms : 0; }

sk_status sk_run_stats(const sk_run* run, uint32_t cls, sk_kernel_stats* out) {
  if (!run || !out || cls > KC_N) return SK_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof *out);
  for (int i = 0; i < KC_N; i++) {
    if (cls != (uint32_t)KC_N && (int)cls != i) continue;
    out->launches += run->stats[i].launches;
    out->ms += run->stats[i].ms;
    out->bytes += run->stats[i].bytes;
    out->items += run->stats[i].items;
  }
  return SK_OK;
}

uint32_t sk_run_kernel_count(const sk_run* run) { return run ? (uint32_t)run->kernels.size() : 0; }

sk_status sk_run_kernel_record(const sk_run* run, uint32_t i, char* name, size_t name_len, uint32_t* cls,
                               sk_kernel_stats* out) {
  if (!run || i >= run->kernels.size()) return SK_ERR_INVALID_ARGUMENT;
  const KRecord& k = run->kernels[i];
  if (name && name_len) std::snprintf(name, name_len, "%s", k.name.c_str());
  if (cls) *cls = (uint32_t)k.cls;
  if (out) {
    out->launches = k.launches;
    out->ms = k.ms;
    out->bytes = k.bytes;
    out->items = k.items;
  }
  return SK_OK;
}

int64_t sk_result_size(sk_run* run, const char* pred) {
  if (!run || !pred) return -1;
  return run->eng->size(pred);
}

int64_t sk_result_arity(sk_run* run, const char* pred) {
  if (!run || !pred) return -1;
  int r = run->eng->program().rel_of(pred);
  return r < 0 ? -1 : (int64_t)run->eng->program().preds[r].arity;
}

sk_status sk_result_read(sk_run* run, const char* pred, sk_val* rows, int64_t* w, uint8_t* inf, uint8_t* is_w) {
  if (!run || !pred) return SK_ERR_INVALID_ARGUMENT;
  EvalErr e;
  return guarded(&e, [&] { run->eng->read(pred, rows, w, inf, is_w); });
}

sk_status sk_result_read_sym(sk_run* run, const char* pred, uint32_t* const* cols, int64_t* w, uint8_t* inf) {
  if (!run || !pred) return SK_ERR_INVALID_ARGUMENT;
  EvalErr e;
  return guarded(&e, [&] { run->eng->read_sym(pred, cols, w, inf); });
}

uint64_t sk_result_iterations(const sk_run* run, uint32_t stratum) {
  return run && run->eng ? run->eng->iterations(stratum) : 0;
}

uint64_t sk_result_universe(const sk_run* run) { return run && run->eng ? run->eng->universe() : 0; }

sk_status sk_run_semi_naive(sk_ctx* ctx, const sk_program* prog, uint64_t n_facts, const char* const* fact_preds,
                            const uint32_t* fact_arity, const sk_val* fact_rows, const int64_t* fact_w,
                            const uint8_t* fact_inf, const uint8_t* fact_is_w, sk_run** out) {
  sk_status s = sk_prepare(ctx, prog, out);
  if (s) return s;
  sk_run* run = *out;
  u64 off = 0;
  for (u64 i = 0; i < n_facts; i++) {
    bool isw = fact_is_w && fact_is_w[i];
    int64_t wv = fact_w ? fact_w[i] : 0;
    uint8_t fv = fact_inf ? fact_inf[i] : 0;
    s = sk_run_add_facts(run, fact_preds[i], 1, fact_rows + off, isw ? &wv : nullptr, isw ? &fv : nullptr);
    off += fact_arity[i];
    if (s) return s;
  }
  s = sk_run_load(run);
  if (s) return s;
  return sk_run_execute(run);
}

void sk_free(void* p) { std::free(p); }

}  // extern "C"
