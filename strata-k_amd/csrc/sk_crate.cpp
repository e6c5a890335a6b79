// sk_crate.cpp — the reference GPU crate's public functions
// (crates/strata-gpu/src/lib.rs:60-149) on top of the device engine.
//
// transitive_closure_bool / shortest_paths_trop run the closure program
// through the same semi-naive engine as sk_run_semi_naive (interpreter
// semantics: checked ⊗, negative-cycle cap universe+2).  count_triangles
// orients each undirected edge from lower to higher id, drops self-loops and
// duplicates (wcoj.rs:99-131) and runs the leapfrog operator; count_4cliques
// extends it by one more intersection level (wcoj.rs:49-77).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sk_engine.h"
#include "strata_k.h"

using namespace sk;

namespace sk {
u64 clique4_count(const Exec& x, const u32* a, const u32* b, u64 n, bool force_heavy, u32 tcap);
bool asp_simplify(const Exec& x, const int32_t* head, const u32* bs, const int32_t* lits, u64 n_rules, u64 n_lits,
                  const u8* poss, const u8* cert, u64 n_atoms, int32_t* out_head, u32* out_bs, int32_t* out_lits,
                  u64* out_rules, u64* out_lits_n, u8* scratch = nullptr);
void intern_lookup(const Exec& x, const u64* table, u64 nt, const u64* q, u64 nq, u32* ids);
}  // namespace sk

namespace {

// Closure program: path(X,Y) :- edge(X,Y).  path(X,Z) :- edge(X,Y), path(Y,Z).
struct TcProg {
  sk_pred preds[2];
  sk_term terms[10];
  sk_atom atoms[5];
  sk_rule rules[2];
  sk_program p;
  explicit TcProg(uint32_t sem) {
    preds[0] = {"edge", 2, sem, 0, 0};
    preds[1] = {"path", 2, sem, 0, 0};
    auto V = [](uint32_t s) { return sk_term{SK_T_VAR, s, 0}; };
    // rule 0: head path(0,1), body edge(0,1)
    terms[0] = V(0); terms[1] = V(1);
    terms[2] = V(0); terms[3] = V(1);
    // rule 1: head path(0,2), body edge(0,1), path(1,2)
    terms[4] = V(0); terms[5] = V(2);
    terms[6] = V(0); terms[7] = V(1);
    terms[8] = V(1); terms[9] = V(2);
    atoms[0] = {"path", 0, 0, 2, 0};
    atoms[1] = {"edge", 0, 2, 2, 0};
    atoms[2] = {"path", 0, 4, 2, 0};
    atoms[3] = {"edge", 0, 6, 2, 0};
    atoms[4] = {"path", 0, 8, 2, 0};
    rules[0] = {0, 1, 1, 0, 2, 0};
    rules[1] = {2, 3, 2, 0, 3, 0};
    p = {preds, 2, rules, 2, atoms, 5, terms, 10, 1};
  }
};

// An sk_free-able result buffer of n entries (at least one byte).
template <class T>
T* out_alloc(int64_t n) {
  return (T*)std::malloc(std::max<size_t>(1, (size_t)std::max<int64_t>(n, 0) * sizeof(T)));
}

}  // namespace

extern "C" {

sk_status sk_transitive_closure_bool(sk_ctx* ctx, const uint32_t* src, const uint32_t* dst, uint64_t n,
                                     uint32_t** out_src, uint32_t** out_dst, uint64_t* out_n) {
  if (!ctx || !out_src || !out_dst || !out_n || (n && (!src || !dst))) return SK_ERR_INVALID_ARGUMENT;
  TcProg tp(SK_BOOL);
  sk_run* run = nullptr;
  sk_status s = sk_prepare(ctx, &tp.p, &run);
  if (s) return s;
  const uint32_t* cols[2] = {src, dst};
  if (!(s = sk_run_add_facts_sym(run, "edge", n, 2, cols, nullptr, nullptr)) && !(s = sk_run_load(run)) &&
      !(s = sk_run_execute(run))) {
    // the result is read straight into the caller's buffers (no zeroed
    // staging vectors, no second copy)
    const int64_t m = sk_result_size(run, "path");
    uint32_t* a = out_alloc<uint32_t>(m);
    uint32_t* b = out_alloc<uint32_t>(m);
    uint32_t* oc[2] = {a, b};
    if (!a || !b) s = SK_ERR_OUT_OF_BUDGET;
    else s = sk_result_read_sym(run, "path", oc, nullptr, nullptr);
    if (!s) {
      *out_src = a;
      *out_dst = b;
      *out_n = (uint64_t)m;
    } else {
      std::free(a);
      std::free(b);
    }
  }
  sk_run_free(run);
  return s;
}

sk_status sk_shortest_paths_trop(sk_ctx* ctx, const uint32_t* src, const uint32_t* dst, const int64_t* w, uint64_t n,
                                 uint32_t** out_src, uint32_t** out_dst, int64_t** out_w, uint64_t* out_n) {
  if (!ctx || !out_src || !out_dst || !out_w || !out_n || (n && (!src || !dst || !w))) return SK_ERR_INVALID_ARGUMENT;
  TcProg tp(SK_TROP);
  sk_run* run = nullptr;
  sk_status s = sk_prepare(ctx, &tp.p, &run);
  if (s) return s;
  const uint32_t* cols[2] = {src, dst};
  int64_t dummy = 0;
  if (!(s = sk_run_add_facts_sym(run, "edge", n, 2, cols, n ? w : &dummy, nullptr)) && !(s = sk_run_load(run)) &&
      !(s = sk_run_execute(run))) {
    const int64_t m = sk_result_size(run, "path");
    uint32_t* a = out_alloc<uint32_t>(m);
    uint32_t* b = out_alloc<uint32_t>(m);
    int64_t* ww = out_alloc<int64_t>(m);
    uint32_t* oc[2] = {a, b};
    if (!a || !b || !ww) s = SK_ERR_OUT_OF_BUDGET;
    else s = sk_result_read_sym(run, "path", oc, ww, nullptr);
    if (!s) {
      *out_src = a;
      *out_dst = b;
      *out_w = ww;
      *out_n = (uint64_t)m;
    } else {
      std::free(a);
      std::free(b);
      std::free(ww);
    }
  }
  sk_run_free(run);
  return s;
}

sk_status sk_count_triangles(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t* out) {
  if (!ctx || !out || (n && (!a || !b))) return SK_ERR_INVALID_ARGUMENT;
  // tri(X,Y,Z) :- e(X,Y), e(Y,Z), e(X,Z).  ntri(count<X>) :- tri(X,_Y,_Z).
  // Orient u < v and drop self-loops here (wcoj.rs:99-131); the duplicates
  // the orientation creates are removed by the load's device sort.
  std::vector<uint32_t> u, v;
  u.reserve(n);
  v.reserve(n);
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t x = a[i], y = b[i];
    if (x == y) continue;
    u.push_back(x < y ? x : y);
    v.push_back(x < y ? y : x);
  }
  sk_pred preds[2] = {{"e", 2, SK_BOOL, 0, 0}, {"tri", 3, SK_BOOL, 0, 0}};
  auto V = [](uint32_t s) { return sk_term{SK_T_VAR, s, 0}; };
  sk_term terms[9] = {V(0), V(1), V(2), V(0), V(1), V(1), V(2), V(0), V(2)};
  sk_atom atoms[4] = {{"tri", 0, 0, 3, 0}, {"e", 0, 3, 2, 0}, {"e", 0, 5, 2, 0}, {"e", 0, 7, 2, 0}};
  sk_rule rules[1] = {{0, 1, 3, 0, 3, 0}};
  sk_program p = {preds, 2, rules, 1, atoms, 4, terms, 9, 1};
  sk_run* run = nullptr;
  sk_status s = sk_prepare(ctx, &p, &run);
  if (s) return s;
  const uint32_t* cols[2] = {u.data(), v.data()};
  if (!(s = sk_run_add_facts_sym(run, "e", u.size(), 2, cols, nullptr, nullptr)) && !(s = sk_run_load(run)) &&
      !(s = sk_run_execute(run)))
    *out = (uint64_t)sk_result_size(run, "tri");
  sk_run_free(run);
  return s;
}

sk_status sk_count_4cliques(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t* out) {
  if (!ctx || !out || (n && (!a || !b))) return SK_ERR_INVALID_ARGUMENT;
  *out = 0;
  if (n == 0) return SK_OK;
  try {
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    Exec x = c.exec();
    // the raw pairs go to the device as they are; simplification,
    // orientation and the CSR are built there
    DBuf<u32> da(c.pool.get(), n), db(c.pool.get(), n);
    c.h2d(da.get(), a, n * 4);
    c.h2d(db.get(), b, n * 4);
    *out = clique4_count(x, da.get(), db.get(), n, c.k4_force_heavy, c.k4_tcap);
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  } catch (const std::bad_alloc&) {
    return SK_ERR_OUT_OF_BUDGET;
  }
  return SK_OK;
}

sk_status sk_ctx_set_clique_test(sk_ctx* ctx, int force_heavy, uint32_t lds_cap) {
  if (!ctx || lds_cap > 1024) return SK_ERR_INVALID_ARGUMENT;
  ctx->c.k4_force_heavy = force_heavy != 0;
  ctx->c.k4_tcap = lds_cap ? lds_cap : 1024;
  return SK_OK;
}

sk_status sk_intern_terms(sk_ctx* ctx, const uint64_t* keys, uint64_t n, uint32_t* out_ids) {
  if (!ctx || (n && (!keys || !out_ids))) return SK_ERR_INVALID_ARGUMENT;
  if (n == 0) return SK_OK;
  try {
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    Exec x = c.exec();
    DBuf<u64> a(c.pool.get(), n), b(c.pool.get(), n), u(c.pool.get(), n), q(c.pool.get(), n);
    DBuf<u32> ids(c.pool.get(), n);
    SK_HIP(hipMemcpyAsync(a.get(), keys, n * 8, hipMemcpyHostToDevice, c.stream));
    SK_HIP(hipMemcpyAsync(q.get(), keys, n * 8, hipMemcpyHostToDevice, c.stream));
    u64* sk;
    radix_sort<u64>(x, a.get(), b.get(), nullptr, nullptr, nullptr, nullptr, n, 64, &sk, nullptr, nullptr);
    u64 m = unique_runs<u64>(x, sk, nullptr, nullptr, n, u.get(), nullptr, nullptr);
    intern_lookup(x, u.get(), m, q.get(), n, ids.get());
    SK_HIP(hipMemcpyAsync(out_ids, ids.get(), n * 4, hipMemcpyDeviceToHost, c.stream));
    SK_HIP(hipStreamSynchronize(c.stream));
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  }
  return SK_OK;
}

// ------------------------------------------------------- sort primitives
//
// The reference crate's device sort helpers over host slices
// (crates/strata-gpu/src/sort.rs:396-413, 518-600), on the LSD radix sort,
// run collapse (Bool: keep one; weighted: keep the minimum) and the
// context's radix autotuner.

}  // extern "C"

namespace {
template <class F>
sk_status sort_guard(sk_ctx* ctx, F f) {
  try {
    (void)hipSetDevice(ctx->c.device);
    f(ctx->c);
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  } catch (const std::bad_alloc&) {
    return SK_ERR_OUT_OF_BUDGET;
  }
  return SK_OK;
}

int bits_needed(const uint64_t* k, uint64_t n) {
  uint64_t o = 0;
  for (uint64_t i = 0; i < n; i++) o |= k[i];
  int b = 0;
  while (b < 64 && (o >> b)) b++;
  return b;
}

// sort (+ unique) of packed keys, weights optional; returns the row count
u64 sort_keys(Ctx& c, const uint64_t* keys, const int64_t* w, uint64_t n, bool unique, uint64_t* ok, int64_t* ow) {
  Exec x = c.exec();
  DBuf<u64> a(c.pool.get(), n), b(c.pool.get(), n), u(c.pool.get(), unique ? n : 1);
  DBuf<i64> wa(c.pool.get(), w ? n : 1), wb(c.pool.get(), w ? n : 1), wu(c.pool.get(), (w && unique) ? n : 1);
  SK_HIP(hipMemcpyAsync(a.get(), keys, n * 8, hipMemcpyHostToDevice, c.stream));
  if (w) SK_HIP(hipMemcpyAsync(wa.get(), w, n * 8, hipMemcpyHostToDevice, c.stream));
  u64* sk;
  i64* sw;
  u8* sf;
  radix_sort<u64>(x, a.get(), b.get(), w ? wa.get() : nullptr, w ? wb.get() : nullptr, nullptr, nullptr, n,
                  bits_needed(keys, n), &sk, &sw, &sf);
  u64 m = n;
  if (unique) {
    m = unique_runs<u64>(x, sk, w ? sw : nullptr, nullptr, n, u.get(), w ? wu.get() : nullptr, nullptr);
    sk = u.get();
    sw = wu.get();
  }
  SK_HIP(hipMemcpyAsync(ok, sk, m * 8, hipMemcpyDeviceToHost, c.stream));
  if (w && ow) SK_HIP(hipMemcpyAsync(ow, sw, m * 8, hipMemcpyDeviceToHost, c.stream));
  SK_HIP(hipStreamSynchronize(c.stream));
  return m;
}
}  // namespace

extern "C" {

sk_status sk_sort_u64(sk_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out) {
  if (!ctx || (n && (!keys || !out))) return SK_ERR_INVALID_ARGUMENT;
  if (n == 0) return SK_OK;
  return sort_guard(ctx, [&](Ctx& c) { sort_keys(c, keys, nullptr, n, false, out, nullptr); });
}

sk_status sk_sort_unique_u64(sk_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out, uint64_t* out_n) {
  if (!ctx || !out_n || (n && (!keys || !out))) return SK_ERR_INVALID_ARGUMENT;
  *out_n = 0;
  if (n == 0) return SK_OK;
  return sort_guard(ctx, [&](Ctx& c) { *out_n = sort_keys(c, keys, nullptr, n, true, out, nullptr); });
}

sk_status sk_sort_unique_pairs(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n, uint32_t* out_a,
                               uint32_t* out_b, uint64_t* out_n) {
  if (!ctx || !out_n || (n && (!a || !b || !out_a || !out_b))) return SK_ERR_INVALID_ARGUMENT;
  *out_n = 0;
  if (n == 0) return SK_OK;
  std::vector<uint64_t> k(n), o(n);
  for (uint64_t i = 0; i < n; i++) k[i] = ((uint64_t)a[i] << 32) | b[i];
  sk_status s = sort_guard(ctx, [&](Ctx& c) { *out_n = sort_keys(c, k.data(), nullptr, n, true, o.data(), nullptr); });
  for (uint64_t i = 0; i < *out_n; i++) {
    out_a[i] = (uint32_t)(o[i] >> 32);
    out_b[i] = (uint32_t)o[i];
  }
  return s;
}

sk_status sk_sort_min_pairs(sk_ctx* ctx, const uint32_t* a, const uint32_t* b, const int64_t* w, uint64_t n,
                            uint32_t* out_a, uint32_t* out_b, int64_t* out_w, uint64_t* out_n) {
  if (!ctx || !out_n || (n && (!a || !b || !w || !out_a || !out_b || !out_w))) return SK_ERR_INVALID_ARGUMENT;
  *out_n = 0;
  if (n == 0) return SK_OK;
  std::vector<uint64_t> k(n), o(n);
  for (uint64_t i = 0; i < n; i++) k[i] = ((uint64_t)a[i] << 32) | b[i];
  sk_status s = sort_guard(ctx, [&](Ctx& c) { *out_n = sort_keys(c, k.data(), w, n, true, o.data(), out_w); });
  for (uint64_t i = 0; i < *out_n; i++) {
    out_a[i] = (uint32_t)(o[i] >> 32);
    out_b[i] = (uint32_t)o[i];
  }
  return s;
}

sk_status sk_radix_sort_device(sk_ctx* ctx, uint64_t* keys, uint64_t* keys_alt, int64_t* w, int64_t* w_alt,
                               uint64_t n, int key_bits, int* result_in_alt) {
  if (!ctx || !result_in_alt || (n && (!keys || !keys_alt)) || (!w != !w_alt) || key_bits < 0 || key_bits > 64)
    return SK_ERR_INVALID_ARGUMENT;
  *result_in_alt = 0;
  return sort_guard(ctx, [&](Ctx& c) {
    Exec x = c.exec();
    u64* sk;
    i64* sw;
    u8* sf;
    radix_sort<u64>(x, keys, keys_alt, w, w_alt, nullptr, nullptr, n, key_bits, &sk, &sw, &sf);
    *result_in_alt = sk == keys_alt ? 1 : 0;
  });
}

sk_status sk_ctx_set_autotune(sk_ctx* ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return SK_ERR_INVALID_ARGUMENT;
  RadixTune& t = ctx->c.tune;
  t.forced = mode == 0 ? 0 : -1;
  t.lazy = mode != 0;
  if (mode == 2)
    return sort_guard(ctx, [&](Ctx& c) {
      Exec x = c.exec();
      radix_autotune(x, &t);
    });
  return SK_OK;
}

sk_status sk_ctx_force_radix(sk_ctx* ctx, int candidate) {
  if (!ctx || candidate < -1 || candidate >= radix_candidates()) return SK_ERR_INVALID_ARGUMENT;
  ctx->c.tune.forced = candidate;
  return SK_OK;
}

sk_status sk_ctx_radix_pick(sk_ctx* ctx, int wide, int weighted, uint64_t n, int key_bits, int* candidate,
                            double* est_ms) {
  if (!ctx || !candidate || key_bits <= 0 || key_bits > 128) return SK_ERR_INVALID_ARGUMENT;
  RadixTune& t = ctx->c.tune;
  int cls = 0;
  while (cls < RadixTune::kClasses - 1 && (n >> (RadixTune::kMinLog2 + 2 * cls + 1)) > 0) cls++;
  const int kw = wide ? 1 : 0, hw = weighted ? 1 : 0;
  int best = 0;
  double bt = -1;
  if (t.forced < 0 && t.tuned[kw][hw][cls] && n >= (1ull << RadixTune::kMinLog2)) {
    for (int c = 0; c < radix_candidates(); c++) {
      const int D = radix_candidate_digit(c);
      const double tt = ((key_bits + D - 1) / D) * (double)t.pass_ms[kw][hw][cls][c] *
                        ((double)n / (double)(1ull << (RadixTune::kMinLog2 + 2 * cls)));
      if (bt < 0 || tt < bt) {
        bt = tt;
        best = c;
      }
    }
  } else if (t.forced > 0) {
    best = t.forced;
  }
  *candidate = best;
  if (est_ms) *est_ms = bt;
  return SK_OK;
}

sk_status sk_ctx_radix_choice(sk_ctx* ctx, int wide, int weighted, uint64_t n, int key_bits, int* digit_bits,
                              int* items, double* est_ms) {
  if (!ctx || !digit_bits || !items) return SK_ERR_INVALID_ARGUMENT;
  int best = 0;
  const sk_status s = sk_ctx_radix_pick(ctx, wide, weighted, n, key_bits, &best, est_ms);
  if (s != SK_OK) return s;
  *digit_bits = radix_candidate_digit(best);
  *items = radix_candidate_items(best, wide != 0);
  return SK_OK;
}

int sk_radix_candidates(void) { return radix_candidates(); }

sk_status sk_ctx_radix_timing(sk_ctx* ctx, int wide, int weighted, int cls, float* ms, int n_ms) {
  if (!ctx || !ms || cls < 0 || cls >= RadixTune::kClasses || n_ms < 0) return SK_ERR_INVALID_ARGUMENT;
  const RadixTune& t = ctx->c.tune;
  const int kw = wide ? 1 : 0, hw = weighted ? 1 : 0;
  if (!t.tuned[kw][hw][cls]) return SK_ERR_INVALID_ARGUMENT;
  for (int c = 0; c < n_ms && c < radix_candidates(); c++) ms[c] = t.pass_ms[kw][hw][cls][c];
  return SK_OK;
}

sk_status sk_radix_candidate(int c, int wide, int* digit_bits, int* items, int* block, int* sweep) {
  if (c < 0 || c >= radix_candidates() || !digit_bits || !items || !block || !sweep) return SK_ERR_INVALID_ARGUMENT;
  *digit_bits = radix_candidate_digit(c);
  *items = radix_candidate_items(c, wide != 0);
  *block = radix_candidate_block(c);
  *sweep = radix_candidate_sweep(c) ? 1 : 0;
  return SK_OK;
}

sk_status sk_asp_simplify_filter(sk_ctx* ctx, const int32_t* head, const uint32_t* body_start,
                                 const int32_t* body_lit, uint64_t n_rules, const uint8_t* poss,
                                 const uint8_t* cert, uint64_t n_atoms, int32_t* out_head,
                                 uint32_t* out_body_start, int32_t* out_body_lit, uint64_t* out_n_rules,
                                 uint64_t* out_n_lits) {
  if (!ctx || !body_start || !out_body_start || !out_n_rules || !out_n_lits || (n_rules && (!head || !out_head)) ||
      (n_atoms && (!poss || !cert)))
    return SK_ERR_INVALID_ARGUMENT;
  const uint64_t n_lits = body_start[n_rules];
  if (n_lits && (!body_lit || !out_body_lit)) return SK_ERR_INVALID_ARGUMENT;
  bool ok = true;
  sk_status s = sort_guard(ctx, [&](Ctx& c) {
    Exec x = c.exec();
    Pool* p = c.pool.get();
    DBuf<int32_t> dh(p, n_rules ? n_rules : 1), dl(p, n_lits ? n_lits : 1), oh(p, n_rules ? n_rules : 1),
        ol(p, n_lits ? n_lits : 1);
    DBuf<u32> db(p, n_rules + 1), ob(p, n_rules + 1);
    DBuf<u8> dp(p, n_atoms ? n_atoms : 1), dc(p, n_atoms ? n_atoms : 1);
    if (n_rules) SK_HIP(hipMemcpyAsync(dh.get(), head, n_rules * 4, hipMemcpyHostToDevice, c.stream));
    SK_HIP(hipMemcpyAsync(db.get(), body_start, (n_rules + 1) * 4, hipMemcpyHostToDevice, c.stream));
    if (n_lits) SK_HIP(hipMemcpyAsync(dl.get(), body_lit, n_lits * 4, hipMemcpyHostToDevice, c.stream));
    if (n_atoms) {
      SK_HIP(hipMemcpyAsync(dp.get(), poss, n_atoms, hipMemcpyHostToDevice, c.stream));
      SK_HIP(hipMemcpyAsync(dc.get(), cert, n_atoms, hipMemcpyHostToDevice, c.stream));
    }
    u64 nr = 0, nl = 0;
    ok = asp_simplify(x, dh.get(), db.get(), dl.get(), n_rules, n_lits, dp.get(), dc.get(), n_atoms, oh.get(),
                      ob.get(), ol.get(), &nr, &nl);
    if (!ok) return;
    if (nr) SK_HIP(hipMemcpyAsync(out_head, oh.get(), nr * 4, hipMemcpyDeviceToHost, c.stream));
    SK_HIP(hipMemcpyAsync(out_body_start, ob.get(), (nr + 1) * 4, hipMemcpyDeviceToHost, c.stream));
    if (nl) SK_HIP(hipMemcpyAsync(out_body_lit, ol.get(), nl * 4, hipMemcpyDeviceToHost, c.stream));
    SK_HIP(hipStreamSynchronize(c.stream));
    *out_n_rules = nr;
    *out_n_lits = nl;
  });
  if (s) return s;
  return ok ? SK_OK : SK_ERR_INVALID_ARGUMENT;
}

sk_status sk_asp_simplify_filter_device(sk_ctx* ctx, const int32_t* head, const uint32_t* body_start,
                                        const int32_t* body_lit, uint64_t n_rules, uint64_t n_lits,
                                        const uint8_t* poss, const uint8_t* cert, uint64_t n_atoms,
                                        int32_t* out_head, uint32_t* out_body_start, int32_t* out_body_lit,
                                        uint64_t* out_n_rules, uint64_t* out_n_lits) {
  if (!ctx || !body_start || !out_body_start || !out_n_rules || !out_n_lits || (n_rules && (!head || !out_head)) ||
      (n_atoms && (!poss || !cert)) || (n_lits && (!body_lit || !out_body_lit)))
    return SK_ERR_INVALID_ARGUMENT;
  bool ok = true;
  sk_status s = sort_guard(ctx, [&](Ctx& c) {
    Exec x = c.exec();
    u64 nr = 0, nl = 0;
    ok = asp_simplify(x, head, body_start, body_lit, n_rules, n_lits, poss, cert, n_atoms, out_head, out_body_start,
                      out_body_lit, &nr, &nl);
    *out_n_rules = nr;
    *out_n_lits = nl;
  });
  if (s) return s;
  return ok ? SK_OK : SK_ERR_INVALID_ARGUMENT;
}

sk_status sk_asp_simplify_filter_device_noalloc(sk_ctx* ctx, const int32_t* head, const uint32_t* body_start,
                                                const int32_t* body_lit, uint64_t n_rules, uint64_t n_lits,
                                                const uint8_t* poss, const uint8_t* cert, uint64_t n_atoms,
                                                int32_t* out_head, uint32_t* out_body_start, int32_t* out_body_lit,
                                                uint64_t* out_n_rules, uint64_t* out_n_lits) {
  if (!ctx || !body_start || !out_body_start || !out_n_rules || !out_n_lits || (n_rules && (!head || !out_head)) ||
      (n_atoms && (!poss || !cert)) || (n_lits && (!body_lit || !out_body_lit)))
    return SK_ERR_INVALID_ARGUMENT;
  Ctx& c = ctx->c;
  if (n_rules > kAspMaxRules || n_atoms > kAspMaxAtoms) {
    c.last = EvalErr{};
    c.last.code = SK_ERR_INVALID_ARGUMENT;
    c.last.msg = "allocation-free ASP filter: at most 2^27 rules and 2^22 atoms (its static device scratch)";
    return SK_ERR_INVALID_ARGUMENT;
  }
  bool ok = true;
  try {
    (void)hipSetDevice(c.device);
    if (!c.asp_scratch) return SK_ERR_DEVICE;
    // no pool (nothing may be allocated) and no timer (no events created)
    Exec x{nullptr, c.stream, nullptr, c.derr, c.host_words, c.read_ev, nullptr};
    u64 nr = 0, nl = 0;
    ok = asp_simplify(x, head, body_start, body_lit, n_rules, n_lits, poss, cert, n_atoms, out_head, out_body_start,
                      out_body_lit, &nr, &nl, c.asp_scratch);
    *out_n_rules = nr;
    *out_n_lits = nl;
  } catch (const DeviceError&) {
    return SK_ERR_DEVICE;
  } catch (const BudgetError&) {
    return SK_ERR_OUT_OF_BUDGET;
  } catch (const std::exception&) {
    return SK_ERR_DEVICE;
  }
  return ok ? SK_OK : SK_ERR_INVALID_ARGUMENT;
}

uint64_t sk_binary_plan_intermediate(const uint32_t* a, const uint32_t* b, uint64_t n) {
  // the id-oriented simple graph: (min, max) pairs, self-loops dropped,
  // duplicates removed; then the sum over sources of out-degree squared
  std::vector<uint64_t> k;
  k.reserve(n);
  for (uint64_t i = 0; i < n; i++)
    if (a[i] != b[i]) k.push_back(((uint64_t)std::min(a[i], b[i]) << 32) | std::max(a[i], b[i]));
  std::sort(k.begin(), k.end());
  k.erase(std::unique(k.begin(), k.end()), k.end());
  uint64_t s = 0, run = 0;
  for (size_t i = 0; i < k.size(); i++) {
    if (i > 0 && (k[i] >> 32) != (k[i - 1] >> 32)) {
      s += run * run;
      run = 0;
    }
    run++;
  }
  s += run * run;
  return k.empty() ? 0 : s;
}

}  // extern "C"
