// sk_join.hip — join-side kernels of the semi-naive engine on gfx950.
//
// A rule body is evaluated as a chain of steps over a binding table (one u32
// column per live variable, plus the Trop accumulator):
//   scan    relation (full or delta) -> table, with constant / repeated-var filters
//   join    table x relation index: per-row range count, device prefix sum,
//           load-balanced emit (outputs split evenly over blocks; each output
//           finds its row by binary search over the row offsets)
//   anti    negated literal: drop rows whose ground atom is present
//   project table -> packed head keys (+ weight)
//   leapfrog  sorted-list intersection for triangle bodies
// Trop ⊗ is checked i64 addition in body order with +inf absorbing; the
// first overflow is recorded in the device error record (crates/strata-ir/src/trop.rs:56-62).
#include <algorithm>
#include <climits>

#include "sk_join.h"

namespace sk {

__device__ __forceinline__ u32 lane_id_j() { return threadIdx.x & 63; }
__device__ __forceinline__ u32 wave_id_j() { return threadIdx.x >> 6; }

__device__ __forceinline__ u64 wave_incl_scan_j(u64 v) {
  u32 l = lane_id_j();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u64 y = __shfl_up(v, o, 64);
    if (l >= (u32)o) v += y;
  }
  return v;
}

// Inclusive prefix maximum of a non-negative value over the 64 lanes of a
// wave, by DPP lane moves (no LDS): shifts by 1, 2, 4 and 8 within each row
// of 16 lanes, then row 0's last lane broadcast into row 1 and row 2's into
// row 3, then lane 31 into rows 2 and 3.  Lanes whose source is outside
// their row read 0, the identity.
__device__ __forceinline__ u32 wave_incl_max_u32(u32 v) {
  v = max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false));  // row_shr:1
  v = max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false));  // row_shr:2
  v = max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false));  // row_shr:4
  v = max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false));  // row_shr:8
  v = max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15
  v = max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return v;
}

// Inclusive prefix sum of a u32 over the 64 lanes of a wave, by the same DPP
// lane moves as wave_incl_max_u32 (lanes whose source is outside their row
// read 0, the identity).
__device__ __forceinline__ u32 wave_incl_sum_u32(u32 v) {
  v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

static unsigned gridj(u64 n, u64 per = kBlock) {
  u64 g = (n + per - 1) / per;
  if (g == 0) g = 1;
  if (g > 0x7fffffffull) g = 0x7fffffffull;
  return (unsigned)g;
}

// Checked Trop product a ⊗ b: +inf absorbs; overflow is recorded (first one
// wins) and the result is meaningless (the round's result is discarded).
__device__ __forceinline__ void trop_mul(i64 a, u8 ai, i64 b, u8 bi, i64* r, u8* ri, DevErr* err) {
  if (ai | bi) {
    *r = 0;
    *ri = 1;
    return;
  }
  i64 s;
  if (__builtin_add_overflow(a, b, &s)) {
    if (atomicCAS(&err->code, 0, 2) == 0) {
      err->lhs = a;
      err->rhs = b;
    }
    s = 0;
  }
  *r = s;
  *ri = 0;
}

template <class K>
__device__ __forceinline__ u32 col_of(K key, int arity, int bits, int j) {
  K mask = (((K)1) << bits) - 1;
  return (u32)((key >> (bits * (arity - 1 - j))) & mask);
}

// ==================================================================== scan

// flags[i] = row i of the relation passes the atom's filters.
template <class K>
__global__ void k_scan_flags(const K* __restrict__ keys, u64 n, ScanDesc d, u8* __restrict__ flags) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = keys[i];
  bool ok = true;
  for (int q = 0; q < d.nconst; q++) ok &= col_of<K>(k, d.arity, d.bits, d.const_col[q]) == d.const_val[q];
  for (int q = 0; q < d.neq; q++)
    ok &= col_of<K>(k, d.arity, d.bits, d.eq_a[q]) == col_of<K>(k, d.arity, d.bits, d.eq_b[q]);
  flags[i] = ok;
}

// Write table rows: pos == nullptr means every row passes (identity positions).
template <class K>
__global__ void k_scan_write(const K* __restrict__ keys, const i64* __restrict__ w, const u8* __restrict__ f, u64 n,
                             ScanDesc d, const u8* __restrict__ flags, const u64* __restrict__ pos, Table out) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags && !flags[i]) return;
  u64 o = pos ? pos[i] : i;
  K k = keys[i];
  for (int c = 0; c < out.ncols; c++) out.col[c][o] = col_of<K>(k, d.arity, d.bits, d.out_col[c]);
  if (out.w) {
    // acc = otimes_id ⊗ w = 0 + w (cannot overflow)
    out.w[o] = w ? w[i] : 0;
    if (out.inf) out.inf[o] = f ? f[i] : (u8)0;
  }
}

template <class K>
u64 scan_rel(const Exec& x, const K* keys, const i64* w, const u8* f, u64 n, const ScanDesc& d, Table* out,
             TableAlloc alloc, void* actx) {
  bool filtered = d.nconst > 0 || d.neq > 0;
  if (!filtered) {
    alloc(actx, out, n);
    if (n) {
      Scope sc(x.timer, "scan_write", KC_JOIN, (double)n * (sizeof(K) + 4.0 * out->ncols), n);
      k_scan_write<K><<<gridj(n), kBlock, 0, x.stream>>>(keys, w, f, n, d, nullptr, nullptr, *out);
      SK_HIP(hipGetLastError());
    }
    return n;
  }
  if (n == 0) {
    alloc(actx, out, 0);
    return 0;
  }
  DBuf<u8> flags(x.pool, n);
  DBuf<u64> pos(x.pool, n);
  DBuf<u64> tot(x.pool, 1);
  {
    Scope sc(x.timer, "scan_flags", KC_JOIN, (double)n * (sizeof(K) + 1.0), n);
    k_scan_flags<K><<<gridj(n), kBlock, 0, x.stream>>>(keys, n, d, flags.get());
  }
  scan_u8(x, flags.get(), pos.get(), n, tot.get());
  u64 m = read_u64(x, tot.get());
  alloc(actx, out, m);
  if (m) {
    Scope sc(x.timer, "scan_write", KC_JOIN, (double)n * (sizeof(K) + 1.0) + m * 4.0 * out->ncols, n);
    k_scan_write<K><<<gridj(n), kBlock, 0, x.stream>>>(keys, w, f, n, d, flags.get(), pos.get(), *out);
    SK_HIP(hipGetLastError());
  }
  return m;
}

// ==================================================================== join

// Probe prefix of index order for table row r, packed as the high bits of
// the index key.
template <class K>
__device__ __forceinline__ K probe_prefix(const Table& in, const ProbeDesc& p, u64 r) {
  K pre = 0;
  for (int j = 0; j < p.nprobe; j++) {
    u32 v = p.psrc[j] >= 0 ? in.col[p.psrc[j]][r] : p.pconst[j];
    pre = (pre << p.bits) | (K)v;
  }
  return pre;
}

template <class K>
__device__ __forceinline__ bool filt_ok(K key, const ProbeDesc& p) {
  bool ok = true;
  for (int q = 0; q < p.nfilt; q++)
    ok &= col_of<K>(key, p.arity, p.bits, p.fa[q]) == col_of<K>(key, p.arity, p.bits, p.fb[q]);
  return ok;
}

template <class K>
__device__ __forceinline__ void join_count_row(const Table& in, const ProbeDesc& p, const K* __restrict__ skeys,
                                               const u64* __restrict__ csr, u64 r, u64* __restrict__ lo_out,
                                               u32* __restrict__ cnt) {
  u64 lo, hi;
  if (p.nprobe == 0) {
    lo = 0;
    hi = p.sn;
  } else if (csr) {
    u32 v = p.psrc[0] >= 0 ? in.col[p.psrc[0]][r] : p.pconst[0];
    lo = csr[v];
    hi = csr[v + 1];
  } else {
    K pre = probe_prefix<K>(in, p, r);
    int rem = p.bits * (p.arity - p.nprobe);
    K klo = pre << rem;
    K khi = rem ? (klo | ((((K)1) << rem) - 1)) : klo;
    u64 a = 0, b = p.sn;
    while (a < b) {
      u64 m = (a + b) >> 1;
      if (skeys[m] < klo) a = m + 1;
      else b = m;
    }
    lo = a;
    b = p.sn;
    while (a < b) {
      u64 m = (a + b) >> 1;
      if (skeys[m] <= khi) a = m + 1;
      else b = m;
    }
    hi = a;
  }
  u64 c = hi - lo;
  if (p.nfilt) {
    c = 0;
    for (u64 s = lo; s < hi; s++) c += filt_ok<K>(skeys[s], p);
  }
  lo_out[r] = lo;
  cnt[r] = (u32)c;
}

template <class K>
__global__ void k_join_count(Table in, ProbeDesc p, const K* __restrict__ skeys, const u64* __restrict__ csr,
                             u64* __restrict__ lo_out, u32* __restrict__ cnt) {
  u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= in.n) return;
  join_count_row<K>(in, p, skeys, csr, r, lo_out, cnt);
}

// Output value for output column c of emitted row (table row r, relation row key).
template <class K>
__device__ __forceinline__ u32 out_val(const OutDesc& o, int c, const Table& in, u64 r, K skey, const ProbeDesc& p) {
  int kind = o.kind[c];
  if (kind == 0) return in.col[o.idx[c]][r];
  if (kind == 1) return col_of<K>(skey, p.arity, p.bits, o.idx[c]);
  return o.cnst[c];
}

template <class K, class KO, bool FINAL, bool TROP>
__device__ __forceinline__ void emit_one(const Table& in, const ProbeDesc& p, const OutDesc& o, u64 r, u64 s,
                                         const K* __restrict__ skeys, const i64* __restrict__ sw,
                                         const u8* __restrict__ sf, u64 oi, DevErr* err) {
  K sk = skeys[s];
  if constexpr (FINAL) {
    if (o.cd.on) {
      u64 c;
      if (!cell_of(o.cd, [&](int j) { return out_val<K>(o, j, in, r, sk, p); }, &c)) return;
      i64 res = 0;
      u8 rf = 0;
      if constexpr (TROP) {
        u8 af = in.inf ? in.inf[r] : (u8)0;
        u8 bf = sf ? sf[s] : (u8)0;
        DevErr* e = in.part_col >= 0 ? &in.perr[in.col[in.part_col][r]] : err;
        trop_mul(in.w[r], af, sw[s], bf, &res, &rf, e);
      }
      cell_insert(o.cd, c, res, rf);
      return;
    }
    KO key = 0;
    for (int c = 0; c < o.ncols; c++) key = (key << o.bits) | (KO)out_val<K>(o, c, in, r, sk, p);
    ((KO*)o.keys)[oi] = key;
  } else {
    for (int c = 0; c < o.ncols; c++) o.col[c][oi] = out_val<K>(o, c, in, r, sk, p);
  }
  if constexpr (TROP) {
    i64 res;
    u8 rf;
    u8 af = in.inf ? in.inf[r] : (u8)0;
    u8 bf = sf ? sf[s] : (u8)0;
    DevErr* e = in.part_col >= 0 ? &in.perr[in.col[in.part_col][r]] : err;
    trop_mul(in.w[r], af, sw[s], bf, &res, &rf, e);
    o.w[oi] = res;
    if (o.inf) o.inf[oi] = rf;
  }
}

constexpr int kEmitItems = 8;
constexpr u64 kEmitTile = (u64)kBlock * kEmitItems;
constexpr int kEmitLds = 2048;

// Load-balanced emit: block b owns outputs [b*tile, (b+1)*tile).  The rows
// spanning the tile are found by binary search over the exclusive offsets;
// when they fit, their offsets are staged in LDS and each output searches
// there.  Output writes and relation reads are coalesced along the output.
template <class K, class KO, bool FINAL, bool TROP>
__global__ __launch_bounds__(kBlock) void k_join_emit_lb(Table in, ProbeDesc p, OutDesc o, const K* __restrict__ skeys,
                                                         const i64* __restrict__ sw, const u8* __restrict__ sf,
                                                         const u64* __restrict__ lo, const u64* __restrict__ off,
                                                         u64 total, DevErr* err) {
  __shared__ u64 s_off[kEmitLds];
  __shared__ u64 s_r[2];
  const u64 o0 = (u64)blockIdx.x * kEmitTile;
  const u64 o1 = min(o0 + kEmitTile, total);
  const u64 nrows = in.n;
  if (threadIdx.x < 2) {
    // row containing output q: last r with off[r] <= q
    u64 q = threadIdx.x == 0 ? o0 : o1 - 1;
    u64 a = 0, b = nrows;
    while (a < b) {
      u64 m = (a + b) >> 1;
      if (off[m] <= q) a = m + 1;
      else b = m;
    }
    s_r[threadIdx.x] = a - 1;
  }
  __syncthreads();
  const u64 r0 = s_r[0], r1 = s_r[1];
  const u64 span = r1 - r0 + 1;
  const bool staged = span <= (u64)kEmitLds;
  if (staged)
    for (u64 q = threadIdx.x; q < span; q += kBlock) s_off[q] = off[r0 + q];
  __syncthreads();
  for (int k = 0; k < kEmitItems; k++) {
    u64 oi = o0 + (u64)k * kBlock + threadIdx.x;
    if (oi >= o1) break;
    u64 a, b;
    u64 r;
    if (staged) {
      a = 0;
      b = span;
      while (a < b) {
        u64 m = (a + b) >> 1;
        if (s_off[m] <= oi) a = m + 1;
        else b = m;
      }
      r = r0 + a - 1;
    } else {
      a = r0;
      b = r1 + 1;
      while (a < b) {
        u64 m = (a + b) >> 1;
        if (off[m] <= oi) a = m + 1;
        else b = m;
      }
      r = a - 1;
    }
    u64 s = lo[r] + (oi - (staged ? s_off[r - r0] : off[r]));
    emit_one<K, KO, FINAL, TROP>(in, p, o, r, s, skeys, sw, sf, oi, err);
  }
}

// Filtered emit (repeated unbound variables inside the probed atom): one
// thread per table row walks its range.
template <class K, class KO, bool FINAL, bool TROP>
__global__ void k_join_emit_rows(Table in, ProbeDesc p, OutDesc o, const K* __restrict__ skeys,
                                 const i64* __restrict__ sw, const u8* __restrict__ sf, const u64* __restrict__ lo,
                                 const u32* __restrict__ cnt, const u64* __restrict__ off, DevErr* err) {
  u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= in.n || cnt[r] == 0) return;
  u64 oi = off[r];
  u64 s = lo[r];
  for (u32 k = 0; k < cnt[r]; s++) {
    if (!filt_ok<K>(skeys[s], p)) continue;
    emit_one<K, KO, FINAL, TROP>(in, p, o, r, s, skeys, sw, sf, oi, err);
    oi++;
    k++;
  }
}

// Dense-head emit (the head rows are ⊕-applied to cells): the same output
// tiling as k_join_emit_lb (output k*kBlock + t of a tile goes to thread t,
// so relation reads stay coalesced), but each thread runs its kEmitItems
// outputs in phases — all row searches interleaved, then all relation and
// table reads, then all cell reads, then the applications — so that every
// lane has kEmitItems independent memory requests in flight instead of one
// dependent chain per output.
// PDOM: some head column is ranked within its program (batched runs); the
// lookup is compiled only into that variant, which keeps the common one at
// its register budget.
template <class K, bool TROP, bool PDOM>
__device__ __forceinline__ void join_apply_tile(const Table& in, const ProbeDesc& p, const OutDesc& o,
                                                const K* __restrict__ skeys, const i64* __restrict__ sw,
                                                const u8* __restrict__ sf, const u64* __restrict__ lo,
                                                const u64* __restrict__ off, const u64 o0, const u64 o1,
                                                const u64 nrows, DevErr* err) {
  __shared__ u64 s_off[kEmitLds];
  __shared__ u64 s_r[2];
  if (threadIdx.x < 2) {
    u64 q = threadIdx.x == 0 ? o0 : o1 - 1;
    u64 a = 0, b = nrows;
    while (a < b) {
      u64 m = (a + b) >> 1;
      if (off[m] <= q) a = m + 1;
      else b = m;
    }
    s_r[threadIdx.x] = a - 1;
  }
  __syncthreads();
  const u64 r0 = s_r[0], r1 = s_r[1];
  const u64 span = r1 - r0 + 1;
  const bool staged = span <= (u64)kEmitLds;
  if (staged)
    for (u64 q = threadIdx.x; q < span; q += kBlock) s_off[q] = off[r0 + q];
  __syncthreads();

  // phase 1: the row of every output — interleaved, branch-free binary
  // searches (one LDS read per output per step, all issued together)
  u64 q[kEmitItems], row[kEmitItems], rbeg[kEmitItems];
  bool ok[kEmitItems];
#pragma unroll
  for (int k = 0; k < kEmitItems; k++) {
    q[k] = o0 + (u64)k * kBlock + threadIdx.x;
    ok[k] = q[k] < o1;
    q[k] = min(q[k], o1 - 1);
  }
  if (staged) {
    u32 a[kEmitItems], b[kEmitItems];
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) {
      a[k] = 0;
      b[k] = (u32)span;
    }
    for (u64 wdt = span; wdt; wdt >>= 1) {
      u64 v[kEmitItems];
      u32 m[kEmitItems];
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) {
        m[k] = min((a[k] + b[k]) >> 1, (u32)span - 1);
        v[k] = s_off[m[k]];
      }
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) {
        const bool live = a[k] < b[k];
        const bool right = v[k] <= q[k];
        a[k] = (live && right) ? m[k] + 1 : a[k];
        b[k] = (live && !right) ? m[k] : b[k];
      }
    }
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) {
      const u32 ai = a[k] - 1;  // s_off[0] <= o0 <= q[k], so a[k] >= 1
      row[k] = r0 + ai;
      rbeg[k] = s_off[ai];
    }
  } else {
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) {
      u64 lo2 = r0, hi2 = r1 + 1;
      while (lo2 < hi2) {
        u64 m = (lo2 + hi2) >> 1;
        if (off[m] <= q[k]) lo2 = m + 1;
        else hi2 = m;
      }
      row[k] = lo2 - 1;
      rbeg[k] = off[lo2 - 1];
    }
  }

  // phase 2: relation records, then head columns one at a time across all
  // outputs (uniform branches outside, independent reads inside)
  u64 sidx[kEmitItems];
  K key[kEmitItems];
#pragma unroll
  for (int k = 0; k < kEmitItems; k++) sidx[k] = lo[row[k]] + (q[k] - rbeg[k]);
#pragma unroll
  for (int k = 0; k < kEmitItems; k++) key[k] = skeys[sidx[k]];
  u64 cell[kEmitItems];
#pragma unroll
  for (int k = 0; k < kEmitItems; k++) cell[k] = 0;
  u32 pidc[PDOM ? kEmitItems : 1];  // column 0's codes (program ids in batched runs)
  if constexpr (PDOM) {
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) pidc[k] = 0;
  }
  for (int j = 0; j < o.ncols; j++) {
    u32 code[kEmitItems];
    const int kind = o.kind[j];
    if (kind == 0) {
      const u32* col = in.col[o.idx[j]];
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) code[k] = col[row[k]];
    } else if (kind == 1) {
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) code[k] = col_of<K>(key[k], p.arity, p.bits, o.idx[j]);
    } else {
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) code[k] = o.cnst[j];
    }
    const u32* db = o.cd.dom_bits[j];
    const int sh = o.cd.shift[j], eb = o.cd.ebits[j];
    bool pdom_col = false;
    if constexpr (PDOM) {
      if (j == 0) {
#pragma unroll
        for (int k = 0; k < kEmitItems; k++) pidc[k] = code[k];
      }
      if (o.cd.pdom_off[j]) {
        pdom_col = true;
#pragma unroll
        for (int k = 0; k < kEmitItems; k++) {
          u64 rk = 0;
          const bool in_dom = cell_rank(o.cd, j, code[k], pidc[k], &rk) && !(rk >> eb);
          ok[k] = ok[k] && in_dom;
          cell[k] |= rk << sh;
        }
      }
    }
    if (pdom_col) {
    } else if (!db) {
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) {
        const bool in_dom = code[k] < o.cd.ncodes && !((u64)code[k] >> eb);
        ok[k] = ok[k] && in_dom;
        cell[k] |= (u64)code[k] << sh;
      }
    } else {
      const u64* dp = o.cd.dom_pre[j];
      u32 wbits[kEmitItems];
      u64 pre[kEmitItems];
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) {
        const u32 wi = min(code[k], (u32)(o.cd.ncodes - 1)) >> 5;
        wbits[k] = db[wi];
        pre[k] = dp[wi];
      }
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) {
        const u32 m = 1u << (code[k] & 31);
        const u64 rk = pre[k] + (u64)__popc(wbits[k] & (m - 1));
        const bool in_dom = code[k] < o.cd.ncodes && (wbits[k] & m) && !(rk >> eb);
        ok[k] = ok[k] && in_dom;
        cell[k] |= rk << sh;
      }
    }
  }
  // a valid output whose value lies outside the domain is a planning bug
  {
    bool badv = false;
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) badv |= (q[k] == o0 + (u64)k * kBlock + threadIdx.x) && !ok[k] &&
                                                 (o0 + (u64)k * kBlock + threadIdx.x < o1);
    if (badv) atomicOr(o.cd.bad, 1u);
  }
  i64 res[kEmitItems];
  u8 rf[kEmitItems];
#pragma unroll
  for (int k = 0; k < kEmitItems; k++) {
    res[k] = 0;
    rf[k] = 0;
  }
  if constexpr (TROP) {
    i64 aw[kEmitItems], bw[kEmitItems];
    u8 af[kEmitItems], bf[kEmitItems];
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) {
      aw[k] = in.w[row[k]];
      bw[k] = sw[sidx[k]];
      af[k] = 0;
      bf[k] = 0;
    }
    if (in.inf) {
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) af[k] = in.inf[row[k]];
    }
    if (sf) {
#pragma unroll
      for (int k = 0; k < kEmitItems; k++) bf[k] = sf[sidx[k]];
    }
#pragma unroll
    for (int k = 0; k < kEmitItems; k++) {
      if (!ok[k]) continue;
      DevErr* e = in.part_col >= 0 ? &in.perr[in.col[in.part_col][row[k]]] : err;
      trop_mul(aw[k], af[k], bw[k], bf[k], &res[k], &rf[k], e);
    }
  }
  // phase 3: every cell read before any application (unconditional: an
  // output that is not applied reads cell 0 and ignores it)
  u64 seen[kEmitItems];
#pragma unroll
  for (int k = 0; k < kEmitItems; k++) {
    const u64 c = ok[k] ? cell[k] : 0;
    if constexpr (TROP) seen[k] = o.cd.narrow ? (u64)o.cd.cell32[c] : o.cd.cell[c];
    else seen[k] = o.cd.present[c >> 5];
  }
  // phase 4: apply
#pragma unroll
  for (int k = 0; k < kEmitItems; k++)
    if (ok[k]) cell_apply_t<TROP>(o.cd, cell[k], res[k], rf[k], seen[k]);
}

template <class K, bool TROP, bool PDOM>
__global__ __launch_bounds__(kBlock) void k_join_apply(Table in, ProbeDesc p, OutDesc o, const K* __restrict__ skeys,
                                                       const i64* __restrict__ sw, const u8* __restrict__ sf,
                                                       const u64* __restrict__ lo, const u64* __restrict__ off,
                                                       u64 total, DevErr* err) {
  const u64 o0 = (u64)blockIdx.x * kEmitTile;
  join_apply_tile<K, TROP, PDOM>(in, p, o, skeys, sw, sf, lo, off, o0, min(o0 + kEmitTile, total), in.n, err);
}

template <class K, class KO, bool FINAL, bool TROP>
static void emit_launch(const Exec& x, const Table& in, const ProbeDesc& p, const OutDesc& o, const K* skeys,
                        const i64* sw, const u8* sf, const u64* lo, const u32* cnt, const u64* off, u64 total) {
  if (FINAL && o.cd.on && p.nfilt == 0) {
    u64 nb = (total + kEmitTile - 1) / kEmitTile;
    bool pdom = false;
    for (int j = 0; j < o.cd.ncols; j++) pdom = pdom || o.cd.pdom_off[j] != nullptr;
    if (pdom)
      k_join_apply<K, TROP, true><<<(unsigned)nb, kBlock, 0, x.stream>>>(in, p, o, skeys, sw, sf, lo, off, total,
                                                                         x.err);
    else
      k_join_apply<K, TROP, false><<<(unsigned)nb, kBlock, 0, x.stream>>>(in, p, o, skeys, sw, sf, lo, off, total,
                                                                          x.err);
  } else if (p.nfilt == 0) {
    u64 nb = (total + kEmitTile - 1) / kEmitTile;
    k_join_emit_lb<K, KO, FINAL, TROP><<<(unsigned)nb, kBlock, 0, x.stream>>>(in, p, o, skeys, sw, sf, lo, off, total,
                                                                               x.err);
  } else {
    k_join_emit_rows<K, KO, FINAL, TROP><<<gridj(in.n), kBlock, 0, x.stream>>>(in, p, o, skeys, sw, sf, lo, cnt, off,
                                                                                x.err);
  }
  SK_HIP(hipGetLastError());
}

template <class K>
u64 join(const Exec& x, const Table& in, const ProbeDesc& p, const K* skeys, const i64* sw, const u8* sf,
         const u64* csr, OutDesc* o, bool final_step, bool k128_out, bool trop, TableAlloc alloc, void* actx,
         HeadAlloc halloc, void* hctx) {
  if (in.n == 0) {
    if (final_step) halloc(hctx, o, 0);
    else alloc(actx, o->table, 0);
    return 0;
  }
  DBuf<u64> lo(x.pool, in.n);
  DBuf<u32> cnt(x.pool, in.n);
  DBuf<u64> off(x.pool, in.n);
  DBuf<u64> tot(x.pool, 1);
  {
    Scope sc(x.timer, "join_count", KC_JOIN, (double)in.n * (4.0 * p.nprobe + 12.0), in.n);
    k_join_count<K><<<gridj(in.n), kBlock, 0, x.stream>>>(in, p, skeys, csr, lo.get(), cnt.get());
    SK_HIP(hipGetLastError());
  }
  scan_u32(x, cnt.get(), off.get(), in.n, tot.get());
  u64 total = read_u64(x, tot.get());
  if (final_step) halloc(hctx, o, total);
  else {
    alloc(actx, o->table, total);
    for (int c = 0; c < o->ncols; c++) o->col[c] = o->table->col[c];
    o->w = o->table->w;
    o->inf = o->table->inf;
  }
  if (total == 0) return 0;
  // algorithmic bytes: every emitted row reads one relation record and
  // writes one output record
  double rec_in = sizeof(K) + (trop ? 8.0 + (sf ? 1.0 : 0) : 0);
  // (a head applied to cells reads one cell word per row instead of writing a record)
  double rec_out = (final_step && o->cd.on)
                       ? (trop ? (o->cd.narrow ? 4.0 : 8.0) : 4.0)
                       : (final_step ? (k128_out ? 16.0 : 8.0) : 4.0 * o->ncols) + (trop ? 8.0 + (o->inf ? 1.0 : 0) : 0);
  Scope sc(x.timer, "join_emit", KC_JOIN, (double)total * (rec_in + rec_out), total);
  if (final_step) {
    if (k128_out) {
      if (trop) emit_launch<K, u128, true, true>(x, in, p, *o, skeys, sw, sf, lo.get(), cnt.get(), off.get(), total);
      else emit_launch<K, u128, true, false>(x, in, p, *o, skeys, sw, sf, lo.get(), cnt.get(), off.get(), total);
    } else {
      if (trop) emit_launch<K, u64, true, true>(x, in, p, *o, skeys, sw, sf, lo.get(), cnt.get(), off.get(), total);
      else emit_launch<K, u64, true, false>(x, in, p, *o, skeys, sw, sf, lo.get(), cnt.get(), off.get(), total);
    }
  } else {
    if (trop) emit_launch<K, u64, false, true>(x, in, p, *o, skeys, sw, sf, lo.get(), cnt.get(), off.get(), total);
    else emit_launch<K, u64, false, false>(x, in, p, *o, skeys, sw, sf, lo.get(), cnt.get(), off.get(), total);
  }
  return total;
}

// ============================================================== anti-join

template <class K>
__global__ void k_anti_flags(Table in, AntiDesc a, const K* __restrict__ skeys, u64 sn, u8* __restrict__ keep) {
  u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= in.n) return;
  K key = 0;
  for (int j = 0; j < a.arity; j++) {
    u32 v = a.src[j] >= 0 ? in.col[a.src[j]][r] : a.cnst[j];
    key = (key << a.bits) | (K)v;
  }
  u64 lo = 0, hi = sn;
  while (lo < hi) {
    u64 m = (lo + hi) >> 1;
    if (skeys[m] < key) lo = m + 1;
    else hi = m;
  }
  bool present = lo < sn && skeys[lo] == key;
  keep[r] = present ? 0 : 1;
}

__global__ void k_compact_table(Table in, const u8* __restrict__ keep, const u64* __restrict__ pos, Table out) {
  u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= in.n || !keep[r]) return;
  u64 o = pos[r];
  for (int c = 0; c < out.ncols; c++) out.col[c][o] = in.col[out.srccol[c]][r];
  if (out.w) {
    out.w[o] = in.w[r];
    if (out.inf) out.inf[o] = in.inf ? in.inf[r] : (u8)0;
  }
}

template <class K>
u64 anti_join(const Exec& x, const Table& in, const AntiDesc& a, const K* skeys, u64 sn, Table* out,
              TableAlloc alloc, void* actx) {
  if (in.n == 0) {
    alloc(actx, out, 0);
    return 0;
  }
  DBuf<u8> keep(x.pool, in.n);
  DBuf<u64> pos(x.pool, in.n);
  DBuf<u64> tot(x.pool, 1);
  if (a.never_present) {
    fill_u8(x, keep.get(), 1, in.n);
  } else {
    Scope sc(x.timer, "anti_flags", KC_JOIN, (double)in.n * (4.0 * a.arity + 1.0), in.n);
    k_anti_flags<K><<<gridj(in.n), kBlock, 0, x.stream>>>(in, a, skeys, sn, keep.get());
    SK_HIP(hipGetLastError());
  }
  scan_u8(x, keep.get(), pos.get(), in.n, tot.get());
  u64 m = read_u64(x, tot.get());
  alloc(actx, out, m);
  if (m) {
    Scope sc(x.timer, "compact_table", KC_JOIN, (double)in.n * 1.0 + m * 4.0 * out->ncols, in.n);
    k_compact_table<<<gridj(in.n), kBlock, 0, x.stream>>>(in, keep.get(), pos.get(), *out);
    SK_HIP(hipGetLastError());
  }
  return m;
}

// ================================================================ project

template <class KO>
__global__ void k_project(Table in, OutDesc o) {
  u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= in.n) return;
  if (o.cd.on) {
    u64 cell;
    if (!cell_of(o.cd, [&](int j) { return o.kind[j] == 0 ? in.col[o.idx[j]][r] : o.cnst[j]; }, &cell)) return;
    cell_insert(o.cd, cell, in.w ? in.w[r] : 0, in.inf ? in.inf[r] : (u8)0);
    return;
  }
  KO key = 0;
  for (int c = 0; c < o.ncols; c++) {
    u32 v = o.kind[c] == 0 ? in.col[o.idx[c]][r] : o.cnst[c];
    key = (key << o.bits) | (KO)v;
  }
  ((KO*)o.keys)[r] = key;
  if (o.w) {
    o.w[r] = in.w ? in.w[r] : 0;
    if (o.inf) o.inf[r] = in.inf ? in.inf[r] : (u8)0;
  }
}

void project(const Exec& x, const Table& in, OutDesc* o, bool k128_out, HeadAlloc halloc, void* hctx) {
  halloc(hctx, o, in.n);
  if (!in.n) return;
  Scope sc(x.timer, "project", KC_JOIN, (double)in.n * (4.0 * o->ncols + (k128_out ? 16 : 8)), in.n);
  if (k128_out) k_project<u128><<<gridj(in.n), kBlock, 0, x.stream>>>(in, *o);
  else k_project<u64><<<gridj(in.n), kBlock, 0, x.stream>>>(in, *o);
  SK_HIP(hipGetLastError());
}

// ========================================================== aggregation

// Group key (packed) and aggregand of each table row.  The aggregand must be
// an Int code (>= first_int); its value comes from the Int table.
template <class K>
__global__ void k_agg_prep(Table in, AggDesc a, const i64* __restrict__ int_vals, K* __restrict__ gk,
                           i64* __restrict__ val, int* __restrict__ bad) {
  u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= in.n) return;
  K key = 0;
  for (int c = 0; c < a.nkey; c++) {
    u32 v = a.key_kind[c] == 0 ? in.col[a.key_idx[c]][r] : a.key_cnst[c];
    key = (key << a.bits) | (K)v;
  }
  gk[r] = key;
  i64 v = 0;
  if (a.op != 0) {
    u32 code = a.val_col >= 0 ? in.col[a.val_col][r] : 0;
    if (code < a.first_int || code >= a.first_int + a.n_ints) {
      atomicOr(bad, 1);
    } else {
      v = int_vals[code - a.first_int];
    }
  }
  val[r] = v;
}

// Segmented reduction over sorted group keys: one output per run.
template <class K>
__global__ void k_agg_reduce(const K* __restrict__ gk, const i64* __restrict__ val, u64 n, int op,
                             const u64* __restrict__ pos, K* __restrict__ ok, i64* __restrict__ ov) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = gk[i];
  if (i > 0 && gk[i - 1] == k) return;
  i64 acc = op == 0 ? 0 : val[i];
  u64 j = i;
  u64 cnt = 0;
  for (; j < n && gk[j] == k; j++) {
    cnt++;
    i64 v = val[j];
    if (j == i) continue;
    if (op == 1) acc = (i64)((u64)acc + (u64)v);
    else if (op == 2) acc = v < acc ? v : acc;
    else if (op == 3) acc = v > acc ? v : acc;
  }
  if (op == 0) acc = (i64)cnt;
  u64 o = pos[i];
  ok[o] = k;
  ov[o] = acc;
}

__global__ void k_run_flags64(const u64* __restrict__ k, u64 n, u8* __restrict__ f) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || k[i] != k[i - 1]);
}
__global__ void k_run_flags128(const u128* __restrict__ k, u64 n, u8* __restrict__ f) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || k[i] != k[i - 1]);
}

template <class K>
u64 aggregate(const Exec& x, const Table& in, const AggDesc& a, const i64* int_vals, int key_bits,
              std::vector<K>* keys_out, std::vector<i64>* vals_out, bool* bad_aggregand) {
  *bad_aggregand = false;
  keys_out->clear();
  vals_out->clear();
  u64 n = in.n;
  if (n == 0) return 0;
  DBuf<K> gk(x.pool, n), gk2(x.pool, n);
  DBuf<i64> v(x.pool, n), v2(x.pool, n);
  DBuf<int> bad(x.pool, 1);
  memset_async(x, bad.get(), 0, sizeof(int));
  k_agg_prep<K><<<gridj(n), kBlock, 0, x.stream>>>(in, a, int_vals, gk.get(), v.get(), bad.get());
  SK_HIP(hipGetLastError());
  K* sk;
  i64* sv;
  radix_sort<K>(x, gk.get(), gk2.get(), v.get(), v2.get(), nullptr, nullptr, n, key_bits, &sk, &sv, nullptr);
  DBuf<u8> fl(x.pool, n);
  DBuf<u64> pos(x.pool, n);
  DBuf<u64> tot(x.pool, 1);
  if constexpr (sizeof(K) == 8) k_run_flags64<<<gridj(n), kBlock, 0, x.stream>>>((const u64*)sk, n, fl.get());
  else k_run_flags128<<<gridj(n), kBlock, 0, x.stream>>>((const u128*)sk, n, fl.get());
  scan_u8(x, fl.get(), pos.get(), n, tot.get());
  u64 m = read_u64(x, tot.get());
  DBuf<K> ok(x.pool, m);
  DBuf<i64> ov(x.pool, m);
  k_agg_reduce<K><<<gridj(n), kBlock, 0, x.stream>>>(sk, sv, n, a.op, pos.get(), ok.get(), ov.get());
  SK_HIP(hipGetLastError());
  int hb = 0;
  SK_HIP(hipMemcpyAsync(&hb, bad.get(), sizeof(int), hipMemcpyDeviceToHost, x.stream));
  keys_out->resize(m);
  vals_out->resize(m);
  SK_HIP(hipMemcpyAsync(keys_out->data(), ok.get(), m * sizeof(K), hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipMemcpyAsync(vals_out->data(), ov.get(), m * sizeof(i64), hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  *bad_aggregand = hb != 0;
  return m;
}

// ============================================================== leapfrog
//
// Triangle body R1(u,v), R2(v,w), R3(u,w) (any column order): for every
// driver row (u, v) of R1 the candidates w are the intersection of two
// sorted adjacency lists, adj2(v) from R2 indexed on its v column and
// adj3(u) from R3 indexed on its u column.  Count pass, device prefix sum,
// emit pass; with the driver scanned in key order and w ascending, the
// bindings (u, v, w) come out sorted.

// Calls f(w) for every w in both sorted lists, ascending.  Lists of similar
// length are merged; when one is much shorter (its length times the longer
// one's log2 below the merge's length) each of its values is found in the
// longer list by a binary search over the part not yet passed.
template <class F>
__device__ __forceinline__ void intersect_each(const u32* __restrict__ A, u64 na, const u32* __restrict__ B, u64 nb,
                                               F f) {
  if (na > nb) {
    const u32* t = A;
    A = B;
    B = t;
    const u64 tn = na;
    na = nb;
    nb = tn;
  }
  if (na == 0) return;
  const u64 lg = 64 - (u64)__builtin_clzll(nb);  // bits of nb
  if (na * (lg + 1) < na + nb) {
    u64 j = 0;
    for (u64 i = 0; i < na && j < nb; i++) {
      const u32 x = A[i];
      u64 lo = j, hi = nb;  // first position in [j, nb) holding a value >= x
      while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (B[mid] < x) lo = mid + 1;
        else hi = mid;
      }
      j = lo;
      if (j < nb && B[j] == x) {
        f(x);
        j++;
      }
    }
    return;
  }
  u64 i = 0, j = 0;
  while (i < na && j < nb) {
    const u32 x = A[i], y = B[j];
    if (x == y) f(x);
    i += (x <= y);
    j += (y <= x);
  }
}

// A driver row whose two lists hold more than kTriHeavy values together is
// left to a wave of its own (k_tri_count_heavy, k_tri_emit_heavy): one lane
// per row would hold its whole wave for the length of the longest merge.
// (RMAT scale 22, ms per listing: 0 -> 59.6, 32 -> 59.7, 64 -> 60.0,
// 128 -> 62.1, 256 -> 69.2; rows this short waste most of a wave's lanes on
// sparser graphs.)
#ifndef SK_TRI_HEAVY
#define SK_TRI_HEAVY 64
#endif
constexpr u64 kTriHeavy = SK_TRI_HEAVY;

// Rows per thread of the light count: a block of kBlock threads takes
// kTriRows * kBlock consecutive driver rows and reserves its heavy-list slots
// with one pair of device atomics (same-address device atomics serialize:
// one pair per wave cost 24 ms over 1 M waves at RMAT scale 22).
constexpr int kTriRows = 16;
static_assert(kTriRows * (kBlock / 64) == 64, "one slot count per (row step, wave): a wave scans them");

// Per driver row (u, v): the two lists' lengths from the CSRs.  A light row
// counts its common values by a merge (or binary searches, for a short list
// against a long one) and writes cnt; a heavy row is appended to the heavy
// list (heavy, hshort), in (row step, wave, lane) order within the block.
template <class K>
__global__ __launch_bounds__(kBlock) void k_tri_count(const K* __restrict__ d_keys, u64 nd, TriDesc t,
                                                      const u64* __restrict__ csr2, const u32* __restrict__ adj2,
                                                      const u64* __restrict__ csr3, const u32* __restrict__ adj3,
                                                      u32* __restrict__ cnt, u32* __restrict__ heavy,
                                                      u32* __restrict__ hshort, unsigned long long* __restrict__ hstat) {
  __shared__ u32 s_off[kTriRows * (kBlock / 64)];  // per (row step, wave): heavy rows, then first slot
  __shared__ u64 s_len[kBlock / 64];
  __shared__ u32 s_hsh[kTriRows][kBlock];  // a heavy row's shorter list length
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const u64 e0 = (u64)blockIdx.x * (kTriRows * kBlock) + threadIdx.x;
  u32 hbits = 0;  // bit r: row step r is heavy
  u64 lens = 0;   // the heavy rows' list values
#pragma unroll 1
  for (int r = 0; r < kTriRows; r++) {
    const u64 e = e0 + (u64)r * kBlock;
    if (e >= nd) continue;
    const K k = d_keys[e];
    const u32 u = col_of<K>(k, 2, t.bits, t.u_col);
    const u32 v = col_of<K>(k, 2, t.bits, 1 - t.u_col);
    const u64 a0 = csr2[v], a1 = csr2[v + 1];
    const u64 b0 = csr3[u], b1 = csr3[u + 1];
    if ((a1 - a0) + (b1 - b0) > kTriHeavy) {
      hbits |= 1u << r;
      s_hsh[r][threadIdx.x] = (u32)min(a1 - a0, b1 - b0);
      lens += (a1 - a0) + (b1 - b0);
      cnt[e] = 0;  // written by the heavy pass
      continue;
    }
    u32 c = 0;
    intersect_each(adj2 + a0, a1 - a0, adj3 + b0, b1 - b0, [&](u32) { c++; });
    cnt[e] = c;
  }
  // the block's heavy rows and list values, then its slots: one wave scans
  // the (row step, wave) counts
#pragma unroll
  for (int r = 0; r < kTriRows; r++) {
    const u64 m = __ballot((hbits >> r) & 1u);
    if (lane == 0) s_off[r * (kBlock / 64) + wv] = (u32)__popcll(m);
  }
  for (u32 o = 32; o > 0; o >>= 1) lens += __shfl_xor(lens, (int)o, 64);
  if (lane == 0) s_len[wv] = lens;
  __syncthreads();
  if (wv == 0) {
    const u32 c = s_off[lane];
    const u32 incl = wave_incl_sum_u32(c);
    u64 base = 0;
    if (lane == 63 && incl) {
      u64 lt = 0;
      for (int w = 0; w < kBlock / 64; w++) lt += s_len[w];
      base = atomicAdd(&hstat[0], (unsigned long long)incl);
      atomicAdd(&hstat[1], (unsigned long long)lt);
    }
    base = __shfl(base, 63, 64);
    s_off[lane] = (u32)base + incl - c;  // heavy-list positions fit u32 (nd < 2^32)
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kTriRows; r++) {
    const bool hv = (hbits >> r) & 1u;
    const u64 m = __ballot(hv);
    if (hv) {
      const u64 h = (u64)s_off[r * (kBlock / 64) + wv] + (u64)__popcll(m & ((1ull << lane) - 1));
      heavy[h] = (u32)(e0 + (u64)r * kBlock);
      hshort[h] = s_hsh[r][threadIdx.x];
    }
  }
}

// The lists of a heavy row, the shorter first.
template <class K>
__device__ __forceinline__ void tri_heavy_lists(const K* __restrict__ d_keys, u64 e, const TriDesc& t,
                                                const u64* __restrict__ csr2, const u32* __restrict__ adj2,
                                                const u64* __restrict__ csr3, const u32* __restrict__ adj3, u32* u,
                                                u32* v, const u32** A, u64* na, const u32** B, u64* nb,
                                                bool* a_in3 = nullptr) {
  const K k = d_keys[e];
  *u = col_of<K>(k, 2, t.bits, t.u_col);
  *v = col_of<K>(k, 2, t.bits, 1 - t.u_col);
  const u64 a0 = csr2[*v], a1 = csr2[*v + 1], b0 = csr3[*u], b1 = csr3[*u + 1];
  *A = adj2 + a0;
  *na = a1 - a0;
  *B = adj3 + b0;
  *nb = b1 - b0;
  if (a_in3) *a_in3 = *na > *nb;
  if (*na > *nb) {
    const u32* tp = *A;
    *A = *B;
    *B = tp;
    const u64 tn = *na;
    *na = *nb;
    *nb = tn;
  }
}

// x in the sorted, duplicate-free list B[0, nb)
__device__ __forceinline__ bool tri_member(const u32* __restrict__ B, u64 nb, u32 x) {
  u64 lo = 0, hi = nb;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (B[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < nb && B[lo] == x;
}

// Heavy rows keep, per 64 values of their shorter list, a word of hit bits
// (hmask, at the row's word offset hwoff[h]): the count pass finds the
// common values once, and the emit pass only expands the bits.
#ifndef SK_TRI_LDS
#define SK_TRI_LDS 1024
#endif
constexpr u64 kTriLds = SK_TRI_LDS;  // longer lists up to this length are searched in LDS
// Threads per block of the heavy-row kernels: one wave, so that a block's
// slot and LDS are not held by its three siblings' shorter rows.
#ifndef SK_TRI_BLOCK
#define SK_TRI_BLOCK 64
#endif
constexpr int kTriBlock = SK_TRI_BLOCK;

__global__ void k_tri_heavy_words(const u32* __restrict__ hshort, u64 nh, u32* __restrict__ nw) {
  const u64 h = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < nh) nw[h] = (u32)((hshort[h] + 63) / 64);
}

// A wave-uniform 64-bit value as the compiler's scalar (read from lane 0).
__device__ __forceinline__ u64 uni_u64(u64 x) {
  const u32 lo = (u32)__builtin_amdgcn_readfirstlane((int)(u32)x);
  const u32 hi = (u32)__builtin_amdgcn_readfirstlane((int)(u32)(x >> 32));
  return ((u64)hi << 32) | lo;
}

// Membership of G values in the sorted, duplicate-free list L[0, nb) held in
// LDS (nb >= 1), searched together: the step sizes depend on nb alone, so every
// lane takes the same ceil(log2 nb) steps with no divergence, each one LDS read,
// a compare and a select per value.  b tracks the last position whose value is
// <= x (position 0 when there is none); x is in L exactly when L[b] == x.
template <int G>
__device__ __forceinline__ void tri_member_lds_n(const u32* L, u32 nb, const u32 (&x)[G], bool (&hit)[G]) {
  const char* Lb = reinterpret_cast<const char*>(L);
  u32 b[G];  // in bytes
#pragma unroll
  for (int g = 0; g < G; g++) b[g] = 0;
  for (u32 n = nb; n > 1;) {
    const u32 h = n >> 1;
#pragma unroll
    for (int g = 0; g < G; g++) {
      const u32 p = b[g] + 4 * h;
      b[g] = *reinterpret_cast<const u32*>(Lb + p) <= x[g] ? p : b[g];
    }
    n -= h;
  }
#pragma unroll
  for (int g = 0; g < G; g++) hit[g] = *reinterpret_cast<const u32*>(Lb + b[g]) == x[g];
}

// G chunks of 64 values of the shorter list, from chunk k0: their values are
// loaded together, searched together, and each chunk's ballot is its hit word.
template <int G>
__device__ __forceinline__ u32 tri_heavy_chunks(const u32* L, const u32* __restrict__ A, u64 na,
                                                const u32* __restrict__ B, u64 nb, bool lds, u64 k0, u32 lane,
                                                u64* __restrict__ words) {
  u32 x[G];
  bool in[G], hit[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const u64 i = (k0 + g) * 64 + lane;
    in[g] = i < na;
    x[g] = in[g] ? A[i] : 0u;
  }
  if (lds) {
    tri_member_lds_n<G>(L, (u32)nb, x, hit);
  } else {
#pragma unroll
    for (int g = 0; g < G; g++) hit[g] = tri_member(B, nb, x[g]);
  }
  u32 c = 0;
#pragma unroll
  for (int g = 0; g < G; g++) {
    const u64 m = __ballot(in[g] && hit[g]);
    if (lane == 0) words[k0 + g] = m;
    c += (u32)__popcll(m);
  }
  return c;
}

// One wave per heavy row: lane l tests values l, l + 64, ... of the shorter
// list against the longer one, in LDS when the longer list fits kTriLds
// values (its loads all issued before the first LDS write), else by binary
// search in global memory.  Up to four chunks of 64 values go through the
// search together; each chunk's ballot is the row's next hit word.
template <class K>
__global__ __launch_bounds__(kTriBlock) void k_tri_count_heavy(const K* __restrict__ d_keys, TriDesc t,
                                                            const u64* __restrict__ csr2,
                                                            const u32* __restrict__ adj2,
                                                            const u64* __restrict__ csr3,
                                                            const u32* __restrict__ adj3,
                                                            const u32* __restrict__ heavy, u64 nh,
                                                            const u64* __restrict__ hwoff, u64* __restrict__ hmask,
                                                            u64* __restrict__ hrec, u32* __restrict__ cnt) {
  __shared__ u32 sB[kTriBlock / 64][kTriLds];
  const u64 w = (u64)blockIdx.x * (kTriBlock / 64) + (threadIdx.x >> 6);
  if (w >= nh) return;  // whole waves; no block barrier below
  const u32 lane = threadIdx.x & 63;
  u32* L = sB[threadIdx.x >> 6];
  // the row's values are the same in every lane: held as scalars, so that the
  // chunk and search loops below branch on scalars
  const u64 e = (u32)__builtin_amdgcn_readfirstlane((int)heavy[w]);
  u64* words = (u64*)uni_u64((u64)(hmask + hwoff[w]));
  u32 u, v;
  const u32 *A, *B;
  u64 na, nb;
  bool in3 = false;
  tri_heavy_lists<K>(d_keys, e, t, csr2, adj2, csr3, adj3, &u, &v, &A, &na, &B, &nb, &in3);
  in3 = __builtin_amdgcn_readfirstlane((int)in3) != 0;
  {
    const u32* Ab = in3 ? adj3 : adj2;  // the lists' arrays: offsets as scalars
    const u32* Bb = in3 ? adj2 : adj3;
    A = Ab + uni_u64((u64)(A - Ab));
    B = Bb + uni_u64((u64)(B - Bb));
  }
  na = uni_u64(na);
  nb = uni_u64(nb);
  const bool lds = nb <= kTriLds;
  if (lds) {
    constexpr int kPer = (int)(kTriLds / 64);
    const u32 nb32 = (u32)nb, nbk = (nb32 + 63) / 64;
    u32 tv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const u32 i = (u32)k * 64 + lane;
      tv[k] = 0;
      if ((u32)k < nbk && i < nb32) tv[k] = B[i];
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const u32 i = (u32)k * 64 + lane;
      if ((u32)k < nbk && i < nb32) L[i] = tv[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  const u64 nc = (na + 63) / 64;
  u32 c = 0;
  for (u64 k0 = 0; k0 < nc; k0 += 4) {
    switch (min(nc - k0, (u64)4)) {
      case 4: c += tri_heavy_chunks<4>(L, A, na, B, nb, lds, k0, lane, words); break;
      case 3: c += tri_heavy_chunks<3>(L, A, na, B, nb, lds, k0, lane, words); break;
      case 2: c += tri_heavy_chunks<2>(L, A, na, B, nb, lds, k0, lane, words); break;
      default: c += tri_heavy_chunks<1>(L, A, na, B, nb, lds, k0, lane, words); break;
    }
  }
  if (lane == 0) {
    cnt[e] = c;
    // the row for the emit pass: where its shorter list starts (bit 63: in
    // adj3, else adj2) and its binding (u, v)
    hrec[2 * w] = ((u64)in3 << 63) | (u64)(A - (in3 ? adj3 : adj2));
    hrec[2 * w + 1] = ((u64)u << 32) | v;
  }
}

// After the scan: each heavy row's output offset, in heavy-list order.
__global__ void k_tri_heavy_off(const u32* __restrict__ heavy, u64 nh, const u64* __restrict__ off,
                                u64* __restrict__ hoff) {
  const u64 h = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < nh) hoff[h] = off[heavy[h]];
}

// The r-th (from 0) set bit of m, which has more than r set bits: halve the
// window six times, keeping the half that holds it.
__device__ __forceinline__ u32 select_bit64(u64 m, u32 r) {
  u32 pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const u32 c = (u32)__popcll(m & ((1ull << w) - 1));
    if (r >= c) {
      r -= c;
      m >>= w;
      pos += (u32)w;
    }
  }
  return pos;
}

// A hit word of the heavy emit, as its lanes read it: the word, the output
// index of its hit q (of the chunk's hits) less q, where in adj2/adj3 its 64
// list values start (bits 0-39; bit 63: adj3) with the chunk's hits before it
// (bits 40-52), and its row's binding (u, v).
struct TriWord {
  u64 m, ob, ao, uv;
  u64 klo, khi;  // the word's key without its hit value (when that value is one column): low, high 64 bits
};

// One wave per 64 heavy rows, over their hit words as one range (rows' words
// are contiguous, as hwoff is a scan in heavy-row order), 64 words at a time:
// lane j takes word j, finds its row (each row's first word marked in LDS with
// the row, spread by a prefix maximum; a row begun in an earlier chunk carries
// over with its hits so far), and the hits before it in its row (prefix sum of
// the words' bit counts).  Then the chunk's hits, 64 at a time, one per lane:
// hit q's word by the same marking (each word marks the rank of its first
// hit), its bit by selecting set bit q - (hits before the word).  A hit is the
// common value A[64 c + bit] of its row's word c, written at the row's output
// offset plus its rank among the row's hits (ascending, as A is); a row's hits
// fill consecutive outputs, so 64 lanes write mostly consecutive keys.
template <class KO>
__global__ __launch_bounds__(kTriBlock) void k_tri_emit_heavy(const u32* __restrict__ adj2, const u32* __restrict__ adj3,
                                                           u64 nh, const u64* __restrict__ hrec,
                                                           const u32* __restrict__ hshort,
                                                           const u64* __restrict__ hwoff,
                                                           const u64* __restrict__ hoff,
                                                           const u64* __restrict__ hmask, OutDesc o) {
  __shared__ TriWord sw[kTriBlock / 64][64];
  __shared__ u32 smk[kTriBlock / 64][64];
  const u64 g = (u64)blockIdx.x * (kTriBlock / 64) + (threadIdx.x >> 6);
  const u64 h0 = g * 64;
  if (h0 >= nh) return;  // whole waves; no block barrier below
  const u32 lane = threadIdx.x & 63;
  TriWord* W = sw[threadIdx.x >> 6];
  u32* mk = smk[threadIdx.x >> 6];
  const u32 nrows = (u32)min((u64)64, nh - h0);
  // when the hit value w is exactly one column, a key is the word's key
  // without it (folded once per word) OR w shifted into its column
  int wcols = 0, wshift = 0;
  for (int c = 0; c < o.ncols; c++)
    if (o.kind[c] == 0 && o.idx[c] == 2) {
      wcols++;
      wshift = o.bits * (o.ncols - 1 - c);
    }
  const bool wone = wcols == 1;
  u64 ra = 0, ruv = 0, rw = 0, ro = 0, rn = 0;  // lane l: row h0 + l
  if (lane < nrows) {
    const u64 h = h0 + lane;
    ra = hrec[2 * h];
    ruv = hrec[2 * h + 1];
    rw = hwoff[h];
    ro = hoff[h];
    rn = hshort[h];
  }
  const u64 w0 = __shfl(rw, 0, 64);
  const u64 w1 = __shfl(rw + (rn + 63) / 64, (int)nrows - 1, 64);
  mk[lane] = 0;
  u32 crow = 0;   // the row of the previous chunk's last word
  u64 chits = 0;  // and its hits up to the end of that chunk
  for (u64 c0 = w0; c0 < w1; c0 += 64) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // a row without words starts where the next row does: the later row wins
    if (lane < nrows && rw >= c0 && rw < c0 + 64) atomicMax(&mk[rw - c0], lane + 1);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const u32 rmk = mk[lane];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    mk[lane] = 0;
    const u32 pr = wave_incl_max_u32(rmk);
    const u32 rj = pr ? pr - 1 : crow;
    const u64 wi = c0 + lane;
    const u64 m = wi < w1 ? hmask[wi] : 0ull;
    const u32 pc = (u32)__popcll(m);
    const u32 incl = wave_incl_sum_u32(pc);
    const u32 ex = incl - pc;
    const u64 rs = __shfl(rw, (int)rj, 64);  // the row's first word
    const u32 ps = __shfl(ex, (int)(rs > c0 ? rs - c0 : 0), 64);
    const u64 before = (u64)(ex - ps) + (rs < c0 ? chits : 0ull);  // hits before this word in its row
    const u64 a = __shfl(ra, (int)rj, 64);
    TriWord tw;
    tw.m = m;
    tw.ob = __shfl(ro, (int)rj, 64) + before - ex;
    tw.ao = (a & (1ull << 63)) | ((u64)ex << 40) | ((a & ((1ull << 40) - 1)) + (wi - rs) * 64);
    tw.uv = __shfl(ruv, (int)rj, 64);
    if (wone) {
      KO kp = 0;
      for (int c = 0; c < o.ncols; c++) {
        // binding slots: 0 = u, 1 = v, 2 = w (0 here)
        const u32 val = o.kind[c] != 0 ? o.cnst[c] : o.idx[c] == 0 ? (u32)(tw.uv >> 32) : o.idx[c] == 1 ? (u32)tw.uv : 0u;
        kp = (kp << o.bits) | (KO)val;
      }
      tw.klo = (u64)kp;
      tw.khi = sizeof(KO) > 8 ? (u64)(kp >> 32 >> 32) : 0ull;
    }
    W[lane] = tw;
    crow = (u32)__builtin_amdgcn_readlane((int)rj, 63);
    chits = __shfl(before + pc, 63, 64);
    const u32 nhit = (u32)__builtin_amdgcn_readlane((int)incl, 63);
    u32 cw = 0;  // the word of the previous batch's last hit
    for (u32 t = 0; t < nhit; t += 64) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // words with hits start at distinct ranks
      if (pc && ex >= t && ex < t + 64) mk[ex - t] = lane + 1;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const u32 hmk = mk[lane];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      mk[lane] = 0;
      const u32 pw = wave_incl_max_u32(hmk);
      const u32 j = pw ? pw - 1 : cw;
      cw = (u32)__builtin_amdgcn_readlane((int)j, 63);
      const u32 q = t + lane;
      if (q < nhit) {
        const TriWord r = W[j];
        const u32 bit = select_bit64(r.m, q - (u32)((r.ao >> 40) & 0x1FFFu));
        const u32* A = (r.ao >> 63) ? adj3 : adj2;
        const u32 x = A[(r.ao & ((1ull << 40) - 1)) + bit];
        KO key = 0;
        if (wone) {
          key = (((KO)r.khi << 32 << 32) | (KO)r.klo) | ((KO)x << wshift);
        } else {
          const u32 u = (u32)(r.uv >> 32), v = (u32)r.uv;
          for (int c = 0; c < o.ncols; c++) {
            // binding slots: 0 = u, 1 = v, 2 = w
            const u32 val = o.kind[c] != 0 ? o.cnst[c] : o.idx[c] == 0 ? u : o.idx[c] == 1 ? v : x;
            key = (key << o.bits) | (KO)val;
          }
        }
        ((KO*)o.keys)[r.ob + q] = key;
      }
    }
  }
}

template <class K, class KO>
__global__ void k_tri_emit(const K* __restrict__ d_keys, u64 nd, TriDesc t, const u64* __restrict__ csr2,
                           const u32* __restrict__ adj2, const u64* __restrict__ csr3, const u32* __restrict__ adj3,
                           const u64* __restrict__ off, const u32* __restrict__ cnt, OutDesc o) {
  u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nd) return;
  const u32 need = cnt[e];  // the count pass's result: rows without a triangle skip the merge
  if (!need) return;
  K k = d_keys[e];
  u32 u = col_of<K>(k, 2, t.bits, t.u_col);
  u32 v = col_of<K>(k, 2, t.bits, 1 - t.u_col);
  u64 a = csr2[v], a1 = csr2[v + 1];
  u64 b = csr3[u], b1 = csr3[u + 1];
  if ((a1 - a) + (b1 - b) > kTriHeavy) return;  // k_tri_emit_heavy's
  u64 oi = off[e];
  const u64 oend = oi + need;
  // a plain merge here (the adaptive search made this pass slower, 384 ->
  // 500 ms on RMAT scale 22); the count pass finds the same set, and the
  // merge stops at the row's last triangle
  while (a < a1 && b < b1 && oi < oend) {
    u32 x = adj2[a], y = adj3[b];
    if (x == y) {
      KO key = 0;
      for (int c = 0; c < o.ncols; c++) {
        // binding slots: 0 = u, 1 = v, 2 = w
        const u32 val = o.kind[c] != 0 ? o.cnst[c] : o.idx[c] == 0 ? u : o.idx[c] == 1 ? v : x;
        key = (key << o.bits) | (KO)val;
      }
      ((KO*)o.keys)[oi++] = key;
    }
    a += (x <= y);
    b += (y <= x);
  }
}

template <class K>
u64 triangles(const Exec& x, const K* d_keys, u64 nd, const TriDesc& t, const u64* csr2, const u32* adj2,
              const u64* csr3, const u32* adj3, OutDesc* o, bool k128_out, bool count_only, HeadAlloc halloc,
              void* hctx, double* sum_list_len) {
  DBuf<u32> cnt(x.pool, nd ? nd : 1);
  DBuf<u64> off(x.pool, nd ? nd : 1);
  DBuf<u64> tot(x.pool, 1);
  if (nd == 0) {
    if (!count_only) halloc(hctx, o, 0);
    return 0;
  }
  // heavy rows are listed by u32 index and take a one-wave block each
  if (nd > 0x7ffffff0ull) throw BudgetError{nd};
  double lb = sum_list_len ? *sum_list_len : 0;
  DBuf<u32> heavy(x.pool, nd), hshort(x.pool, nd);
  DBuf<u64> hstat(x.pool, 2);
  memset_async(x, hstat.get(), 0, 16);
  {
    // bytes credited once the heavy rows' share of the lists is known
    Scope sc(x.timer, "leapfrog_count", KC_LEAPFROG, 0.0, nd);
    k_tri_count<K><<<gridj(nd, (u64)kTriRows * kBlock), kBlock, 0, x.stream>>>(d_keys, nd, t, csr2, adj2, csr3,
                                                                             adj3, cnt.get(),
                                                      heavy.get(), hshort.get(), (unsigned long long*)hstat.get());
    SK_HIP(hipGetLastError());
  }
  u64 hs[2];
  read_u64s(x, hstat.get(), hs, 2);
  const u64 nh = hs[0];
  const double heavy_vals = (double)hs[1], light_vals = std::max(0.0, lb - heavy_vals);
  const unsigned hgrid = (unsigned)((nh + kTriBlock / 64 - 1) / (kTriBlock / 64));
  if (x.timer) x.timer->credit("leapfrog_count", KC_LEAPFROG, (double)nd * sizeof(K) + light_vals * 4.0, 0);
  // the heavy rows' hit words: ceil(shorter list / 64) per row
  DBuf<u32> hnw(x.pool, nh ? nh : 1);
  DBuf<u64> hwoff(x.pool, nh ? nh : 1), hwtot(x.pool, 1), hrec(x.pool, 2 * (nh ? nh : 1)), hoff(x.pool, nh ? nh : 1);
  DBuf<u64> hmask;
  if (nh) {
    k_tri_heavy_words<<<gridj(nh), kBlock, 0, x.stream>>>(hshort.get(), nh, hnw.get());
    SK_HIP(hipGetLastError());
    scan_u32(x, hnw.get(), hwoff.get(), nh, hwtot.get());
    hmask.reset(x.pool, std::max<u64>(read_u64(x, hwtot.get()), 1));
    Scope sc(x.timer, "leapfrog_count_heavy", KC_LEAPFROG, (double)nh * 4.0 + heavy_vals * 4.0, nh);
    k_tri_count_heavy<K><<<hgrid, kTriBlock, 0, x.stream>>>(d_keys, t, csr2, adj2, csr3, adj3, heavy.get(), nh,
                                                         hwoff.get(), hmask.get(), hrec.get(), cnt.get());
    SK_HIP(hipGetLastError());
  }
  scan_u32(x, cnt.get(), off.get(), nd, tot.get());
  u64 total = read_u64(x, tot.get());
  if (count_only) return total;
  halloc(hctx, o, total);
  if (total == 0) return 0;
  {
    Scope sc(x.timer, "leapfrog_emit", KC_LEAPFROG,
             (double)nd * sizeof(K) + light_vals * 4.0 + (double)total * (k128_out ? 16.0 : 8.0), nd);
    if (k128_out)
      k_tri_emit<K, u128><<<gridj(nd), kBlock, 0, x.stream>>>(d_keys, nd, t, csr2, adj2, csr3, adj3, off.get(),
                                                               cnt.get(), *o);
    else
      k_tri_emit<K, u64><<<gridj(nd), kBlock, 0, x.stream>>>(d_keys, nd, t, csr2, adj2, csr3, adj3, off.get(),
                                                              cnt.get(), *o);
    SK_HIP(hipGetLastError());
  }
  if (nh) {
    k_tri_heavy_off<<<gridj(nh), kBlock, 0, x.stream>>>(heavy.get(), nh, off.get(), hoff.get());
    SK_HIP(hipGetLastError());
    // per heavy row its record (16), word and output offsets (16), list
    // length (4); per hit its shorter-list value (4) and the output written
    Scope sc(x.timer, "leapfrog_emit_heavy", KC_LEAPFROG, 0.0, nh);
    const unsigned egrid = (unsigned)((nh + 64 * (kTriBlock / 64) - 1) / (64 * (kTriBlock / 64)));
    if (k128_out)
      k_tri_emit_heavy<u128><<<egrid, kTriBlock, 0, x.stream>>>(adj2, adj3, nh, hrec.get(), hshort.get(), hwoff.get(),
                                                             hoff.get(), hmask.get(), *o);
    else
      k_tri_emit_heavy<u64><<<egrid, kTriBlock, 0, x.stream>>>(adj2, adj3, nh, hrec.get(), hshort.get(), hwoff.get(),
                                                            hoff.get(), hmask.get(), *o);
    SK_HIP(hipGetLastError());
  }
  return total;
}

// Σ over driver rows of |adj2(v)| + |adj3(u)| (the leapfrog's list reads).
template <class K>
__global__ __launch_bounds__(kBlock) void k_tri_listlen(const K* __restrict__ d_keys, u64 nd, TriDesc t,
                                                        const u64* __restrict__ csr2, const u64* __restrict__ csr3,
                                                        unsigned long long* __restrict__ acc) {
  __shared__ u64 part[kBlock / 64];
  u64 s = 0;
  for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < nd; e += (u64)gridDim.x * blockDim.x) {
    K k = d_keys[e];
    u32 u = col_of<K>(k, 2, t.bits, t.u_col);
    u32 v = col_of<K>(k, 2, t.bits, 1 - t.u_col);
    s += (csr2[v + 1] - csr2[v]) + (csr3[u + 1] - csr3[u]);
  }
  s = wave_incl_scan_j(s);
  if (lane_id_j() == 63) part[wave_id_j()] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 tt = 0;
    for (int w = 0; w < kBlock / 64; w++) tt += part[w];
    if (tt) atomicAdd(acc, (unsigned long long)tt);
  }
}

template <class K>
double tri_list_len(const Exec& x, const K* d_keys, u64 nd, const TriDesc& t, const u64* csr2, const u64* csr3) {
  if (!nd) return 0;
  DBuf<u64> acc(x.pool, 1);
  memset_async(x, acc.get(), 0, 8);
  k_tri_listlen<K><<<(unsigned)std::min<u64>(gridj(nd), 2048), kBlock, 0, x.stream>>>(d_keys, nd, t, csr2, csr3,
                                                                                     (unsigned long long*)acc.get());
  SK_HIP(hipGetLastError());
  return (double)read_u64(x, acc.get());
}

// ================================================================== pull
//
// A round of `H(r, x) :- H(j, x), E(j, r)` (or the mirrored column order)
// evaluated from the destination side: E's index by r lists each head row's
// in-edges (j, w); the last round's delta of H is read through per-row
// records {lanes of row j in the delta, position of row j's first delta
// cell}.  The candidates are exactly the push join's (every (delta cell,
// edge) pair), so the round's result, fresh set and errors are the same;
// what changes is that each head cell is written once per chunk instead of
// once per candidate.

// Threads per block of the pull kernels (k_pull, k_pull_packed, k_bitpull):
// one wave, one chunk.  With four waves a block held its LDS and slot until
// its slowest chunk was done (chunks range from one index row to 256):
// config 2, 0.759 -> 0.714 ms per packed pull launch.
#ifndef SK_PULL_BLOCK
#define SK_PULL_BLOCK 64
#endif
constexpr int kPullBlock = SK_PULL_BLOCK;

// Pairs per lane and step of the pull kernel.
constexpr int kPullPairs = 2;

// Per wave of the pull kernel: one batch of 64 index rows.
struct PullLds {
  u64 base[64];               // delta-list position of each index row's source row's first delta cell
  i64 w[64];                  // edge weight
  u32 excl[64];               // exclusive prefix of the index rows' pair counts
  u32 slot[64];               // which of the chunk's destination rows the index row feeds
  i64 acc[kPullRows * 64];    // running minimum per (destination row, lane value)
  u8 have[kPullRows * 64];    // (destination row, lane value) has a candidate
  u8 start[64 * kPullPairs];  // 1 + the index row whose pairs start at this position of the step
};

// Rows of at most 64 lanes (lbits <= 6).  A wave takes one chunk (up to
// kPullRows destination rows) in batches of 64 index rows: lane l loads row
// l's source, weight and the source's delta record, a wave scan counts the
// (edge, delta lane) pairs, and the pairs are dealt round-robin to lanes, so
// every lane does work whatever the masks' density.  A lane finds its pair's
// edge from the step's edge starts: every edge whose first pair falls in the
// step marks that position in LDS, a ballot of the marks gives the start
// mask, and the edge of pair p is the one marked at the highest start at or
// below p (or, with none in the step, the edge the step began in).  The
// pair's delta cell sits at base + rank in the delta list, where it reads
// the cell's weight and lane value (pairs of one edge read consecutive
// positions).  Candidates fold into the wave's LDS minimum per (destination
// row, lane value), applied to the cells once per chunk.  The destination
// cells are peeked at the start, overlapping the loads.

template <bool TROP>
__global__ __launch_bounds__(kPullBlock) void k_pull(PullDesc pd, CellDesc cd, DevErr* err) {
  __shared__ PullLds sh[kPullBlock / 64];
  const u32 lane = threadIdx.x & 63;
  const u64 ch = (u64)blockIdx.x * (kPullBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ch >= pd.nchunks) return;  // whole waves leave; no block barrier below
  if (pd.gate && (pd.gate->done | pd.gate->leave | pd.gate->stamp)) return;  // a push round runs instead
  PullLds& L = sh[threadIdx.x >> 6];
  const PullChunk c = pd.chunks[ch];
  const u32 nr = c.nr_off & 0xFFu;
  const u64 lo = c.lo, hi = c.lo + c.n;
  const int lb = pd.lbits;
  const u32 nl = 1u << lb;
  const u32 ncell = nr << lb;  // <= kPullRows * 64
  for (u32 i = lane; i < kPullRows * 64; i += 64) {
    L.acc[i] = LLONG_MAX;
    L.have[i] = 0;
  }
  u64 seen[kPullRows];
#pragma unroll
  for (u32 u = 0; u < kPullRows; u++) {
    const u32 i = lane + 64 * u;
    seen[u] = 0;
    if (i < ncell) seen[u] = cell_peek(cd, ((u64)c.row[i >> lb] << lb) + (i & (nl - 1)));
  }
  for (u64 e0 = lo; e0 < hi; e0 += 64) {
    const u64 e = e0 + lane;
    u64 m = 0, base = 0;
    i64 we = 0;
    u32 slot = 0;
    if (e < hi) {
      const u32 q = pd.adj[e];
      if constexpr (TROP) we = pd.w[e];
      if (q < pd.nrows) {
        // one 16-byte record per source row: its delta lanes and delta-list position
        const ulonglong2 rec = *(const ulonglong2*)(pd.rowrec + 2 * (u64)q);
        m = rec.x;
        base = rec.y;
      }
      const u32 el = (u32)(e - lo);
      for (u32 s = 1; s < nr; s++) slot += el >= ((c.nr_off >> (8 * s)) & 0xFFu) ? 1u : 0u;
    }
    const u32 cnt = (u32)__popcll(m);
    const u32 incl = (u32)wave_incl_scan_j(cnt);
    const u32 total = __shfl(incl, 63);
    const u32 excl = incl - cnt;
    L.base[lane] = base;
    L.w[lane] = we;
    L.excl[lane] = excl;
    L.slot[lane] = slot;
#pragma unroll
    for (int u = 0; u < kPullPairs; u++) L.start[u * 64 + lane] = 0;
    const u64 le = ~0ull >> (63 - lane);  // lanes at or below this one
    u32 carry = 0;  // the edge holding the step's first pair (pair 0 starts an edge)
    // kPullPairs pairs per lane per step: their delta-weight loads are
    // independent, so they overlap instead of queueing
    for (u32 p0 = 0; p0 < total; p0 += 64 * kPullPairs) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (cnt && excl >= p0 && excl < p0 + 64 * kPullPairs) L.start[excl - p0] = (u8)(lane + 1);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      u32 rs[kPullPairs];
#pragma unroll
      for (int u = 0; u < kPullPairs; u++) rs[u] = L.start[u * 64 + lane];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int u = 0; u < kPullPairs; u++) L.start[u * 64 + lane] = 0;
      u32 jj[kPullPairs], kk[kPullPairs], xx[kPullPairs];
      i64 vv[kPullPairs];
#pragma unroll
      for (int u = 0; u < kPullPairs; u++) {
        // the edge holding pair pidx: marked at the highest start at or
        // below this lane, else the edge the window began in
        const u64 sm = __ballot(rs[u] != 0);
        const u64 mine = sm & le;
        const int h = mine ? 63 - __builtin_clzll(mine) : 0;
        const u32 from = (u32)__shfl((int)rs[u], h, 64);
        const u32 j = mine ? from - 1 : carry;
        if (sm) carry = (u32)__builtin_amdgcn_readlane((int)rs[u], 63 - __builtin_clzll(sm)) - 1;
        const u32 pidx = p0 + (u32)u * 64 + lane;
        jj[u] = j;
        kk[u] = pidx - L.excl[j];
      }
#pragma unroll
      for (int u = 0; u < kPullPairs; u++) {
        // the pair's delta cell: its weight and its lane value
        vv[u] = 0;
        xx[u] = 0;
        if (p0 + (u32)u * 64 + lane < total) {
          const u64 di = L.base[jj[u]] + kk[u];
          xx[u] = pd.dlane[di];
          if constexpr (TROP) vv[u] = pd.dw[di];
        }
      }
#pragma unroll
      for (int u = 0; u < kPullPairs; u++) {
        if (p0 + (u32)u * 64 + lane >= total) continue;
        const u32 j = jj[u];
        const u32 xl = xx[u];
        const u32 a = L.slot[j] * 64 + xl;
        if constexpr (TROP) {
          i64 res;
          u8 rf;
          if (pd.delta_first) trop_mul(vv[u], 0, L.w[j], 0, &res, &rf, err);
          else trop_mul(L.w[j], 0, vv[u], 0, &res, &rf, err);
          atomicMin((long long*)&L.acc[a], (long long)res);
        }
        L.have[a] = 1;  // plain byte stores: racing writers all store 1
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (u32 u = 0; u < kPullRows; u++) {
    const u32 i = lane + 64 * u;
    if (i >= ncell) continue;
    const u32 sl = i >> lb, xl = i & (nl - 1);
    const u32 a = sl * 64 + xl;
    if (!L.have[a]) continue;
    const u64 cl = ((u64)c.row[sl] << lb) + xl;
    cell_apply_t<TROP>(cd, cl, TROP ? L.acc[a] : 0, 0, seen[u]);
  }
}


// Pairs per lane per step of the packed pull kernel (their loads are issued
// together before the LDS minimum).
#ifndef SK_PULL_PAIRS_PACKED
#define SK_PULL_PAIRS_PACKED 4
#endif
constexpr int kPullPairsP = SK_PULL_PAIRS_PACKED;
// LDS minima of one destination row are kPullAccStride words apart.
#ifndef SK_PULL_ACC_STRIDE
#define SK_PULL_ACC_STRIDE 64
#endif
constexpr u32 kPullAccStride = SK_PULL_ACC_STRIDE;


// Waves per SIMD the packed pull kernel is compiled for (its VGPR budget).
#ifndef SK_PULL_MIN_WAVES
#define SK_PULL_MIN_WAVES 1
#endif

// Per wave of the packed pull kernel (narrow Trop cells below 2^26).  One
// 16-byte record per index row of the batch, read with one LDS load per
// pair: {delta-list position of the source row's run | the LDS offset of
// the minima of the destination row it feeds << 48, exclusive prefix of the
// batch's pair counts | edge weight << 32}.
struct PullLdsP {
  ulonglong2 rec[64];
  u32 acc[kPullRows * kPullAccStride];  // running minimum per (destination row, lane value); kCellTop32 = none
  alignas(4) u8 start[64 * kPullPairsP];  // 1 + the index row whose pairs start at this position of the step
};
static_assert(kPullRows * kPullAccStride < (1u << 16), "minima offsets fit 16 bits");
static_assert(kPullPairsP == 4, "a lane's start marks are one 32-bit word");

// The pull round of k_pull<true> over a packed delta (CellDesc::dpack,
// rowrec8): a source row's delta run is one array of words
// (weight << 6) | lane at the record's position, its length the record's
// count, so a pair costs one 4-byte load and the row record 8 bytes.  Edge
// and delta weights are below 2^32, so their sum cannot overflow; a sum at or
// above narrow_lim raises the widen flag and is dropped, as a narrow cell
// drops it.  Pairs are dealt to lanes and folded into a u32 LDS minimum per
// (destination row, lane value) as in k_pull.
__global__ __launch_bounds__(kPullBlock, SK_PULL_MIN_WAVES) void k_pull_packed(PullDesc pd, CellDesc cd) {
  __shared__ PullLdsP sh[kPullBlock / 64];
  const u32 lane = threadIdx.x & 63;
  const u64 ch = (u64)blockIdx.x * (kPullBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ch >= pd.nchunks) return;  // whole waves leave; no block barrier below
  if (pd.gate && (pd.gate->done | pd.gate->leave | pd.gate->stamp)) return;  // a push round runs instead
  PullLdsP& L = sh[threadIdx.x >> 6];
  const PullChunk c = pd.chunks[ch];
  const u32 nr = c.nr_off & 0xFFu;
  const u64 lo = c.lo, hi = c.lo + c.n;
  const int lb = pd.lbits;
  const u32 nl = 1u << lb;
  const u32 ncell = nr << lb;  // <= kPullRows * 64
  // packed deltas come with narrow_lim = kPackLim; clamping keeps every sum
  // below 2^27 (a lower limit would only widen more often, which is exact)
  const u32 lim = (u32)min((u64)cd.narrow_lim, kPackLim);
  for (u32 i = lane; i < kPullRows * kPullAccStride; i += 64) L.acc[i] = kCellTop32;
  u64 seen[kPullRows];
#pragma unroll
  for (u32 u = 0; u < kPullRows; u++) {
    const u32 i = lane + 64 * u;
    seen[u] = 0;
    if (i < ncell) seen[u] = cell_peek(cd, ((u64)c.row[i >> lb] << lb) + (i & (nl - 1)));
  }
  bool wide = false;
  for (u64 e0 = lo; e0 < hi; e0 += 64) {
    const u64 e = e0 + lane;
    u32 cnt = 0, we = 0, slot = 0;
    u64 base = 0;
    if (e < hi) {
      const u32 q = pd.adj[e];
      // an edge weight outside [0, lim) (a lower stratum's relation is not
      // bounded by the staged weights) makes every candidate it forms reach
      // lim, i.e. widen, where the wide cells redo the round exactly
      const i64 wraw = pd.w[e];
      we = (wraw < 0 || (u64)wraw >= lim) ? lim : (u32)wraw;
      if (q < pd.nrows) {
        const u64 rec = pd.rowrec8[q];
        base = rec & ((1ull << 40) - 1);
        cnt = (u32)(rec >> 40);
      }
      const u32 el = (u32)(e - lo);
      for (u32 s = 1; s < nr; s++) slot += el >= ((c.nr_off >> (8 * s)) & 0xFFu) ? 1u : 0u;
    }
    const u32 incl = (u32)wave_incl_scan_j(cnt);
    const u32 total = __shfl(incl, 63);
    const u32 excl = incl - cnt;
    L.rec[lane] = make_ulonglong2(base | (u64)(slot * kPullAccStride) << 48, (u64)excl | (u64)we << 32);
    // the step's start marks: position u * 64 + l is byte u of word l, so a
    // lane reads and clears its four marks with one LDS access each
    u32* smark = reinterpret_cast<u32*>(L.start);
    smark[lane] = 0;
    u32 carry = 0;  // the edge holding the step's first pair (pair 0 starts an edge)
    for (u32 p0 = 0; p0 < total; p0 += 64 * kPullPairsP) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (cnt && excl >= p0 && excl < p0 + 64 * kPullPairsP) {
        const u32 pos = excl - p0;
        L.start[(pos & 63u) * 4 + (pos >> 6)] = (u8)(lane + 1);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const u32 mw = smark[lane];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      smark[lane] = 0;
      ulonglong2 rr[kPullPairsP];
      u32 vv[kPullPairsP];
#pragma unroll
      for (int u = 0; u < kPullPairsP; u++) {
        // the last index row started at or below this position (rows start
        // in index order), else the row the step's previous positions were in
        const u32 pm = wave_incl_max_u32((mw >> (8 * u)) & 0xFFu);
        const u32 j = pm ? pm - 1 : carry;
        carry = (u32)__builtin_amdgcn_readlane((int)j, 63);
        rr[u] = L.rec[j];
      }
#pragma unroll
      for (int u = 0; u < kPullPairsP; u++) {
        const u32 pidx = p0 + (u32)u * 64 + lane;
        vv[u] = kCellTop32;
        if (pidx < total)
          vv[u] = pd.dpack[(rr[u].x & ((1ull << 40) - 1)) + (pidx - (u32)rr[u].y)];
      }
      // lanes past the batch's pairs apply a no-op minimum (kCellTop32) at
      // an offset inside the row's minima
#pragma unroll
      for (int u = 0; u < kPullPairsP; u++) {
        const u32 cand = (vv[u] >> 6) + (u32)(rr[u].y >> 32);  // < 2^27
        const bool real = p0 + (u32)u * 64 + lane < total;
        const bool fits = cand < lim;
        wide |= real && !fits;
        atomicMin(&L.acc[(u32)(rr[u].x >> 48) + (vv[u] & 63u)], real && fits ? cand : kCellTop32);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (u32 u = 0; u < kPullRows; u++) {
    const u32 i = lane + 64 * u;
    if (i >= ncell) continue;
    const u32 sl = i >> lb, xl = i & (nl - 1);
    const u32 v = L.acc[sl * kPullAccStride + xl];
    if (v == kCellTop32) continue;
    const u64 cl = ((u64)c.row[sl] << lb) + xl;
    cell_apply_t<true>(cd, cl, (i64)v, 0, seen[u]);
  }
  if (__ballot(wide) && lane == 0) atomicOr(cd.widen, 1u);
}

// ============================================================ device-driven push rounds
//
// The push join of one delta variant with the delta's size and the join's
// output count on the device, for rounds captured in a graph: each grid is
// fixed and strides over the device-side count, and every kernel returns at
// once when the rounds are done or have left, or the round is not a push
// round (LoopState::stamp != 1).

constexpr u32 kDevGrid = 4096;      // blocks of the striding row and tile kernels

template <class K>
__global__ __launch_bounds__(kBlock) void k_scan_delta_dev(const K* __restrict__ keys, const i64* __restrict__ w,
                                                           const LoopState* __restrict__ ls, int arity, int bits,
                                                           ScanDesc sd, Table out) {
  if (ls->done | ls->leave | (ls->stamp ^ 1u)) return;  // push rounds only
  const u64 n = ls->n;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (u64)gridDim.x * kBlock) {
    const K key = keys[i];
    for (int c = 0; c < out.ncols; c++) out.col[c][i] = col_of<K>(key, arity, bits, sd.out_col[c]);
    if (out.w) out.w[i] = w[i];
  }
}

template <class K>
__global__ __launch_bounds__(kBlock) void k_join_count_dev(Table in, ProbeDesc p, const K* __restrict__ skeys,
                                                           const u64* __restrict__ csr, u64* __restrict__ lo_out,
                                                           u32* __restrict__ cnt, const LoopState* __restrict__ ls) {
  if (ls->done | ls->leave | (ls->stamp ^ 1u)) return;  // push rounds only
  const u64 n = ls->n;
  for (u64 r = (u64)blockIdx.x * kBlock + threadIdx.x; r < n; r += (u64)gridDim.x * kBlock)
    join_count_row<K>(in, p, skeys, csr, r, lo_out, cnt);
}

// Exclusive scan of cnt[0, ls->n) in three launches of fixed grids: block b
// sums its share, one block scans the shares (writing the total), block b
// rescans its share from its offset.
__device__ __forceinline__ u64 dscan_block_sum(u64 v, u64* tmp) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) tmp[threadIdx.x >> 6] = v;
  __syncthreads();
  u64 t = 0;
#pragma unroll
  for (int q = 0; q < kBlock / 64; q++) t += tmp[q];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(kBlock) void k_dscan_reduce(const u32* __restrict__ cnt, const LoopState* __restrict__ ls,
                                                         u64* __restrict__ bsum) {
  __shared__ u64 tmp[kBlock / 64];
  if (ls->done | ls->leave | (ls->stamp ^ 1u)) return;  // push rounds only
  const u64 n = ls->n, per = (n + gridDim.x - 1) / gridDim.x;
  const u64 a = (u64)blockIdx.x * per, e = a + per < n ? a + per : n;
  u64 v = 0;
  for (u64 i = a + threadIdx.x; i < e; i += kBlock) v += cnt[i];
  const u64 t = dscan_block_sum(v, tmp);
  if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_dscan_sums(u64* __restrict__ bsum, u32 nb, u64* __restrict__ total,
                                                       const LoopState* __restrict__ ls) {
  __shared__ u64 part[kBlock];
  if (ls->done | ls->leave | (ls->stamp ^ 1u)) return;  // push rounds only
  // thread t scans entries [t * per, (t + 1) * per) serially, then the block
  // scans the threads' sums
  const u32 per = (nb + kBlock - 1) / kBlock;
  const u32 a = threadIdx.x * per, e = a + per < nb ? a + per : nb;
  u64 sum = 0;
  for (u32 i = a; i < e; i++) sum += bsum[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (u32 o = 1; o < kBlock; o <<= 1) {
    const u64 y = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += y;
    __syncthreads();
  }
  u64 run = part[threadIdx.x] - sum;
  for (u32 i = a; i < e; i++) {
    const u64 v = bsum[i];
    bsum[i] = run;
    run += v;
  }
  if (threadIdx.x == kBlock - 1) *total = part[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_dscan_down(const u32* __restrict__ cnt, const LoopState* __restrict__ ls,
                                                       const u64* __restrict__ boff, u64* __restrict__ out) {
  __shared__ u64 tmp[kBlock / 64];
  if (ls->done | ls->leave | (ls->stamp ^ 1u)) return;  // push rounds only
  const u64 n = ls->n, per = (n + gridDim.x - 1) / gridDim.x;
  const u64 a = (u64)blockIdx.x * per, e = a + per < n ? a + per : n;
  u64 carry = boff[blockIdx.x];
  const u32 lane = threadIdx.x & 63;
  for (u64 t0 = a; t0 < e; t0 += kBlock) {
    const u64 i = t0 + threadIdx.x;
    const u64 v = i < e ? cnt[i] : 0;
    u64 inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const u64 y = __shfl_up(inc, o, 64);
      if (lane >= (u32)o) inc += y;
    }
    if (lane == 63) tmp[threadIdx.x >> 6] = inc;
    __syncthreads();
    u64 wbase = 0, tile = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) {
      if (q < (int)(threadIdx.x >> 6)) wbase += tmp[q];
      tile += tmp[q];
    }
    if (i < e) out[i] = carry + wbase + inc - v;
    carry += tile;
    __syncthreads();
  }
}

template <class K, bool TROP, bool PDOM>
__global__ __launch_bounds__(kBlock) void k_join_apply_dev(Table in, ProbeDesc p, OutDesc o, const K* __restrict__ skeys,
                                                           const i64* __restrict__ sw, const u64* __restrict__ lo,
                                                           const u64* __restrict__ off,
                                                           const u64* __restrict__ total_dev,
                                                           const LoopState* __restrict__ ls, DevErr* err) {
  if (ls->done | ls->leave | (ls->stamp ^ 1u)) return;  // push rounds only
  const u64 total = *total_dev, nrows = ls->n;
  for (u64 t = blockIdx.x; t * kEmitTile < total; t += gridDim.x) {
    const u64 o0 = t * kEmitTile;
    join_apply_tile<K, TROP, PDOM>(in, p, o, skeys, sw, nullptr, lo, off, o0, min(o0 + kEmitTile, total), nrows, err);
    __syncthreads();  // the tile's LDS row offsets are rewritten by the next tile
  }
}

template <class KD, class KE>
void push_round_dev(const Exec& x, const KD* dkeys, const i64* dw, int darity, int dbits, const ScanDesc& sd,
                    const Table& t, const ProbeDesc& p, const KE* skeys, const i64* sw, const u64* csr,
                    const OutDesc& o, u64* lo, u32* cnt, u64* off, u64* bsum, u64* jtot, u64 cap,
                    const LoopState* ls) {
  if (!o.cd.on || !o.cd.trop || p.nfilt || !t.w || t.inf) throw std::runtime_error("push_round_dev: shape");
  const unsigned rg = (unsigned)std::min<u64>(gridj(cap), kDevGrid);
  k_scan_delta_dev<KD><<<rg, kBlock, 0, x.stream>>>(dkeys, dw, ls, darity, dbits, sd, t);
  SK_HIP(hipGetLastError());
  k_join_count_dev<KE><<<rg, kBlock, 0, x.stream>>>(t, p, skeys, csr, lo, cnt, ls);
  SK_HIP(hipGetLastError());
  k_dscan_reduce<<<kDevScanBlocks, kBlock, 0, x.stream>>>(cnt, ls, bsum);
  SK_HIP(hipGetLastError());
  k_dscan_sums<<<1, kBlock, 0, x.stream>>>(bsum, kDevScanBlocks, jtot, ls);
  SK_HIP(hipGetLastError());
  k_dscan_down<<<kDevScanBlocks, kBlock, 0, x.stream>>>(cnt, ls, bsum, off);
  SK_HIP(hipGetLastError());
  bool pdom = false;
  for (int j = 0; j < o.cd.ncols; j++) pdom = pdom || o.cd.pdom_off[j] != nullptr;
  if (pdom)
    k_join_apply_dev<KE, true, true><<<kDevGrid, kBlock, 0, x.stream>>>(t, p, o, skeys, sw, lo, off, jtot, ls, x.err);
  else
    k_join_apply_dev<KE, true, false><<<kDevGrid, kBlock, 0, x.stream>>>(t, p, o, skeys, sw, lo, off, jtot, ls, x.err);
  SK_HIP(hipGetLastError());
}

template void push_round_dev<u64, u64>(const Exec&, const u64*, const i64*, int, int, const ScanDesc&, const Table&,
                                       const ProbeDesc&, const u64*, const i64*, const u64*, const OutDesc&, u64*,
                                       u32*, u64*, u64*, u64*, u64, const LoopState*);
template void push_round_dev<u128, u64>(const Exec&, const u128*, const i64*, int, int, const ScanDesc&, const Table&,
                                        const ProbeDesc&, const u64*, const i64*, const u64*, const OutDesc&, u64*,
                                        u32*, u64*, u64*, u64*, u64, const LoopState*);
template void push_round_dev<u64, u128>(const Exec&, const u64*, const i64*, int, int, const ScanDesc&, const Table&,
                                        const ProbeDesc&, const u128*, const i64*, const u64*, const OutDesc&, u64*,
                                        u32*, u64*, u64*, u64*, u64, const LoopState*);
template void push_round_dev<u128, u128>(const Exec&, const u128*, const i64*, int, int, const ScanDesc&,
                                         const Table&, const ProbeDesc&, const u128*, const i64*, const u64*,
                                         const OutDesc&, u64*, u32*, u64*, u64*, u64*, u64, const LoopState*);

void pull_apply(const Exec& x, const PullDesc& p, const CellDesc& cd, bool trop, double bytes, u64 items) {
  if (!p.nchunks) return;
  Scope sc(x.timer, "join_pull", KC_JOIN, bytes, items);
  const u64 waves_per_block = kPullBlock / 64;
  const unsigned nb = (unsigned)((p.nchunks + waves_per_block - 1) / waves_per_block);
  if (trop && p.dpack) k_pull_packed<<<nb, kPullBlock, 0, x.stream>>>(p, cd);
  else if (trop) k_pull<true><<<nb, kPullBlock, 0, x.stream>>>(p, cd, x.err);
  else k_pull<false><<<nb, kPullBlock, 0, x.stream>>>(p, cd, x.err);
  SK_HIP(hipGetLastError());
}

u64 pull_chunks(const Exec& x, const u64* csr, u64 nrows, DBuf<PullChunk>* out, u64* rows_used) {
  std::vector<u64> h(nrows + 1);
  SK_HIP(hipMemcpyAsync(h.data(), csr, (nrows + 1) * sizeof(u64), hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  if (rows_used) {
    *rows_used = 0;
    for (u64 r = 0; r < nrows; r++) *rows_used += h[r + 1] > h[r];
  }
  std::vector<PullChunk> v;
  PullChunk cur{};
  bool open = false;
  auto flush = [&] {
    if (open) v.push_back(cur);
    open = false;
  };
  for (u64 r = 0; r < nrows; r++) {
    const u64 a = h[r], d = h[r + 1] - a;
    if (!d) continue;
    if (d > kPullPack) {
      flush();
      for (u64 s = a; s < a + d; s += 256) {
        PullChunk c{};
        c.lo = s;
        c.n = (u32)std::min<u64>(256, a + d - s);
        c.nr_off = 1;
        c.row[0] = (u32)r;
        v.push_back(c);
      }
      continue;
    }
    if (open && (cur.n + d > kPullPack || (cur.nr_off & 0xFFu) == kPullRows)) flush();
    if (!open) {
      cur = PullChunk{};
      cur.lo = a;
      open = true;
    }
    const u32 s = cur.nr_off & 0xFFu;
    if (s) cur.nr_off |= cur.n << (8 * s);  // cur.n < kPullPack fits a byte
    cur.row[s] = (u32)r;
    cur.nr_off = (cur.nr_off & ~0xFFu) | (s + 1);
    cur.n += (u32)d;
  }
  flush();
  const u64 n = v.size();
  out->reset(x.pool, n ? n : 1);
  if (n) SK_HIP(hipMemcpyAsync(out->get(), v.data(), n * sizeof(PullChunk), hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  return n;
}

// ============================================================ bit pull
//
// Bool rules of the pull shape over rows of any width: a destination row is
// the OR of the delta rows its row-graph entries name, 32 lanes per word
// operation and no atomic per lane.  Per wave and batch of 64 graph
// entries, lane l loads entry l's source row and its stamp; a ballot keeps
// the rows holding delta cells, which are compacted into LDS, and the wave
// then ORs those rows' delta words.  Rows of W >= 64 words go word block by
// word block (one word per lane); narrower rows split the wave into 64 / W
// slots that take the active rows in turn, folded by shuffles at the end.
// The result is applied with one atomic OR per word that gains bits; the
// bits it gained (and only those) are marked touched, exactly as the push
// join marks each new cell.

struct BitGraphDesc {
  int arity = 0, bits = 0, hcols = 0, lbits = 0;
  int eh[kMaxCols] = {}, ed[kMaxCols] = {};
  u64 nrows = 0;
};

template <class K>
__global__ void k_bitpull_graph(const K* __restrict__ ek, u64 n, BitGraphDesc g, CellDesc cd,
                                u64* __restrict__ out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const K key = ek[i];
  u64 d = 0, s = 0;
  bool ok = true;
  for (int c = 0; c < g.hcols; c++) {
    if (g.eh[c] < 0) continue;  // the lane column
    u64 rd = 0, rs = 0;
    const u32 ph = col_of<K>(key, g.arity, g.bits, g.eh[0] < 0 ? 0 : g.eh[0]);
    const u32 pd = col_of<K>(key, g.arity, g.bits, g.ed[0] < 0 ? 0 : g.ed[0]);
    ok = ok && cell_rank(cd, c, col_of<K>(key, g.arity, g.bits, g.eh[c]), ph, &rd) && !(rd >> cd.ebits[c]);
    ok = ok && cell_rank(cd, c, col_of<K>(key, g.arity, g.bits, g.ed[c]), pd, &rs) && !(rs >> cd.ebits[c]);
    d |= rd << (cd.shift[c] - g.lbits);
    s |= rs << (cd.shift[c] - g.lbits);
  }
  out[i] = ok ? ((d << 32) | s) : (g.nrows << 32);
}

__device__ __forceinline__ void bit_apply(const CellDesc& cd, u64 wi, u32 acc) {
  if (!acc) return;
  const u32 nw = acc & ~cd.present[wi];
  if (!nw) return;
  const u32 old = atomicOr(&cd.present[wi], nw);
  const u32 nb = nw & ~old;
  if (nb) atomicOr(&cd.touched[wi], nb);
}

// Loads entry e0 + lane's source row; returns the number of entries of the
// batch whose source row holds delta cells, their rows compacted in `act`.
__device__ __forceinline__ u32 bitpull_batch(const BitPullDesc& bd, u32 stamp, u64 e0, u64 e_end, u32 lane,
                                             u32* act) {
  const u64 e = e0 + lane;
  u32 s = 0;
  bool a = false;
  if (e < e_end) {
    s = (u32)bd.graph[e];
    a = bd.rstamp[s] == stamp;
  }
  const u64 m = __ballot(a);
  if (a) act[__popcll(m & ((1ull << lane) - 1ull))] = s;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return (u32)__popcll(m);
}

__global__ __launch_bounds__(kPullBlock) void k_bitpull(BitPullDesc bd, CellDesc cd) {
  __shared__ u32 s_act[kPullBlock];
  const u32 lane = threadIdx.x & 63;
  const u32 wv = threadIdx.x >> 6;
  const u64 ch = (u64)blockIdx.x * (kPullBlock / 64) + __builtin_amdgcn_readfirstlane(wv);
  if (ch >= bd.nchunks) return;  // whole waves leave; no block barrier below
  if (bd.loop && bd.loop->done) return;
  const u32 stamp = bd.loop ? bd.loop->stamp : bd.stamp;
  u32* act = s_act + wv * 64;
  const BitChunk c = bd.chunks[ch];
  const u64 e_end = c.lo + c.n;
  const u32 W = 1u << bd.wshift;
  if (W >= 64) {
    for (u32 wb = 0; wb < W; wb += 64) {
      u32 acc = 0;
      for (u64 e0 = c.lo; e0 < e_end; e0 += 64) {
        const u32 na = bitpull_batch(bd, stamp, e0, e_end, lane, act);
        u32 k = 0;
        for (; k + 4 <= na; k += 4) {
          const u32 v0 = bd.dbits[((u64)act[k] << bd.wshift) + wb + lane];
          const u32 v1 = bd.dbits[((u64)act[k + 1] << bd.wshift) + wb + lane];
          const u32 v2 = bd.dbits[((u64)act[k + 2] << bd.wshift) + wb + lane];
          const u32 v3 = bd.dbits[((u64)act[k + 3] << bd.wshift) + wb + lane];
          acc |= v0 | v1 | v2 | v3;
        }
        for (; k < na; k++) acc |= bd.dbits[((u64)act[k] << bd.wshift) + wb + lane];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      bit_apply(cd, ((u64)c.row << bd.wshift) + wb + lane, acc);
    }
  } else {
    const u32 slot = lane >> bd.wshift, word = lane & (W - 1), nslots = 64u >> bd.wshift;
    u32 acc = 0;
    for (u64 e0 = c.lo; e0 < e_end; e0 += 64) {
      const u32 na = bitpull_batch(bd, stamp, e0, e_end, lane, act);
      for (u32 k = slot; k < na; k += nslots) acc |= bd.dbits[((u64)act[k] << bd.wshift) + word];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    for (u32 o = W; o < 64; o <<= 1) acc |= __shfl_xor(acc, o, 64);
    if (lane < W) bit_apply(cd, ((u64)c.row << bd.wshift) + word, acc);
  }
}

static int ceil_log2_j(u64 n) {
  int b = 0;
  while ((1ull << b) < n) b++;
  return b;
}

template <class K>
u64 bitpull_graph(const Exec& x, const K* ekeys, u64 n, int arity, int bits, const CellDesc& cd, int hcols,
                  const int* eh, const int* ed, int lbits, u64 nrows, DBuf<u64>* graph, DBuf<u64>* csr) {
  BitGraphDesc g;
  g.arity = arity;
  g.bits = bits;
  g.hcols = hcols;
  g.lbits = lbits;
  g.nrows = nrows;
  for (int c = 0; c < hcols; c++) {
    g.eh[c] = eh[c];
    g.ed[c] = ed[c];
  }
  DBuf<u64> a(x.pool, n ? n : 1), b(x.pool, n ? n : 1);
  u64* sorted = a.get();
  if (n) {
    {
      Scope sc(x.timer, "bitpull_graph", KC_JOIN, (double)n * (sizeof(K) + 8.0), n);
      k_bitpull_graph<K><<<gridj(n), kBlock, 0, x.stream>>>(ekeys, n, g, cd, a.get());
      SK_HIP(hipGetLastError());
    }
    radix_sort<u64>(x, a.get(), b.get(), nullptr, nullptr, nullptr, nullptr, n, 32 + ceil_log2_j(nrows + 1),
                    &sorted, nullptr, nullptr);
  }
  csr->reset(x.pool, nrows + 1);
  build_csr<u64>(x, sorted, n, 32, nrows, csr->get());
  const u64 valid = read_u64(x, csr->get() + nrows);
  *graph = (sorted == a.get()) ? std::move(a) : std::move(b);
  return valid;
}

u64 bitpull_chunks(const Exec& x, const u64* csr, u64 nrows, DBuf<BitChunk>* out) {
  std::vector<u64> h(nrows + 1);
  SK_HIP(hipMemcpyAsync(h.data(), csr, (nrows + 1) * sizeof(u64), hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  std::vector<BitChunk> v;
  for (u64 r = 0; r < nrows; r++)
    for (u64 s = h[r]; s < h[r + 1]; s += 256) v.push_back(BitChunk{s, (u32)std::min<u64>(256, h[r + 1] - s), (u32)r});
  const u64 n = v.size();
  out->reset(x.pool, n ? n : 1);
  if (n) SK_HIP(hipMemcpyAsync(out->get(), v.data(), n * sizeof(BitChunk), hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  return n;
}

void bitpull_apply(const Exec& x, const BitPullDesc& p, const CellDesc& cd, double bytes, u64 items) {
  if (!p.nchunks) return;
  Scope sc(x.timer, "join_bitpull", KC_JOIN, bytes, items);
  const unsigned nb = (unsigned)((p.nchunks + kPullBlock / 64 - 1) / (kPullBlock / 64));
  k_bitpull<<<nb, kPullBlock, 0, x.stream>>>(p, cd);
  SK_HIP(hipGetLastError());
}

// Grid-stride sum; one atomic per block (a single counter hit once per wave
// serialises the whole launch).
template <class K>
__global__ __launch_bounds__(kBlock) void k_push_cost(const K* __restrict__ dkeys, u64 n, int arity, int bits, int col,
                                                      const u64* __restrict__ csr,
                                                      unsigned long long* __restrict__ acc) {
  __shared__ u64 part[kBlock / 64];
  u64 s = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const u32 v = col_of<K>(dkeys[i], arity, bits, col);
    s += csr[v + 1] - csr[v];
  }
  s = wave_incl_scan_j(s);
  if (lane_id_j() == 63) part[wave_id_j()] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
    for (int w = 0; w < kBlock / 64; w++) t += part[w];
    if (t) atomicAdd(acc, (unsigned long long)t);
  }
}

template <class K>
u64 push_cost(const Exec& x, const K* dkeys, u64 n, int arity, int bits, int col, const u64* csr) {
  if (!n) return 0;
  DBuf<u64> acc(x.pool, 1);
  memset_async(x, acc.get(), 0, 8);
  {
    Scope sc(x.timer, "push_cost", KC_MISC, (double)n * (sizeof(K) + 16.0), n);
    const unsigned nb = (unsigned)std::min<u64>((n + kBlock - 1) / kBlock, 2048);
    k_push_cost<K><<<nb, kBlock, 0, x.stream>>>(dkeys, n, arity, bits, col, csr, (unsigned long long*)acc.get());
    SK_HIP(hipGetLastError());
  }
  return read_u64(x, acc.get());
}

// ================================================================ remap

template <class K, class KO>
__global__ void k_remap(const K* __restrict__ in, u64 n, int arity, int bits_in, int bits_out,
                        const u32* __restrict__ map, KO* __restrict__ out) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = in[i];
  KO o = 0;
  for (int j = 0; j < arity; j++) o = (o << bits_out) | (KO)map[col_of<K>(k, arity, bits_in, j)];
  out[i] = o;
}

template <class K, class KO>
void remap_keys(const Exec& x, const K* in, u64 n, int arity, int bits_in, int bits_out, const u32* map, KO* out) {
  if (!n) return;
  k_remap<K, KO><<<gridj(n), kBlock, 0, x.stream>>>(in, n, arity, bits_in, bits_out, map, out);
  SK_HIP(hipGetLastError());
}

// ================================================================ encode

// Pack staged u32 code columns into keys.
template <class K>
__global__ void k_pack(Cols cols, int arity, int bits, u64 n, K* __restrict__ out) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = 0;
  for (int j = 0; j < arity; j++) k = (k << bits) | (K)cols.c[j][i];
  out[i] = k;
}

template <class K>
void pack_cols(const Exec& x, const Cols& cols, int arity, int bits, u64 n, K* out) {
  if (!n) return;
  k_pack<K><<<gridj(n), kBlock, 0, x.stream>>>(cols, arity, bits, n, out);
  SK_HIP(hipGetLastError());
}

// Mark codes present (universe bitmap).
__global__ void k_mark(const u32* __restrict__ c, u64 n, u32 limit, u32* __restrict__ bitmap) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u32 v = i < n ? c[i] : 0u;
  mark_bit_once(bitmap, v, i < n && v < limit);
}

__global__ __launch_bounds__(kBlock) void k_popcount(const u32* __restrict__ b, u64 n, unsigned long long* acc) {
  __shared__ u64 part[kBlock / 64];
  u64 s = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    s += (u64)__popc(b[i]);
  s = wave_incl_scan_j(s);
  if (lane_id_j() == 63) part[wave_id_j()] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
    for (int w = 0; w < kBlock / 64; w++) t += part[w];
    if (t) atomicAdd(acc, (unsigned long long)t);
  }
}

void mark_codes(const Exec& x, const u32* c, u64 n, u32 limit, u32* bitmap) {
  if (!n) return;
  k_mark<<<gridj(n), kBlock, 0, x.stream>>>(c, n, limit, bitmap);
  SK_HIP(hipGetLastError());
}

u64 popcount_bits(const Exec& x, const u32* bitmap, u64 nwords) {
  DBuf<u64> acc(x.pool, 1);
  memset_async(x, acc.get(), 0, 8);
  if (nwords)
    k_popcount<<<(unsigned)std::min<u64>(gridj(nwords), 2048), kBlock, 0, x.stream>>>(bitmap, nwords,
                                                                                (unsigned long long*)acc.get());
  SK_HIP(hipGetLastError());
  return read_u64(x, acc.get());
}

// ============================================================ instances

#define SK_JOIN_INST(K)                                                                                          \
  template u64 scan_rel<K>(const Exec&, const K*, const i64*, const u8*, u64, const ScanDesc&, Table*, TableAlloc, \
                           void*);                                                                               \
  template u64 join<K>(const Exec&, const Table&, const ProbeDesc&, const K*, const i64*, const u8*, const u64*, \
                       OutDesc*, bool, bool, bool, TableAlloc, void*, HeadAlloc, void*);                         \
  template u64 anti_join<K>(const Exec&, const Table&, const AntiDesc&, const K*, u64, Table*, TableAlloc, void*); \
  template u64 aggregate<K>(const Exec&, const Table&, const AggDesc&, const i64*, int, std::vector<K>*,         \
                            std::vector<i64>*, bool*);                                                          \
  template u64 push_cost<K>(const Exec&, const K*, u64, int, int, int, const u64*);                              \
  template u64 triangles<K>(const Exec&, const K*, u64, const TriDesc&, const u64*, const u32*, const u64*,       \
                            const u32*, OutDesc*, bool, bool, HeadAlloc, void*, double*);                        \
  template double tri_list_len<K>(const Exec&, const K*, u64, const TriDesc&, const u64*, const u64*);           \
  template void pack_cols<K>(const Exec&, const Cols&, int, int, u64, K*);                                       \
  template u64 bitpull_graph<K>(const Exec&, const K*, u64, int, int, const CellDesc&, int, const int*, const int*, \
                                int, u64, DBuf<u64>*, DBuf<u64>*);

SK_JOIN_INST(u64)
SK_JOIN_INST(u128)
template void remap_keys<u64, u64>(const Exec&, const u64*, u64, int, int, int, const u32*, u64*);
template void remap_keys<u64, u128>(const Exec&, const u64*, u64, int, int, int, const u32*, u128*);
template void remap_keys<u128, u64>(const Exec&, const u128*, u64, int, int, int, const u32*, u64*);
template void remap_keys<u128, u128>(const Exec&, const u128*, u64, int, int, int, const u32*, u128*);

// Interned id = position of each key in the sorted distinct table.
__global__ void k_intern(const u64* __restrict__ t, u64 nt, const u64* __restrict__ q, u64 nq, u32* __restrict__ ids) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  u64 k = q[i], lo = 0, hi = nt;
  while (lo < hi) {
    u64 m = (lo + hi) >> 1;
    if (t[m] < k) lo = m + 1;
    else hi = m;
  }
  ids[i] = (u32)lo;
}

void intern_lookup(const Exec& x, const u64* table, u64 nt, const u64* q, u64 nq, u32* ids) {
  if (!nq) return;
  k_intern<<<gridj(nq), kBlock, 0, x.stream>>>(table, nt, q, nq, ids);
  SK_HIP(hipGetLastError());
}

// ====================================================== partition filter

template <class K>
__global__ void k_part_flags(const K* __restrict__ keys, u64 n, int arity, int bits, const u8* __restrict__ dead,
                             u8* __restrict__ keep) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keep[i] = dead[col_of<K>(keys[i], arity, bits, 0)] ? 0 : 1;
}

template <class K>
__global__ void k_part_write(const K* __restrict__ keys, const i64* __restrict__ w, const u8* __restrict__ f, u64 n,
                             const u8* __restrict__ keep, const u64* __restrict__ pos, K* __restrict__ ok,
                             i64* __restrict__ ow, u8* __restrict__ of) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  u64 o = pos[i];
  ok[o] = keys[i];
  if (w) ow[o] = w[i];
  if (f) of[o] = f[i];
}

template <class K>
u64 drop_partitions(const Exec& x, const K* keys, const i64* w, const u8* inf, u64 n, int arity, int bits,
                    const u8* dead, K* ok, i64* ow, u8* of) {
  if (!n) return 0;
  DBuf<u8> keep(x.pool, n);
  DBuf<u64> pos(x.pool, n);
  DBuf<u64> tot(x.pool, 1);
  k_part_flags<K><<<gridj(n), kBlock, 0, x.stream>>>(keys, n, arity, bits, dead, keep.get());
  scan_u8(x, keep.get(), pos.get(), n, tot.get());
  k_part_write<K><<<gridj(n), kBlock, 0, x.stream>>>(keys, w, inf, n, keep.get(), pos.get(), ok, ow, of);
  SK_HIP(hipGetLastError());
  return read_u64(x, tot.get());
}

template u64 drop_partitions<u64>(const Exec&, const u64*, const i64*, const u8*, u64, int, int, const u8*, u64*,
                                  i64*, u8*);
template u64 drop_partitions<u128>(const Exec&, const u128*, const i64*, const u8*, u64, int, int, const u8*, u128*,
                                   i64*, u8*);

}  // namespace sk
