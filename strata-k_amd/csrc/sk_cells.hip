// sk_cells.hip — dense cell state kernels: column domains, loading a
// relation into cells, ⊕-applying derivations, and compacting touched or
// present cells back into sorted packed relations.
#include <climits>
#include <cstring>

#include <algorithm>

#include "sk_cells.h"

namespace sk {

static unsigned gridc(u64 n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class K>
__device__ __forceinline__ u32 code_at(K key, int arity, int bits, int j) {
  K mask = (((K)1) << bits) - 1;
  return (u32)((key >> (bits * (arity - 1 - j))) & mask);
}

// ----------------------------------------------------------------- domains

template <class K>
__global__ void k_cells_mark_col(const K* __restrict__ keys, u64 n, int arity, int bits, int j, u32* bitmap) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u32 v = i < n ? code_at<K>(keys[i], arity, bits, j) : 0u;
  mark_bit_once(bitmap, v, i < n);
}

// Codes below 2^kMarkLdsBits: each block marks a private LDS bitmap over a
// grid-stride share of the rows and ORs its non-zero words into the global
// bitmap once (a column of few distinct codes otherwise sends every row's
// atomic to the same few words).
constexpr int kMarkLdsBits = 18;
template <class K>
__global__ __launch_bounds__(kBlock) void k_cells_mark_col_lds(const K* __restrict__ keys, u64 n, int arity, int bits,
                                                               int j, u32* bitmap) {
  __shared__ u32 sb[(1u << kMarkLdsBits) / 32];
  const u32 nw = (1u << bits) / 32 ? (1u << bits) / 32 : 1u;
  for (u32 i = threadIdx.x; i < nw; i += kBlock) sb[i] = 0;
  __syncthreads();
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (u64)gridDim.x * kBlock) {
    const u32 v = code_at<K>(keys[i], arity, bits, j);
    const u32 m = 1u << (v & 31);
    if (!(sb[v >> 5] & m)) atomicOr(&sb[v >> 5], m);
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < nw; i += kBlock) {
    const u32 w = sb[i];
    if (w && (bitmap[i] & w) != w) atomicOr(&bitmap[i], w);
  }
}

__global__ void k_cells_mark_code(u32* bitmap, u32 code) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicOr(&bitmap[code >> 5], 1u << (code & 31));
}

__global__ void k_cells_or(u32* __restrict__ dst, const u32* __restrict__ src, u64 nwords) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) dst[i] |= src[i];
}

__global__ void k_cells_popc(const u32* __restrict__ b, u64 nwords, u32* __restrict__ cnt) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) cnt[i] = (u32)__popc(b[i]);
}

__global__ void k_cells_list(const u32* __restrict__ b, const u64* __restrict__ pre, u64 nwords,
                             u32* __restrict__ list) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  u32 w = b[i];
  u64 o = pre[i];
  while (w) {
    int bit = __ffs(w) - 1;
    list[o++] = (u32)(i * 32 + bit);
    w &= w - 1;
  }
}

template <class K>
void cells_mark_col(const Exec& x, const K* keys, u64 n, int arity, int bits, int j, u32* bitmap) {
  if (!n) return;
  Scope sc(x.timer, "cells_domain", KC_MISC, (double)n * sizeof(K), n);
  if (bits <= kMarkLdsBits) {
    // enough blocks to fill the chip, each over at least 16 rows per thread
    const u64 nb = std::min<u64>(1024, std::max<u64>(1, n / (kBlock * 16)));
    k_cells_mark_col_lds<K><<<(unsigned)nb, kBlock, 0, x.stream>>>(keys, n, arity, bits, j, bitmap);
  } else {
    k_cells_mark_col<K><<<gridc(n), kBlock, 0, x.stream>>>(keys, n, arity, bits, j, bitmap);
  }
  SK_HIP(hipGetLastError());
}

void cells_mark_code(const Exec& x, u32* bitmap, u32 code) {
  k_cells_mark_code<<<1, 64, 0, x.stream>>>(bitmap, code);
  SK_HIP(hipGetLastError());
}

void cells_or(const Exec& x, u32* dst, const u32* src, u64 nwords) {
  if (!nwords) return;
  k_cells_or<<<gridc(nwords), kBlock, 0, x.stream>>>(dst, src, nwords);
  SK_HIP(hipGetLastError());
}

u64 cells_domain(const Exec& x, const u32* bitmap, u64 nwords, u64* pre, DBuf<u32>* list) {
  DBuf<u32> cnt(x.pool, nwords);
  DBuf<u64> tot(x.pool, 1);
  k_cells_popc<<<gridc(nwords), kBlock, 0, x.stream>>>(bitmap, nwords, cnt.get());
  SK_HIP(hipGetLastError());
  scan_u32(x, cnt.get(), pre, nwords, tot.get());
  u64 n = read_u64(x, tot.get());
  list->reset(x.pool, n ? n : 1);
  k_cells_list<<<gridc(nwords), kBlock, 0, x.stream>>>(bitmap, pre, nwords, list->get());
  SK_HIP(hipGetLastError());
  return n;
}

template <class K>
__global__ void k_part_pairs(const K* __restrict__ keys, u64 n, int arity, int bits, int col, u64* __restrict__ out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const K k = keys[i];
  out[i] = ((u64)code_at<K>(k, arity, bits, 0) << 32) | code_at<K>(k, arity, bits, col);
}

__global__ void k_low32(const u64* __restrict__ in, u64 n, u32* __restrict__ out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (u32)in[i];
}

u64 cells_part_domain(const Exec& x, const std::vector<PartCol>& parts, u64 nparts, DBuf<u64>* off, DBuf<u32>* codes) {
  u64 total = 0;
  for (auto& p : parts) total += p.n;
  DBuf<u64> a(x.pool, total ? total : 1), b(x.pool, total ? total : 1), u(x.pool, total ? total : 1);
  u64 at = 0;
  for (auto& p : parts) {
    if (!p.n) continue;
    if (p.wide)
      k_part_pairs<u128><<<gridc(p.n), kBlock, 0, x.stream>>>((const u128*)p.keys, p.n, p.arity, p.bits, p.col,
                                                              a.get() + at);
    else
      k_part_pairs<u64><<<gridc(p.n), kBlock, 0, x.stream>>>((const u64*)p.keys, p.n, p.arity, p.bits, p.col,
                                                             a.get() + at);
    SK_HIP(hipGetLastError());
    at += p.n;
  }
  int pb = 0;
  while ((1ull << pb) < nparts) pb++;
  u64* sorted = a.get();
  u64 nu = 0;
  if (total) {
    radix_sort<u64>(x, a.get(), b.get(), nullptr, nullptr, nullptr, nullptr, total, 32 + pb, &sorted, nullptr, nullptr);
    nu = unique_runs<u64>(x, sorted, nullptr, nullptr, total, u.get(), nullptr, nullptr);
  }
  off->reset(x.pool, nparts + 1);
  build_csr<u64>(x, u.get(), nu, 32, nparts, off->get());
  codes->reset(x.pool, nu ? nu : 1);
  if (nu) {
    k_low32<<<gridc(nu), kBlock, 0, x.stream>>>(u.get(), nu, codes->get());
    SK_HIP(hipGetLastError());
  }
  std::vector<u64> h(nparts + 1);
  SK_HIP(hipMemcpyAsync(h.data(), off->get(), (nparts + 1) * 8, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  u64 mx = 0;
  for (u64 p = 0; p < nparts; p++) mx = std::max<u64>(mx, h[p + 1] - h[p]);
  return mx;
}

// ------------------------------------------------------ load / ⊕-apply

template <class K>
__global__ void k_cells_load(CellDesc cd, const K* __restrict__ keys, const i64* __restrict__ w,
                             const u8* __restrict__ inf, u64 n, int arity, int bits) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K key = keys[i];
  u64 c;
  if (!cell_of(cd, [&](int j) { return code_at<K>(key, arity, bits, j); }, &c)) return;
  u32 m = 1u << (c & 31);
  if (!cd.trop) {
    atomicOr(&cd.present[c >> 5], m);
  } else if (inf && inf[i]) {
    atomicOr(&cd.infb[c >> 5], m);
  } else if (cd.narrow) {
    // keys are unique: one writer per cell
    if ((u64)w[i] >= (u64)kCellTop32) atomicOr(cd.widen, 1u);
    else cd.cell32[c] = (u32)w[i];
  } else if (w[i] == LLONG_MAX) {
    atomicOr(&cd.maxf[c >> 5], m);
  } else {
    cd.cell[c] = cell_enc(w[i]);
  }
}

template <class K>
__global__ void k_cells_insert(CellDesc cd, const K* __restrict__ keys, const i64* __restrict__ w,
                               const u8* __restrict__ inf, u64 n, int arity, int bits) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K key = keys[i];
  u64 c;
  if (!cell_of(cd, [&](int j) { return code_at<K>(key, arity, bits, j); }, &c)) return;
  cell_insert(cd, c, cd.trop ? w[i] : 0, (cd.trop && inf) ? inf[i] : (u8)0);
}

template <class K>
void cells_load(const Exec& x, const CellDesc& cd, const K* keys, const i64* w, const u8* inf, u64 n, int arity,
                int bits) {
  if (!n) return;
  Scope sc(x.timer, "cells_load", KC_MERGE, (double)n * (sizeof(K) + (cd.trop ? 16.0 : 4.0)), n);
  k_cells_load<K><<<gridc(n), kBlock, 0, x.stream>>>(cd, keys, w, inf, n, arity, bits);
  SK_HIP(hipGetLastError());
}

template <class K>
void cells_insert(const Exec& x, const CellDesc& cd, const K* keys, const i64* w, const u8* inf, u64 n, int arity,
                  int bits) {
  if (!n) return;
  Scope sc(x.timer, "cells_insert", KC_MERGE, (double)n * (sizeof(K) + (cd.trop ? 16.0 : 4.0)), n);
  k_cells_insert<K><<<gridc(n), kBlock, 0, x.stream>>>(cd, keys, w, inf, n, arity, bits);
  SK_HIP(hipGetLastError());
}

// ----------------------------------------------------------------- compact
//
// One block per kBlock bitmap words (kBlock * 32 cells).  Pass 1 counts the
// selected cells per block; after a scan over blocks, pass 2 scans the
// per-word counts inside the block, stages the selected cells' in-block
// indices in LDS (in cell order) and writes the packed rows coalesced.

constexpr u32 kCellsPerBlock = (u32)kBlock * 32;

__device__ __forceinline__ u32 cells_word(const CellDesc& cd, u64 wi, bool touched_only) {
  if (touched_only) return cd.touched[wi];
  if (!cd.trop) return cd.present[wi];
  u32 b = cd.infb ? cd.infb[wi] : 0u;
  if (cd.narrow) {
    const u32* p = cd.cell32 + wi * 32;
#pragma unroll 8
    for (int k = 0; k < 32; k++)
      if (p[k] != kCellTop32) b |= 1u << k;
    return b;
  }
  b |= cd.maxf[wi];
  const u64* p = cd.cell + wi * 32;
#pragma unroll 8
  for (int k = 0; k < 32; k++)
    if (p[k] != kCellTop) b |= 1u << k;
  return b;
}

// Exclusive block scan of one u32 per thread (kBlock = 4 waves of 64).
__device__ __forceinline__ u32 cells_block_scan(u32 v, u32* tmp, u32* total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u32 inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u32 y = __shfl_up(inc, o, 64);
    if (lane >= (u32)o) inc += y;
  }
  if (lane == 63) tmp[wave] = inc;
  __syncthreads();
  u32 base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; w++) {
    u32 t = tmp[w];
    if ((u32)w < wave) base += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_cells_count(CellDesc cd, u64 w0, u64 nwords, int touched_only,
                                                        u32* __restrict__ cnt, u64* __restrict__ pcb) {
  __shared__ u32 tmp[kBlock / 64];
  __shared__ u64 ptmp[kBlock / 64];
  if (cd.gate && (cd.gate->done | cd.gate->leave)) return;  // uniform over the launch
  const u64 wl = (u64)blockIdx.x * kBlock + threadIdx.x;
  const u32 b = wl < nwords ? cells_word(cd, w0 + wl, touched_only != 0) : 0u;
  u32 tot;
  cells_block_scan((u32)__popc(b), tmp, &tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
  if (pcb) {
    // push cost of the touched cells: each costs its row's degree in pc_csr
    // (rows of 2^rlbits cells; w0 == 0 here)
    u64 pc = 0;
    if (b) {
      const u64 wi = w0 + wl;
      const int lb = cd.rlbits;
      if (lb >= 5) {
        const u64 r = wi >> (lb - 5);
        if (r < cd.pc_nv) pc = (u64)__popc(b) * (cd.pc_csr[r + 1] - cd.pc_csr[r]);
      } else {
        const int per = 32 >> lb;
        const u32 m = (1u << (1u << lb)) - 1u;
        for (int q = 0; q < per; q++) {
          const u32 lanes = (b >> (q << lb)) & m;
          const u64 r = wi * (u64)per + q;
          if (lanes && r < cd.pc_nv) pc += (u64)__popc(lanes) * (cd.pc_csr[r + 1] - cd.pc_csr[r]);
        }
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pc += __shfl_xor(pc, o, 64);
    if ((threadIdx.x & 63) == 0) ptmp[threadIdx.x >> 6] = pc;
    __syncthreads();
    if (threadIdx.x == 0) {
      u64 s = 0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; w++) s += ptmp[w];
      pcb[blockIdx.x] = s;
    }
  }
}

template <class K>
__global__ __launch_bounds__(kBlock) void k_cells_write(CellDesc cd, u64 w0, u64 nwords, int touched_only,
                                                        const u64* __restrict__ off, int bits, K* __restrict__ keys,
                                                        i64* __restrict__ w, u8* __restrict__ inf) {
  __shared__ u32 tmp[kBlock / 64];
  __shared__ unsigned short sc[kCellsPerBlock];
  if (cd.gate && (cd.gate->done | cd.gate->leave)) return;  // uniform over the launch
  const u64 wl = (u64)blockIdx.x * kBlock + threadIdx.x;
  const u64 wi = w0 + wl;
  u32 b = wl < nwords ? cells_word(cd, wi, touched_only != 0) : 0u;
  if (touched_only && b) cd.touched[wi] = 0;
  if (touched_only && cd.dbits && wl < nwords) {
    // the delta as a bitmap, every word rewritten; rows holding a delta cell
    // carry this compaction's stamp (racing writers store the same value)
    cd.dbits[wi] = b;
    if (b) {
      const u64 r = wi >> cd.dwshift;
      if (cd.rstamp[r] != cd.stamp) cd.rstamp[r] = cd.stamp;
    }
  }
  u32 tot;
  u32 o = cells_block_scan((u32)__popc(b), tmp, &tot);
  if (touched_only && (cd.rowrec || cd.rowrec8)) {
    // row records (w0 == 0, so word wi sits in lane wi % 64 of its wave)
    const u64 g = off[blockIdx.x] + o;
    const u32 hi = __shfl_down(b, 1, 64);
    if (wl < nwords) {
      if (cd.rlbits == 6) {
        const u64 r = wi >> 1;
        if (!(wi & 1) && r < cd.rnrows) {
          ulonglong2 v;
          v.x = (u64)b | ((u64)hi << 32);
          v.y = g;
          if (cd.rowrec) *(ulonglong2*)(cd.rowrec + 2 * r) = v;
          if (cd.rowrec8) cd.rowrec8[r] = g | ((u64)(__popc(b) + __popc(hi)) << 40);
        }
      } else {
        const int lb = cd.rlbits, per = 32 >> lb;
        const u32 m = (lb == 5) ? ~0u : ((1u << (1u << lb)) - 1u);
        for (int q = 0; q < per; q++) {
          const u64 r = wi * (u64)per + q;
          if (r >= cd.rnrows) break;
          const int sh = q << lb;
          ulonglong2 v;
          v.x = (b >> sh) & m;
          v.y = g + (u32)__popc(b & ((1u << sh) - 1u));
          if (cd.rowrec) *(ulonglong2*)(cd.rowrec + 2 * r) = v;
          if (cd.rowrec8) cd.rowrec8[r] = v.y | ((u64)__popc((u32)v.x) << 40);
        }
      }
    }
  }
  if (touched_only && cd.pact && cd.pact_wshift >= 0 && b) {
    const u64 pr = wi >> cd.pact_wshift;
    const u32 pid = cd.pact_list ? cd.pact_list[pr] : (u32)pr;
    if (!cd.pact[pid]) cd.pact[pid] = 1;
  }
  if (touched_only && cd.grp_mask) {
    // per-group lanes and push cost of this block's rows (one group)
    __shared__ u64 gm[kBlock / 64], gp[kBlock / 64];
    const u32 hi = __shfl_down(b, 1, 64);
    u64 m = 0, pc = 0;
    if (wl < nwords && !(wi & 1)) {
      const u64 r = wi >> 1;
      if (r < cd.rnrows) {
        m = (u64)b | ((u64)hi << 32);
        const u64 z = r & ((1ull << cd.grp_shift) - 1);
        if (m && z < cd.pc_nv) pc = (u64)__popcll(m) * (cd.pc_csr[z + 1] - cd.pc_csr[z]);
        if (!z) cd.grp_start[r >> cd.grp_shift] = off[blockIdx.x] + o;
      }
    }
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      m |= __shfl_xor(m, q, 64);
      pc += __shfl_xor(pc, q, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      gm[threadIdx.x >> 6] = m;
      gp[threadIdx.x >> 6] = pc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      u64 bm = 0, bp = 0;
#pragma unroll
      for (int q = 0; q < kBlock / 64; q++) {
        bm |= gm[q];
        bp += gp[q];
      }
      const u64 g = ((u64)blockIdx.x * kBlock / 2) >> cd.grp_shift;
      if (bp) atomicAdd((unsigned long long*)&cd.grp_cost[g], (unsigned long long)bp);
      if (bm && (bm & ~__hip_atomic_load(&cd.grp_mask[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
        atomicOr((unsigned long long*)&cd.grp_mask[g], (unsigned long long)bm);
    }
  }
  while (b) {
    int bit = __ffs(b) - 1;
    b &= b - 1;
    sc[o++] = (unsigned short)(threadIdx.x * 32 + bit);
  }
  __syncthreads();
  const u64 base = off[blockIdx.x];
  const u64 cbase = w0 * 32 + (u64)blockIdx.x * kCellsPerBlock;
  for (u32 j = threadIdx.x; j < tot; j += kBlock) {
    const u64 c = cbase + sc[j];
    if (keys) {
      K key = 0;
      u32 pid = 0;
      for (int q = 0; q < cd.ncols; q++) {
        u64 r = (c >> cd.shift[q]) & ((1ull << cd.ebits[q]) - 1);
        u32 code = cd.pdom_off[q]  ? cd.pdom_codes[q][cd.pdom_off[q][pid] + r]
                   : cd.dom_list[q] ? cd.dom_list[q][r]
                                    : (u32)r;
        if (q == 0) pid = code;
        key = (key << bits) | (K)code;
      }
      keys[base + j] = key;
    }
    if (touched_only && cd.dlane) cd.dlane[base + j] = (u8)(c & ((1u << cd.rlbits) - 1u));
    if (cd.trop) {
      const u64 e = cell_peek(cd, c);
      i64 wv = 0;
      u8 f = 0;
      if (!cell_code_absent(cd, e)) wv = cell_code_value(cd, e);
      else if (!cd.narrow && (cd.maxf[c >> 5] & (1u << (c & 31)))) wv = LLONG_MAX;
      else f = 1;
      if (w) w[base + j] = wv;
      if (inf) inf[base + j] = f;
      if (touched_only && cd.dpack) cd.dpack[base + j] = ((u32)wv << 6) | (u32)(c & ((1u << cd.rlbits) - 1u));
    }
  }
}

template <class K>
u64 cells_compact(const Exec& x, const CellDesc& cd, u64 ncells, bool touched_only, int bits, DBuf<K>* keys,
                  DBuf<i64>* w, DBuf<u8>* inf, bool want_inf, u64 cell0, u64* push_cost) {
  const u64 w0 = cell0 / 32;  // cell0 is a multiple of 32
  if (cd.rowrec && touched_only && (w0 || cd.rlbits < 0 || cd.rlbits > 6))
    throw std::runtime_error("cells_compact: row records need a touched compaction from cell 0, rows <= 64 cells");
  if (push_cost) *push_cost = ~0ull;
  const u64 nwords = (ncells + 31) / 32;
  const u64 nblocks = (nwords + kBlock - 1) / kBlock;
  DBuf<u32> cnt(x.pool, nblocks);
  DBuf<u64> off(x.pool, nblocks);
  DBuf<u64> tot(x.pool, 2);
  // the push cost rides along with the count when the block counts fit one
  // scan launch
  const bool pc = push_cost && cd.pc_csr && cd.rowrec && touched_only && nblocks <= kScanOneMax;
  DBuf<u64> pcb(x.pool, pc ? nblocks : 1);
  const double scan_bytes = touched_only ? 4.0 * nwords : (cd.trop ? 8.0 * ncells + 8.0 * nwords : 4.0 * nwords);
  {
    Scope sc(x.timer, "cells_count", KC_DEDUP, scan_bytes, ncells);
    k_cells_count<<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, w0, nwords, touched_only ? 1 : 0, cnt.get(),
                                                              pc ? pcb.get() : nullptr);
    SK_HIP(hipGetLastError());
  }
  u64 hv[2] = {0, 0};
  if (nblocks <= kScanOneMax) {
    scan_u32_one(x, cnt.get(), off.get(), nblocks, tot.get(), pc ? pcb.get() : nullptr);
    read_u64s(x, tot.get(), hv, pc ? 2 : 1);
    if (pc) *push_cost = hv[1];
  } else {
    scan_u32(x, cnt.get(), off.get(), nblocks, tot.get());
    hv[0] = read_u64(x, tot.get());
  }
  const u64 n = hv[0];
  if (n == 0) return 0;
  keys->reset(x.pool, n);
  if (cd.trop) {
    w->reset(x.pool, n);
    if (want_inf) inf->reset(x.pool, n);
  }
  const double rec = sizeof(K) + (cd.trop ? 16.0 + (want_inf ? 1.0 : 0) : 0);
  Scope sc(x.timer, "cells_write", KC_DEDUP, scan_bytes + rec * n, ncells);
  k_cells_write<K><<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, w0, nwords, touched_only ? 1 : 0, off.get(), bits,
                                                               keys->get(), cd.trop ? w->get() : nullptr,
                                                               (cd.trop && want_inf) ? inf->get() : nullptr);
  SK_HIP(hipGetLastError());
  return n;
}

// Bitmap-only compaction of a device-driven round: every touched word
// becomes the delta bitmap's word (zeros included) and is cleared; rows
// holding delta cells get the new stamp and their program marked; one
// atomic per wave that saw a touched word sets loop->any.
__global__ __launch_bounds__(kBlock) void k_loop_compact(CellDesc cd, u64 nwords, LoopState* ls) {
  if (ls->done) return;  // uniform over the launch
  const u64 wi = (u64)blockIdx.x * kBlock + threadIdx.x;
  u32 b = 0;
  if (wi < nwords) {
    b = cd.touched[wi];
    if (b) cd.touched[wi] = 0;
    cd.dbits[wi] = b;
    if (b) {
      const u32 st = ls->stamp + 1u;
      const u64 r = wi >> cd.dwshift;
      if (cd.rstamp[r] != st) cd.rstamp[r] = st;
      if (cd.pact) {
        const u64 pr = r >> cd.pact_shift;
        const u32 pid = cd.pact_list ? cd.pact_list[pr] : (u32)pr;
        if (!cd.pact[pid]) cd.pact[pid] = 1;
      }
    }
  }
  if (__ballot(b != 0) && (threadIdx.x & 63) == 0 && !ls->any) atomicOr(&ls->any, 1u);
}

__global__ __launch_bounds__(kBlock) void k_loop_book(LoopState* ls, u8* __restrict__ active_prev,
                                                      u8* __restrict__ active_next, u64* __restrict__ piters,
                                                      u32 nparts) {
  if (ls->done) return;  // uniform
  for (u32 p = threadIdx.x; p < nparts; p += kBlock) {
    if (active_prev[p]) piters[p] += 1;
    active_prev[p] = active_next[p];
    active_next[p] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ls->iters += 1;
    ls->stamp += 1;
    if (!ls->any) ls->done = 1;
    ls->any = 0;
  }
}

void cells_loop_round(const Exec& x, const CellDesc& cd, u64 ncells, LoopState* ls, u8* active_prev,
                      u8* active_next, u64* piters, u32 nparts) {
  const u64 nwords = (std::max<u64>(ncells, 32) + 31) / 32;
  const u64 nblocks = (nwords + kBlock - 1) / kBlock;
  if (!cd.dbits || nblocks > 0x7fffffffull) throw std::runtime_error("cells_loop_round: layout");
  k_loop_compact<<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, nwords, ls);
  SK_HIP(hipGetLastError());
  k_loop_book<<<1, kBlock, 0, x.stream>>>(ls, active_prev, active_next, piters, nparts);
  SK_HIP(hipGetLastError());
}

// The round's bookkeeping of device-driven Trop rounds: leave on a device
// status (2) or past the round cap (3), done on an empty delta; otherwise
// the next round pulls, or pushes when the delta's push cost is below
// push_limit: on the device when the delta fits the push table (push_cap
// rows), else the rounds leave for the host (1).
__global__ void k_loop_book_trop(LoopState* ls, const u64* __restrict__ tot, u64 cap_left, u64 push_limit,
                                 u64 push_cap) {
  if (threadIdx.x != 0 || ls->done || ls->leave) return;
  const DevErr* st = (const DevErr*)(tot + 2);
  if (st->code || st->widen) {
    ls->st = *st;
    ls->leave = 2;
    return;
  }
  ls->iters += 1;
  ls->n = tot[0];
  ls->pc = tot[1];
  ls->stamp = 0;
  if (ls->iters > cap_left) ls->leave = 3;
  else if (tot[0] == 0) ls->done = 1;
  else if (tot[1] < push_limit) {
    if (tot[0] <= push_cap) ls->stamp = 1;
    else ls->leave = 1;
  }
}

template <class K>
void cells_loop_round_trop(const Exec& x, const CellDesc& cd, u64 ncells, int bits, K* keys, i64* w, u8* inf,
                           u32* cnt, u64* off, u64* pcb, u64* tot, const DevErr* status, LoopState* ls,
                           u64 cap_left, u64 push_limit, u64 push_cap) {
  static_assert(sizeof(DevErr) == 3 * sizeof(u64), "status record is three words");
  const u64 nwords = (ncells + 31) / 32;
  const u64 nblocks = (nwords + kBlock - 1) / kBlock;
  if (!cells_into_fits(ncells) || !cd.gate || !cd.rowrec || !cd.pc_csr || cd.rlbits < 0 || cd.rlbits > 6)
    throw std::runtime_error("cells_loop_round_trop: layout");
  k_cells_count<<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, cnt, pcb);
  SK_HIP(hipGetLastError());
  scan_u32_one(x, cnt, off, nblocks, tot, pcb, (const u64*)status, 3);
  k_cells_write<K><<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, off, bits, keys, w, inf);
  SK_HIP(hipGetLastError());
  k_loop_book_trop<<<1, 64, 0, x.stream>>>(ls, tot, cap_left, push_limit, push_cap);
  SK_HIP(hipGetLastError());
}

// Between the count and the write pass of a device-driven batched Trop
// round: a device status (widen, overflow) or any program's overflow record
// leaves the rounds (2) before the write pass consumes the touched cells, so
// the host resumes the round after its join with the previous delta in place.
__global__ __launch_bounds__(kBlock) void k_loop_check_parts(LoopState* ls, const u64* __restrict__ tot,
                                                             const DevErr* __restrict__ part_derr, u32 nparts) {
  __shared__ u32 s_ovf;
  if (ls->done | ls->leave) return;  // uniform over the block
  if (threadIdx.x == 0) s_ovf = 0;
  __syncthreads();
  for (u32 p = threadIdx.x; p < nparts; p += kBlock)
    if (part_derr[p].code) s_ovf = 1;
  __syncthreads();
  const DevErr* st = (const DevErr*)(tot + 2);
  if (threadIdx.x == 0 && (st->code || st->widen || s_ovf)) {
    ls->st = *st;
    ls->leave = 2;
  }
}

// The round's per-program bookkeeping (after the write pass): every program
// active in the previous round gets one round, active_prev takes
// active_next (cleared), iters += 1, then leave (3) once some program
// passed its cap, else done on an empty delta.
__global__ __launch_bounds__(kBlock) void k_loop_book_trop_parts(LoopState* ls, const u64* __restrict__ tot,
                                                                 u8* __restrict__ active_prev,
                                                                 u8* __restrict__ active_next,
                                                                 u64* __restrict__ piters,
                                                                 const u64* __restrict__ capleft, u32 nparts) {
  __shared__ u32 s_cap;
  if (ls->done | ls->leave) return;  // uniform over the block
  if (threadIdx.x == 0) s_cap = 0;
  __syncthreads();
  for (u32 p = threadIdx.x; p < nparts; p += kBlock) {
    if (active_prev[p]) {
      piters[p] += 1;
      if (piters[p] > capleft[p]) s_cap = 1;
    }
    active_prev[p] = active_next[p];
    active_next[p] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ls->iters += 1;
    ls->n = tot[0];
    if (s_cap) ls->leave = 3;
    else if (tot[0] == 0) ls->done = 1;
  }
}

template <class K>
void cells_loop_round_trop_parts(const Exec& x, const CellDesc& cd, u64 ncells, int bits, K* keys, i64* w, u32* cnt,
                                 u64* off, u64* tot, const DevErr* status, LoopState* ls, u8* active_prev,
                                 u8* active_next, u64* piters, const u64* capleft, const DevErr* part_derr,
                                 u32 nparts) {
  static_assert(sizeof(DevErr) == 3 * sizeof(u64), "status record is three words");
  const u64 nwords = (ncells + 31) / 32;
  const u64 nblocks = (nwords + kBlock - 1) / kBlock;
  if (!cells_into_fits(ncells) || !cd.gate || !cd.pact || cd.pact_wshift < 0 || cd.pact != active_next)
    throw std::runtime_error("cells_loop_round_trop_parts: layout");
  k_cells_count<<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, cnt, nullptr);
  SK_HIP(hipGetLastError());
  scan_u32_one(x, cnt, off, nblocks, tot, nullptr, (const u64*)status, 3);
  k_loop_check_parts<<<1, kBlock, 0, x.stream>>>(ls, tot, part_derr, nparts);
  SK_HIP(hipGetLastError());
  k_cells_write<K><<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, off, bits, keys, w, nullptr);
  SK_HIP(hipGetLastError());
  k_loop_book_trop_parts<<<1, kBlock, 0, x.stream>>>(ls, tot, active_prev, active_next, piters, capleft, nparts);
  SK_HIP(hipGetLastError());
}

bool cells_into_fits(u64 ncells) {
  const u64 nwords = (ncells + 31) / 32;
  return (nwords + kBlock - 1) / kBlock <= kScanOneMax;
}

template <class K>
u64 cells_compact_into(const Exec& x, const CellDesc& cd, u64 ncells, int bits, K* keys, i64* w, u8* inf,
                       u64* push_cost, const DevErr* status, DevErr* status_out) {
  static_assert(sizeof(DevErr) == 3 * sizeof(u64), "status record is three words");
  const u64 nwords = (ncells + 31) / 32;
  const u64 nblocks = (nwords + kBlock - 1) / kBlock;
  if (!cells_into_fits(ncells) || nblocks == 0) throw std::runtime_error("cells_compact_into: too many cells");
  if (cd.rowrec && (cd.rlbits < 0 || cd.rlbits > 6)) throw std::runtime_error("cells_compact_into: row records");
  if (push_cost) *push_cost = ~0ull;
  const bool pc = push_cost && cd.pc_csr && cd.rowrec;
  DBuf<u32> cnt(x.pool, nblocks);
  DBuf<u64> off(x.pool, nblocks);
  DBuf<u64> tot(x.pool, 8);
  DBuf<u64> pcb(x.pool, pc ? nblocks : 1);
  const double scan_bytes = 4.0 * nwords;
  {
    Scope sc(x.timer, "cells_count", KC_DEDUP, scan_bytes, ncells);
    k_cells_count<<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, cnt.get(), pc ? pcb.get() : nullptr);
    SK_HIP(hipGetLastError());
  }
  scan_u32_one(x, cnt.get(), off.get(), nblocks, tot.get(), pc ? pcb.get() : nullptr, (const u64*)status,
               status ? 3 : 0);
  {
    // the records' bytes are credited once the count is known
    Scope sc(x.timer, "cells_write", KC_DEDUP, scan_bytes, ncells);
    k_cells_write<K><<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, off.get(), bits, keys,
                                                                 cd.trop ? w : nullptr, cd.trop ? inf : nullptr);
    SK_HIP(hipGetLastError());
  }
  u64 hv[5] = {0, 0, 0, 0, 0};
  read_u64s(x, tot.get(), hv, status ? 5 : (pc ? 2 : 1));
  if (pc) *push_cost = hv[1];
  if (status) std::memcpy((void*)status_out, &hv[2], sizeof(DevErr));
  const u64 n = hv[0];
  if (x.timer) x.timer->credit("cells_write", KC_DEDUP, (sizeof(K) + (cd.trop ? 16.0 + (inf ? 1.0 : 0) : 0)) * n, 0);
  return n;
}

// Touched compaction (from cell 0) of a layout whose rows are 2^rlbits <= 64
// cells, writing the row records (cd.rowrec8, and cd.rowrec when set) and
// per listed cell only what is asked: the packed word ((weight << 6) | lane,
// narrow cells below narrow_lim) into *dpack, or the weight and the lane into
// *w and *lane.  The outputs are (re)allocated to the count, which is
// returned; no keys are written.
u64 cells_compact_rows(const Exec& x, CellDesc cd, u64 ncells, DBuf<u32>* dpack, DBuf<i64>* w, DBuf<u8>* lane) {
  if (!cd.trop || (!cd.rowrec && !cd.rowrec8) || cd.rlbits < 0 || cd.rlbits > 6 || cd.dbits)
    throw std::runtime_error("cells_compact_rows: Trop cells with row records of at most 64 cells");
  if (cd.grp_mask && (cd.rlbits != 6 || cd.grp_shift < 8 || !cd.pc_csr || !cd.grp_cost || !cd.grp_start))
    throw std::runtime_error("cells_compact_rows: group statistics need rows of 64 cells, groups of >= 256 rows");
  if (dpack && !cd.narrow) throw std::runtime_error("cells_compact_rows: packed words need narrow cells");
  const u64 nwords = (ncells + 31) / 32;
  const u64 nblocks = (nwords + kBlock - 1) / kBlock;
  if (nblocks == 0) return 0;
  DBuf<u32> cnt(x.pool, nblocks);
  DBuf<u64> off(x.pool, nblocks);
  DBuf<u64> tot(x.pool, 2);
  const double scan_bytes = 4.0 * nwords;
  {
    Scope sc(x.timer, "cells_count", KC_DEDUP, scan_bytes, ncells);
    k_cells_count<<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, cnt.get(), nullptr);
    SK_HIP(hipGetLastError());
  }
  u64 n = 0;
  if (nblocks <= kScanOneMax) {
    scan_u32_one(x, cnt.get(), off.get(), nblocks, tot.get(), nullptr);
    n = read_u64(x, tot.get());
  } else {
    scan_u32(x, cnt.get(), off.get(), nblocks, tot.get());
    n = read_u64(x, tot.get());
  }
  cd.dpack = nullptr;
  cd.dlane = nullptr;
  if (dpack) {
    dpack->reset(x.pool, n ? n : 1);
    cd.dpack = dpack->get();
  }
  if (w) w->reset(x.pool, n ? n : 1);
  if (lane) {
    lane->reset(x.pool, n ? n : 1);
    cd.dlane = lane->get();
  }
  // the row records are rewritten even when nothing was touched
  const double rec = (dpack ? 4.0 : 0) + (w ? 8.0 : 0) + (lane ? 1.0 : 0);
  Scope sc(x.timer, "cells_write", KC_DEDUP, scan_bytes + rec * n + 8.0 * (double)cd.rnrows, ncells);
  k_cells_write<u64><<<(unsigned)nblocks, kBlock, 0, x.stream>>>(cd, 0, nwords, 1, off.get(), 0, nullptr,
                                                                 w ? w->get() : nullptr, nullptr);
  SK_HIP(hipGetLastError());
  return n;
}

template <class K>
__global__ void k_cells_mark_touched(CellDesc cd, const K* __restrict__ keys, u64 n, int arity, int bits) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const K key = keys[i];
  u64 c;
  if (!cell_of(cd, [&](int j) { return code_at<K>(key, arity, bits, j); }, &c)) return;
  cell_touch(cd, c);
}

template <class K>
void cells_mark_touched(const Exec& x, const CellDesc& cd, const K* keys, u64 n, int arity, int bits) {
  if (!n) return;
  k_cells_mark_touched<K><<<gridc(n), kBlock, 0, x.stream>>>(cd, keys, n, arity, bits);
  SK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ gather

__global__ void k_cells_gather(CellDesc cd, const u64* __restrict__ idx, u64 n, i64* __restrict__ w,
                               u8* __restrict__ present) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 c = idx[i];
  const u32 m = 1u << (c & 31);
  if (!cd.trop) {
    present[i] = (cd.present[c >> 5] & m) ? 1 : 0;
    if (w) w[i] = 0;
    return;
  }
  const u64 e = cell_peek(cd, c);
  i64 wv = 0;
  u8 pr = 1;
  if (!cell_code_absent(cd, e)) wv = cell_code_value(cd, e);
  else if (!cd.narrow && (cd.maxf[c >> 5] & m)) wv = LLONG_MAX;
  else if (!(cd.infb && (cd.infb[c >> 5] & m))) pr = 0;  // +inf: present with weight 0
  w[i] = wv;
  present[i] = pr;
}

void cells_gather(const Exec& x, const CellDesc& cd, const u64* idx, u64 n, i64* w, u8* present) {
  if (!n) return;
  k_cells_gather<<<gridc(n), kBlock, 0, x.stream>>>(cd, idx, n, w, present);
  SK_HIP(hipGetLastError());
}

__global__ void k_cells_widen(const u32* __restrict__ c32, u64* __restrict__ c64, u64 n) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 e = c32[i];
  c64[i] = e == kCellTop32 ? kCellTop : cell_enc((i64)e);
}

void cells_widen(const Exec& x, const u32* c32, u64* c64, u64 n) {
  if (!n) return;
  Scope sc(x.timer, "cells_widen", KC_MISC, (double)n * 12.0, n);
  k_cells_widen<<<gridc(n), kBlock, 0, x.stream>>>(c32, c64, n);
  SK_HIP(hipGetLastError());
}

#define SK_CELLS_INST(K)                                                                                      \
  template void cells_mark_col<K>(const Exec&, const K*, u64, int, int, int, u32*);                          \
  template void cells_load<K>(const Exec&, const CellDesc&, const K*, const i64*, const u8*, u64, int, int);  \
  template void cells_insert<K>(const Exec&, const CellDesc&, const K*, const i64*, const u8*, u64, int, int); \
  template u64 cells_compact<K>(const Exec&, const CellDesc&, u64, bool, int, DBuf<K>*, DBuf<i64>*, DBuf<u8>*, bool, \
                                u64, u64*);                                                                \
  template u64 cells_compact_into<K>(const Exec&, const CellDesc&, u64, int, K*, i64*, u8*, u64*, const DevErr*,   \
                                     DevErr*);                                                                    \
  template void cells_mark_touched<K>(const Exec&, const CellDesc&, const K*, u64, int, int);                 \
  template void cells_loop_round_trop<K>(const Exec&, const CellDesc&, u64, int, K*, i64*, u8*, u32*, u64*, u64*, \
                                         u64*, const DevErr*, LoopState*, u64, u64, u64);                             \
  template void cells_loop_round_trop_parts<K>(const Exec&, const CellDesc&, u64, int, K*, i64*, u32*, u64*, u64*,  \
                                               const DevErr*, LoopState*, u8*, u8*, u64*, const u64*,             \
                                               const DevErr*, u32);
SK_CELLS_INST(u64)
SK_CELLS_INST(u128)

}  // namespace sk
