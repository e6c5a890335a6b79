// sk_serve.hip — kernels of the query server's lane groups (see sk_serve.h).
#include <hip/hip_runtime.h>

#include <climits>
#include <stdexcept>
#include <vector>

#include "sk_serve.h"

namespace sk {

namespace {

constexpr u64 kPosMask = (1ull << 40) - 1;
// Rows per wave of the all-lanes answer pass: a wave reads 64 lanes of each
// of its rows (one 256-byte line of narrow cells), lane t keeping instance t.
constexpr u64 kSrvRowsPerWave = 256;

unsigned grid_of(u64 n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ bool srv_present(const CellDesc& cd, u64 c, i64* w) {
  const u64 e = cell_peek(cd, c);
  if (!cell_code_absent(cd, e)) {
    *w = cell_code_value(cd, e);
    return true;
  }
  if (!cd.narrow && (cd.maxf[c >> 5] & (1u << (c & 31)))) {
    *w = LLONG_MAX;
    return true;
  }
  return false;
}

__device__ __forceinline__ void srv_fill(const CellDesc& cd, u64 c, u64 fill, bool clear_maxf) {
  if (cd.narrow) {
    cd.cell32[c] = (u32)fill;
    return;
  }
  cd.cell[c] = fill;
  if (clear_maxf) atomicAnd(&cd.maxf[c >> 5], ~(1u << (c & 31)));
}

__global__ __launch_bounds__(kBlock) void k_srv_table(const u64* __restrict__ rowrec8, int zbits, u32 g,
                                                      const u32* __restrict__ dpack, const i64* __restrict__ dw,
                                                      const u8* __restrict__ dlane, u64 p0, u64 p1,
                                                      u32* __restrict__ tg, u32* __restrict__ tl,
                                                      u32* __restrict__ ty, u32* __restrict__ ti,
                                                      i64* __restrict__ tw) {
  const u64 z = (u64)blockIdx.x * kBlock + threadIdx.x;
  if (z >> zbits) return;
  const u64 rec = rowrec8[((u64)g << zbits) + z];
  const u32 cnt = (u32)(rec >> 40);
  const u64 pos = rec & kPosMask;
  if (!cnt || pos >= p1 || pos + cnt <= p0) return;
  for (u32 k = 0; k < cnt; k++) {
    const u64 p = pos + k;
    if (p < p0 || p >= p1) continue;
    const u64 i = p - p0;
    u32 l;
    i64 w;
    if (dpack) {
      const u32 v = dpack[p];
      l = v & 63u;
      w = (i64)(v >> 6);
    } else {
      l = dlane[p];
      w = dw[p];
    }
    tg[i] = g;
    tl[i] = l;
    ty[i] = (u32)z;
    ti[i] = g * 64 + l;
    tw[i] = w;
  }
}

// All-lanes answer pass: wave v takes rows [v * kSrvRowsPerWave, ...), lane t
// instance t.  Counts are laid out lane-major (cnt[t * nwaves + v]), so one
// exclusive scan gives each lane's rows contiguous and in node order.
__global__ __launch_bounds__(kBlock) void k_srv_rows_count(CellDesc cd, u64 n_nodes, u64 nwaves, u64 take,
                                                           u32* __restrict__ cnt) {
  const u64 v = (u64)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (v >= nwaves) return;
  const u32 lane = threadIdx.x & 63;
  u32 c = 0;
  if ((take >> lane) & 1) {
    const u64 r0 = v * kSrvRowsPerWave, r1 = r0 + kSrvRowsPerWave < n_nodes ? r0 + kSrvRowsPerWave : n_nodes;
    for (u64 r = r0; r < r1; r++) {
      i64 w;
      c += srv_present(cd, (r << 6) | lane, &w) ? 1u : 0u;
    }
  }
  cnt[(u64)lane * nwaves + v] = c;
}

__global__ __launch_bounds__(kBlock) void k_srv_rows_write(CellDesc cd, u64 n_nodes, u64 nwaves, u64 take,
                                                           u64 clear, u64 fill, int clear_maxf,
                                                           const u64* __restrict__ off, u32* const* __restrict__ oz,
                                                           i64* const* __restrict__ ow) {
  const u64 v = (u64)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (v >= nwaves) return;
  const u32 lane = threadIdx.x & 63;
  const bool tk = (take >> lane) & 1, cl = (clear >> lane) & 1;
  if (!tk && !cl) return;
  u64 pos = 0;
  u32* z = nullptr;
  i64* w = nullptr;
  if (tk) {
    pos = off[(u64)lane * nwaves + v] - off[(u64)lane * nwaves];
    z = oz[lane];
    w = ow[lane];
  }
  const u64 r0 = v * kSrvRowsPerWave, r1 = r0 + kSrvRowsPerWave < n_nodes ? r0 + kSrvRowsPerWave : n_nodes;
  for (u64 r = r0; r < r1; r++) {
    const u64 c = (r << 6) | lane;
    if (tk) {
      i64 val;
      if (srv_present(cd, c, &val)) {
        z[pos] = (u32)r;
        w[pos] = val;
        pos++;
      }
    }
    if (cl) srv_fill(cd, c, fill, clear_maxf != 0);
  }
}

__global__ void k_srv_lane_tot(const u64* __restrict__ off, u64 nwaves, const u64* __restrict__ total,
                               u64* __restrict__ out) {
  const u32 l = threadIdx.x;
  if (l >= 64) return;
  const u64 hi = l < 63 ? off[(u64)(l + 1) * nwaves] : *total;
  out[l] = hi - off[(u64)l * nwaves];
}

// One lane: a thread per row, wave counts by ballot.
__global__ __launch_bounds__(kBlock) void k_srv_lane_count(CellDesc cd, u64 n_nodes, u32 lane,
                                                           u32* __restrict__ cnt) {
  const u64 r = (u64)blockIdx.x * kBlock + threadIdx.x;
  i64 w;
  const bool p = r < n_nodes && srv_present(cd, (r << 6) | lane, &w);
  const u64 b = __ballot(p);
  if ((threadIdx.x & 63) == 0) cnt[(u64)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)] = (u32)__popcll(b);
}

__global__ __launch_bounds__(kBlock) void k_srv_lane_write(CellDesc cd, u64 n_nodes, u32 lane, int clear, u64 fill,
                                                           int clear_maxf, const u64* __restrict__ off,
                                                           u32* __restrict__ oz, i64* __restrict__ ow) {
  const u64 r = (u64)blockIdx.x * kBlock + threadIdx.x;
  const u64 c = (r << 6) | lane;
  i64 val = 0;
  const bool p = oz && r < n_nodes && srv_present(cd, c, &val);
  const u64 b = __ballot(p);
  if (p) {
    const u32 t = threadIdx.x & 63;
    const u64 below = t ? (b & (~0ull >> (64 - t))) : 0ull;
    const u64 pos = off[(u64)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)] + (u64)__popcll(below);
    oz[pos] = (u32)r;
    ow[pos] = val;
  }
  if (clear && r < n_nodes) srv_fill(cd, c, fill, clear_maxf != 0);
}

}  // namespace

void srv_table(const Exec& x, const u64* rowrec8, int zbits, u32 g, const u32* dpack, const i64* dw,
               const u8* dlane, u64 p0, u64 p1, u32* tg, u32* tl, u32* ty, u32* ti, i64* tw) {
  if (p1 <= p0) return;
  Scope sc(x.timer, "srv_table", KC_MISC, 8.0 * (double)(1ull << zbits) + 24.0 * (double)(p1 - p0), p1 - p0);
  k_srv_table<<<grid_of(1ull << zbits), kBlock, 0, x.stream>>>(rowrec8, zbits, g, dpack, dw, dlane, p0, p1, tg, tl,
                                                               ty, ti, tw);
  SK_HIP(hipGetLastError());
}

void srv_lanes(const Exec& x, const CellDesc& gcd, u64 n_nodes, u64 take_mask, u64 clear_mask, bool poison,
               DBuf<u32>* outz, DBuf<i64>* outw, u64* counts) {
  for (int l = 0; l < 64; l++) counts[l] = 0;
  if (!(take_mask | clear_mask) || !n_nodes) return;
  const u64 fill = poison ? 0ull : (gcd.narrow ? (u64)kCellTop32 : kCellTop);
  const int clear_maxf = (!poison && !gcd.narrow) ? 1 : 0;
  const u64 cell_b = gcd.narrow ? 4 : 8;
  const int nl = __builtin_popcountll(take_mask | clear_mask);
  if (nl <= 4) {
    // a few lanes: one strided pass per lane (a load per row)
    const unsigned nb = grid_of(n_nodes);
    for (int l = 0; l < 64; l++) {
      const bool tk = (take_mask >> l) & 1, cl = (clear_mask >> l) & 1;
      if (!tk && !cl) continue;
      DBuf<u64> off;
      if (tk) {
        DBuf<u32> cnt(x.pool, (u64)nb * (kBlock / 64));
        off.reset(x.pool, (u64)nb * (kBlock / 64));
        DBuf<u64> tot(x.pool, 1);
        {
          Scope sc(x.timer, "srv_answer", KC_MISC, (double)cell_b * n_nodes, n_nodes);
          k_srv_lane_count<<<nb, kBlock, 0, x.stream>>>(gcd, n_nodes, (u32)l, cnt.get());
          SK_HIP(hipGetLastError());
        }
        scan_u32(x, cnt.get(), off.get(), (u64)nb * (kBlock / 64), tot.get());
        counts[l] = read_u64(x, tot.get());
        outz[l].reset(x.pool, counts[l] ? counts[l] : 1);
        outw[l].reset(x.pool, counts[l] ? counts[l] : 1);
      }
      Scope sc(x.timer, "srv_answer", KC_MISC, (double)cell_b * n_nodes + 12.0 * counts[l], n_nodes);
      k_srv_lane_write<<<nb, kBlock, 0, x.stream>>>(gcd, n_nodes, (u32)l, cl ? 1 : 0, fill, clear_maxf,
                                                    tk ? off.get() : nullptr, tk ? outz[l].get() : nullptr,
                                                    tk ? outw[l].get() : nullptr);
      SK_HIP(hipGetLastError());
    }
    return;
  }
  // many lanes: whole rows per wave
  const u64 nwaves = (n_nodes + kSrvRowsPerWave - 1) / kSrvRowsPerWave;
  const unsigned nb = (unsigned)((nwaves + kBlock / 64 - 1) / (kBlock / 64));
  DBuf<u64> off;
  DBuf<u64> ptrs(x.pool, 128);
  if (take_mask) {
    DBuf<u32> cnt(x.pool, 64 * nwaves);
    off.reset(x.pool, 64 * nwaves);
    DBuf<u64> tot(x.pool, 1), lt(x.pool, 64);
    {
      Scope sc(x.timer, "srv_answer", KC_MISC, (double)cell_b * 64 * n_nodes, n_nodes);
      k_srv_rows_count<<<nb, kBlock, 0, x.stream>>>(gcd, n_nodes, nwaves, take_mask, cnt.get());
      SK_HIP(hipGetLastError());
    }
    scan_u32(x, cnt.get(), off.get(), 64 * nwaves, tot.get());
    k_srv_lane_tot<<<1, 64, 0, x.stream>>>(off.get(), nwaves, tot.get(), lt.get());
    SK_HIP(hipGetLastError());
    SK_HIP(hipMemcpyAsync(counts, lt.get(), 64 * 8, hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    std::vector<u64> hp(128, 0);
    for (int l = 0; l < 64; l++) {
      if (!((take_mask >> l) & 1)) continue;
      outz[l].reset(x.pool, counts[l] ? counts[l] : 1);
      outw[l].reset(x.pool, counts[l] ? counts[l] : 1);
      hp[l] = (u64)(uintptr_t)outz[l].get();
      hp[64 + l] = (u64)(uintptr_t)outw[l].get();
    }
    SK_HIP(hipMemcpyAsync(ptrs.get(), hp.data(), 128 * 8, hipMemcpyHostToDevice, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
  }
  u64 rows_out = 0;
  for (int l = 0; l < 64; l++) rows_out += counts[l];
  Scope sc(x.timer, "srv_answer", KC_MISC, (double)cell_b * 64 * n_nodes + 12.0 * rows_out, n_nodes);
  k_srv_rows_write<<<nb, kBlock, 0, x.stream>>>(gcd, n_nodes, nwaves, take_mask, clear_mask, fill, clear_maxf,
                                                take_mask ? off.get() : nullptr, (u32* const*)ptrs.get(),
                                                (i64* const*)(ptrs.get() + 64));
  SK_HIP(hipGetLastError());
}

void srv_widen_in_place(const Exec& x, void* cells, u64 n) {
  // ranges [a, b) with 2a >= b, from the top: the u64 words written
  // (bytes [8a, 8b)) never overlap the u32 cells read (bytes [4a, 4b)) nor
  // any cell below a not yet converted; [0, 1) is one thread's read-then-write
  u32* c32 = (u32*)cells;
  u64* c64 = (u64*)cells;
  u64 b = n;
  while (b > 0) {
    const u64 a = b == 1 ? 0 : (b + 1) / 2;
    cells_widen(x, c32 + a, c64 + a, b - a);
    b = a;
  }
}

}  // namespace sk
