// sk_prims.hip — data-parallel primitives for sorted packed-key relations on
// gfx950: device scan, LSD radix sort, run collapse (dedup / Trop min),
// merge-path union and difference, CSR build, pack/unpack, membership.
//
// Every kernel is written for 64-wide wavefronts and 256-thread blocks.
// Global reads and writes are coalesced along the element index; ranking and
// merging happen in LDS.
#include "sk_dev.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace sk {

// ============================================================ device helpers

__device__ __forceinline__ u32 lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ u32 wave_id() { return threadIdx.x >> 6; }
__device__ __forceinline__ u64 lanemask_lt() {
  u32 l = lane_id();
  return l ? (~0ull >> (64 - l)) : 0ull;
}

// Wave-level inclusive scan of u64.
__device__ __forceinline__ u64 wave_incl_scan(u64 v) {
  u32 l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u64 y = __shfl_up(v, o, 64);
    if (l >= (u32)o) v += y;
  }
  return v;
}

// Block-level exclusive scan; `tmp` is __shared__ u64[kWaves]; returns the
// exclusive prefix and the block total.  Must be reached by every thread.
__device__ __forceinline__ u64 block_excl_scan(u64 v, u64* tmp, u64* total) {
  u64 inc = wave_incl_scan(v);
  if (lane_id() == 63) tmp[wave_id()] = inc;
  __syncthreads();
  u64 off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    u64 s = tmp[w];
    if ((u32)w < wave_id()) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Weight order with +inf above every finite value.
__device__ __forceinline__ bool w_less(i64 a, u8 ai, i64 b, u8 bi) {
  if (ai) return false;
  if (bi) return true;
  return a < b;
}

// ===================================================================== fill

__global__ void k_fill_u8(u8* p, u8 v, u64 n) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 stride = (u64)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// One thread per element: the exact block count (the kernels launched with it
// have no grid-stride loop, so the grid is never clipped).
static unsigned grid_for(u64 n, u64 per = kBlock) {
  u64 g = (n + per - 1) / per;
  if (g == 0) g = 1;
  if (g > 0x7fffffffull) throw BudgetError{n};
  return (unsigned)g;
}

__global__ void k_fill_u32(u32* p, u32 v, u64 n) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_u32(const Exec& x, u32* p, u32 v, u64 n) {
  if (!n) return;
  k_fill_u32<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, x.stream>>>(p, v, n);
  SK_HIP(hipGetLastError());
}

void fill_u8(const Exec& x, u8* p, u8 v, u64 n) {
  if (!n) return;
  k_fill_u8<<<grid_for(n, kBlock * 8), kBlock, 0, x.stream>>>(p, v, n);
  SK_HIP(hipGetLastError());
}

void memset_async(const Exec& x, void* p, int v, u64 bytes) { fill_u8(x, (u8*)p, (u8)v, bytes); }

u64 read_u64(const Exec& x, const u64* d) {
  u64 h;
  read_u64s(x, d, &h, 1);
  return h;
}

// The host polls an event behind the copy instead of sleeping in a stream
// synchronisation: these reads sit between dependent launches, where the
// wake-up latency of a blocking wait shows up directly as idle device time.
void read_u64s(const Exec& x, const u64* d, u64* h, int n) {
  if (n < 1 || n > 8) throw std::runtime_error("read_u64s: 1..8 words");
  if (!x.host || !x.read_ev) throw std::runtime_error("read_u64s: execution context without read buffers");
  SK_HIP(hipMemcpyAsync(x.host, d, sizeof(u64) * n, hipMemcpyDeviceToHost, x.stream));
  SK_HIP(hipEventRecord(x.read_ev, x.stream));
  for (;;) {
    const hipError_t q = hipEventQuery(x.read_ev);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) SK_HIP(q);
  }
  for (int i = 0; i < n; i++) h[i] = x.host[i];
  check_fault(x);
}

void check_fault(const Exec& x) {
  if (x.fault_h && *x.fault_h)
    throw DeviceError{hipErrorLaunchFailure, "radix sort: a single-sweep look-back exceeded its spin bound"};
}

// ===================================================================== scan

constexpr int kScanItems = 16;
constexpr u64 kScanTile = (u64)kBlock * kScanItems;  // 4096
constexpr unsigned kScanMaxBlocks = 2048;

template <class T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T* __restrict__ in, u64 n, u64 per,
                                                        u64* __restrict__ sums) {
  __shared__ u64 tmp[kWaves];
  u64 lo = (u64)blockIdx.x * per, hi = min(n, lo + per);
  u64 acc = 0;
  for (u64 i = lo + threadIdx.x; i < hi; i += kBlock) acc += (u64)in[i];
  u64 tot;
  block_excl_scan(acc, tmp, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single-block exclusive scan of block sums in place; writes the grand total.
__global__ __launch_bounds__(kBlock) void k_scan_sums(u64* sums, u32 nb, u64* total) {
  __shared__ u64 tmp[kWaves];
  u64 carry = 0;
  for (u32 base = 0; base < nb; base += kBlock) {
    u32 i = base + threadIdx.x;
    u64 v = i < nb ? sums[i] : 0;
    u64 tot;
    u64 ex = block_excl_scan(v, tmp, &tot);
    if (i < nb) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scan_down(const T* __restrict__ in, u64* __restrict__ out,
                                                      u64 n, u64 per, const u64* __restrict__ offs) {
  __shared__ u64 tmp[kWaves];
  u64 lo = (u64)blockIdx.x * per, hi = min(n, lo + per);
  u64 carry = offs[blockIdx.x];
  for (u64 base = lo; base < hi; base += kBlock) {
    u64 i = base + threadIdx.x;
    u64 v = i < hi ? (u64)in[i] : 0;
    u64 tot;
    u64 ex = block_excl_scan(v, tmp, &tot);
    if (i < hi) out[i] = carry + ex;
    carry += tot;
  }
}

// Small inputs (n <= kScanOneMax): one block of 1024 threads, each scanning
// a contiguous run of at most 16 elements around one block scan -- one
// launch instead of three (neighbouring threads' runs are adjacent, so the
// wave's loads stay within a few cache lines).  Safe in place (each element
// is read before it is written).
constexpr int kScanOneThreads = 1024;
static_assert(kScanOneMax <= (u64)kScanOneThreads * 16, "k_scan_one runs");

__device__ __forceinline__ u64 scan_one_block(u64 v, u64* tmp, u64* total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const u64 inc = wave_incl_scan(v);
  if (lane == 63) tmp[wave] = inc;
  __syncthreads();
  u64 off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanOneThreads / 64; w++) {
    const u64 s = tmp[w];
    if ((u32)w < wave) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

template <class T>
__global__ __launch_bounds__(kScanOneThreads) void k_scan_one(const T* in, u64* out, u64 n, u64* total,
                                                              const u64* side, const u64* extra, int nextra) {
  __shared__ u64 tmp[kScanOneThreads / 64];
  const u64 per = (n + kScanOneThreads - 1) / kScanOneThreads;
  const u64 lo = min(n, (u64)threadIdx.x * per), hi = min(n, lo + per);
  u64 acc = 0, sacc = 0;
  for (u64 i = lo; i < hi; i++) acc += (u64)in[i];
  if (side)
    for (u64 i = lo; i < hi; i++) sacc += side[i];
  u64 tot;
  u64 c = scan_one_block(acc, tmp, &tot);
  for (u64 i = lo; i < hi; i++) {
    const u64 v = (u64)in[i];
    out[i] = c;
    c += v;
  }
  if (side) {
    u64 stot;
    scan_one_block(sacc, tmp, &stot);
    if (threadIdx.x == 0) total[1] = stot;
  }
  if (threadIdx.x == 0 && total) *total = tot;
  if ((int)threadIdx.x < nextra) total[2 + threadIdx.x] = extra[threadIdx.x];
}

void scan_u32_one(const Exec& x, const u32* in, u64* out, u64 n, u64* total, const u64* side, const u64* extra,
                  int nextra) {
  if (n == 0 || n > kScanOneMax) throw std::runtime_error("scan_u32_one: 1..kScanOneMax values");
  if (nextra < 0 || nextra > 6 || (nextra && !total)) throw std::runtime_error("scan_u32_one: extra words");
  Scope sc(x.timer, "scan_one", KC_MISC, (double)n * (4 + 8 + (side ? 8 : 0)), n);
  k_scan_one<u32><<<1, kScanOneThreads, 0, x.stream>>>(in, out, n, total, side, extra, nextra);
  SK_HIP(hipGetLastError());
}

template <class T>
static void scan_impl(const Exec& x, const T* in, u64* out, u64 n, u64* d_total) {
  if (n == 0) {
    if (d_total) memset_async(x, d_total, 0, sizeof(u64));
    return;
  }
  if (n <= kScanOneMax) {
    Scope sc(x.timer, "scan_one", KC_MISC, (double)n * (2 * sizeof(T) + 8), n);
    k_scan_one<T><<<1, kScanOneThreads, 0, x.stream>>>(in, out, n, d_total, nullptr, nullptr, 0);
    SK_HIP(hipGetLastError());
    return;
  }
  u64 nb = (n + kScanTile - 1) / kScanTile;
  if (nb > kScanMaxBlocks) nb = kScanMaxBlocks;
  u64 per = (n + nb - 1) / nb;
  per = (per + kBlock - 1) / kBlock * kBlock;
  nb = (n + per - 1) / per;
  DBuf<u64> sums(x.pool, nb);
  {
    Scope sc(x.timer, "scan_reduce", KC_MISC, (double)n * sizeof(T), n);
    k_scan_reduce<T><<<(unsigned)nb, kBlock, 0, x.stream>>>(in, n, per, sums.get());
  }
  k_scan_sums<<<1, kBlock, 0, x.stream>>>(sums.get(), (u32)nb, d_total);
  {
    Scope sc(x.timer, "scan_down", KC_MISC, (double)n * (sizeof(T) + 8), n);
    k_scan_down<T><<<(unsigned)nb, kBlock, 0, x.stream>>>(in, out, n, per, sums.get());
  }
  SK_HIP(hipGetLastError());
}

void scan_u32(const Exec& x, const u32* in, u64* out, u64 n, u64* t) { scan_impl<u32>(x, in, out, n, t); }
void scan_u64(const Exec& x, const u64* in, u64* out, u64 n, u64* t) { scan_impl<u64>(x, in, out, n, t); }
void scan_u8(const Exec& x, const u8* in, u64* out, u64 n, u64* t) { scan_impl<u8>(x, in, out, n, t); }

// =============================================================== radix sort
//
// One LSD pass over a D-bit digit at `shift` (D = 6, 7 or 8; 2^D <= kBlock
// buckets):
//   k_radix_hist     per-block digit histogram (per-wave LDS counters)
//   scan             exclusive scan of counts in digit-major order
//   k_radix_scatter  per tile of kBlock * IT records: stable in-wave ranking
//                    by ballot match, tile-local counting sort in LDS, then
//                    coalesced runs written to the digit's global position.
// Each block owns a contiguous chunk of whole tiles, so the pass is stable.
// The digit width, the items per thread and the grid cap of each sort come
// from the context's autotuner (RadixTune), or a default when it has none.

// Lanes of the wave (among `valid` lanes) holding the same D-bit digit.
template <int D>
__device__ __forceinline__ u64 match_digit(u32 d, bool valid) {
  u64 m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < D; b++) {
    bool bit = (d >> b) & 1u;
    u64 bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

template <class K, int D>
__device__ __forceinline__ u32 digit_of(K k, int shift) {
  return (u32)(k >> shift) & ((1u << D) - 1u);
}

// Adds the valid lanes' digits to the LDS histogram h: one atomic of 64 when
// the whole wave holds one digit (sorted or clustered keys would otherwise
// send the wave to one address), else one atomic per lane.  The test costs
// one compare; `nvalid` (wave-uniform) counts the valid lanes, a prefix of
// the wave.  Must be reached by every lane.
__device__ __forceinline__ void hist_add(u32* h, u32 d, bool valid, u32 nvalid) {
  const u32 d0 = (u32)__builtin_amdgcn_readfirstlane((int)d);
  const bool uni = nvalid == 64 && __ballot(d != d0) == 0;
  const bool act = uni ? lane_id() == 0 : valid;
  if (act) atomicAdd(&h[uni ? d0 : d], uni ? 64u : 1u);
}

template <class K, int D>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, u64 n, u64 chunk,
                                                       int shift, u32* __restrict__ counts, u32 G) {
  constexpr u32 NB = 1u << D;
  __shared__ u32 h[kWaves][NB];
  for (u32 i = threadIdx.x; i < kWaves * NB; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  u64 lo = (u64)blockIdx.x * chunk, hi = min(n, lo + chunk);
  u32 w = wave_id();
  // kHistBatch keys per thread and step, loaded at clamped addresses so that
  // all of the loads are in flight together; positions past the chunk are
  // masked out of the count
  constexpr int kHistBatch = 8;
  for (u64 base = lo; base < hi; base += (u64)kBlock * kHistBatch) {
    K kv[kHistBatch];
#pragma unroll
    for (int q = 0; q < kHistBatch; q++) {
      const u64 i = base + (u64)q * kBlock + threadIdx.x;
      kv[q] = keys[i < hi ? i : hi - 1];
    }
    // one LDS atomic per key, or one per wave when the wave holds one digit
    // (no D-ballot match)
#pragma unroll
    for (int q = 0; q < kHistBatch; q++) {
      const u64 i0 = base + (u64)q * kBlock + (threadIdx.x & ~63u);
      const u32 nvalid = i0 >= hi ? 0u : (u32)min((u64)64, hi - i0);
      hist_add(h[w], digit_of<K, D>(kv[q], shift), lane_id() < nvalid, nvalid);
    }
  }
  __syncthreads();
  for (u32 d = threadIdx.x; d < NB; d += kBlock) {
    u32 s = 0;
#pragma unroll
    for (int q = 0; q < kWaves; q++) s += h[q][d];
    counts[(u64)d * G + blockIdx.x] = s;
  }
}

template <class K, bool HASW, int D, int IT>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const K* __restrict__ kin, K* __restrict__ kout, const i64* __restrict__ win, i64* __restrict__ wout,
    const u8* __restrict__ fin, u8* __restrict__ fout, u64 n, u64 chunk, int shift,
    const u64* __restrict__ offs, u32 G) {
  constexpr u32 NB = 1u << D;
  static_assert(NB <= (u32)kBlock, "one thread per bucket");
  constexpr int TILE = kBlock * IT;
  __shared__ K s_key[TILE];
  __shared__ i64 s_w[HASW ? TILE : 1];
  __shared__ u8 s_f[HASW ? TILE : 1];
  __shared__ u32 s_wcnt[kWaves][NB];
  __shared__ u32 s_start[NB];
  __shared__ u32 s_tot[NB];
  __shared__ u64 s_gbase[NB];
  __shared__ u64 tmp[kWaves];

  const u32 tid = threadIdx.x, lane = lane_id(), wv = wave_id();
  const u64 lo = (u64)blockIdx.x * chunk, hi = min(n, lo + chunk);
  for (u32 d = tid; d < NB; d += kBlock) s_gbase[d] = offs[(u64)d * G + blockIdx.x];
  const bool has_f = HASW && fin != nullptr;

  for (u64 tb = lo; tb < hi; tb += TILE) {
    for (u32 i = tid; i < kWaves * NB; i += kBlock) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    K key[IT];
    i64 wt[HASW ? IT : 1];
    u8 fl[HASW ? IT : 1];
    u32 rk[IT];
    const u64 wb = tb + (u64)wv * 64 * IT;
    // every key load of the tile first, at clamped addresses, so that they
    // are all in flight together (ranking between loads would serialise them
    // behind its LDS atomics); positions past the chunk are never used.  The
    // weights follow the ranking, in flight while the tile's offsets are
    // scanned.
    const bool full = wb + (u64)IT * 64 <= hi;  // the wave's whole stretch is in the chunk
    if (full) {
#pragma unroll
      for (int it = 0; it < IT; it++) key[it] = kin[wb + (u64)it * 64 + lane];
    } else {
#pragma unroll
      for (int it = 0; it < IT; it++) {
        const u64 idx = wb + (u64)it * 64 + lane;
        key[it] = kin[idx < hi ? idx : hi - 1];
      }
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
      u64 idx = wb + (u64)it * 64 + lane;
      bool valid = idx < hi;
      u32 d = digit_of<K, D>(key[it], shift);
      u64 m = match_digit<D>(d, valid);
      int leader = (m ? (int)__builtin_ctzll(m) : -1);
      u32 old = 0;
      if (valid && (int)lane == leader) old = atomicAdd(&s_wcnt[wv][d], (u32)__popcll(m));
      old = __shfl(old, leader < 0 ? 0 : leader, 64);
      rk[it] = old + (u32)__popcll(m & lanemask_lt());
    }
    if constexpr (HASW) {
      if (full) {
#pragma unroll
        for (int it = 0; it < IT; it++) {
          wt[it] = win[wb + (u64)it * 64 + lane];
          fl[it] = has_f ? fin[wb + (u64)it * 64 + lane] : (u8)0;
        }
      } else {
#pragma unroll
        for (int it = 0; it < IT; it++) {
          const u64 idx = wb + (u64)it * 64 + lane;
          const u64 ic = idx < hi ? idx : hi - 1;
          wt[it] = win[ic];
          fl[it] = has_f ? fin[ic] : (u8)0;
        }
      }
    }
    __syncthreads();
    // digit totals of the tile and per-wave offsets (threads past the last
    // bucket contribute nothing to the scan)
    u32 dsum = 0;
    if (tid < NB) {
      u32 run = 0;
#pragma unroll
      for (int q = 0; q < kWaves; q++) {
        u32 c = s_wcnt[q][tid];
        s_wcnt[q][tid] = run;
        run += c;
      }
      dsum = run;
      s_tot[tid] = run;
    }
    u64 ttot;
    u64 st = block_excl_scan((u64)dsum, tmp, &ttot);
    if (tid < NB) s_start[tid] = (u32)st;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
      u64 idx = wb + (u64)it * 64 + lane;
      if (idx < hi) {
        u32 d = digit_of<K, D>(key[it], shift);
        u32 pos = s_start[d] + s_wcnt[wv][d] + rk[it];
        s_key[pos] = key[it];
        if constexpr (HASW) {
          s_w[pos] = wt[it];
          s_f[pos] = fl[it];
        }
      }
    }
    __syncthreads();
    const u32 tn = (u32)min((u64)TILE, hi - tb);
    for (u32 p = tid; p < tn; p += kBlock) {
      K k = s_key[p];
      u32 d = digit_of<K, D>(k, shift);
      u64 o = s_gbase[d] + (p - s_start[d]);
      kout[o] = k;
      if constexpr (HASW) {
        wout[o] = s_w[p];
        if (has_f) fout[o] = s_f[p];
      }
    }
    __syncthreads();
    if (tid < NB) s_gbase[tid] += s_tot[tid];
    __syncthreads();
  }
}

// ============================================================ single-sweep radix
//
// The pass structure of the candidates marked `sweep`:
//   1. k_os_hist reads the keys once and counts the digits of every pass
//      (per block in LDS, then one global add per (pass, digit) and block);
//   2. k_os_scan turns each pass's counts into its bucket starts;
//   3. each pass is one launch of k_os_pass, whose tiles take their ids in
//      launch order from a counter and rank their records in LDS.  A tile
//      finds how many records of each digit the earlier tiles hold by
//      looking back over their published counts:
//        - at once it publishes its own digit counts (status AGG); tile 0
//          publishes them as a prefix (status PRE);
//        - one thread per digit walks back from the predecessor, adding AGG
//          counts, until it reads a PRE, then publishes PRE with the sum plus
//          its own count.  kOsWindow predecessors are read per step.
//      A tile waits only on tiles with smaller ids, which took their ids
//      earlier and are running, and those publish AGG without waiting on
//      anything, so every walk ends.  The spin bound only guards against a
//      defect: past it the tile stops waiting and sets the context's fault
//      word, which the host turns into an error (an undercounted prefix keeps
//      every write in bounds).
// Status and count share one 64-bit word, written and read with relaxed
// device-scope atomics: nothing else is ordered by them, so no fence and no
// L2 write-back is needed; a publication is one store and a look-back step
// one load served below the XCDs' L2s.  Each pass clears, row by row, the
// words the next pass publishes into (two word arrays alternate).

constexpr u64 kOsAgg = 1ull << 62, kOsPre = 2ull << 62, kOsVal = (1ull << 62) - 1;
#ifndef SK_OS_WINDOW
#define SK_OS_WINDOW 4
#endif
constexpr int kOsWindow = SK_OS_WINDOW;
constexpr u32 kOsSpinBound = 1u << 22;   // sleeping retries of one look-back (~0.1 s) before it gives up
constexpr u32 kOsMaxSlots = 4096;        // passes x buckets counted by k_os_hist (128-bit keys at 8 bits: 16 x 256)
constexpr u64 kOsMaxTiles = 1ull << 16;  // tiles of a pass (their published words: tiles x buckets x 16 B)
constexpr int kOsHistBlock = 1024;

struct OsPass {
  u64* flags;         // this pass: tiles x buckets words
  u64* flags_next;    // the next pass's words, cleared row by row
  u32* tile_ctr;      // this pass's tile counter
  const u64* gstart;  // this pass's bucket starts
  u32* fault;         // pinned host word (device pointer)
};

// Digit counts of every pass: per block in LDS (hist_add), then one global
// add per (pass, bucket) and block.
template <class K, int D>
__global__ __launch_bounds__(kOsHistBlock) void k_os_hist(const K* __restrict__ keys, u64 n, int lo_bit, int P,
                                                          unsigned long long* __restrict__ counts) {
  constexpr u32 NB = 1u << D;
  __shared__ u32 h[kOsMaxSlots];
  const u32 slots = (u32)P * NB;
  const u32 lane = lane_id();
  for (u32 i = threadIdx.x; i < slots; i += kOsHistBlock) h[i] = 0;
  __syncthreads();
  constexpr int kB = 4;  // keys per thread in flight
  const u64 step = (u64)gridDim.x * kOsHistBlock * kB;
  for (u64 base = (u64)blockIdx.x * kOsHistBlock * kB; base < n; base += step) {
    K kv[kB];
    u32 nv[kB];
#pragma unroll
    for (int q = 0; q < kB; q++) {
      const u64 i = base + (u64)q * kOsHistBlock + threadIdx.x;
      kv[q] = keys[i < n ? i : n - 1];
      const u64 i0 = i - lane;  // the wave's first position: valid lanes are a prefix
      nv[q] = i0 >= n ? 0u : (u32)min((u64)64, n - i0);
    }
    for (int p = 0; p < P; p++) {
      const int sh = lo_bit + D * p;
#pragma unroll
      for (int q = 0; q < kB; q++) hist_add(h + p * NB, digit_of<K, D>(kv[q], sh), lane < nv[q], nv[q]);
    }
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < slots; i += kOsHistBlock)
    if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
}

// Bucket starts of pass blockIdx.x: exclusive scan of its NB counts (one
// thread per bucket; NB is a multiple of 64).
template <int D>
__global__ __launch_bounds__(256) void k_os_scan(const u64* __restrict__ counts, u64* __restrict__ gstart) {
  constexpr u32 NB = 1u << D;
  __shared__ u64 part[4];
  const u32 d = threadIdx.x;
  const u64 v = counts[(u64)blockIdx.x * NB + d];
  const u64 inc = wave_incl_scan(v);
  if (lane_id() == 63) part[wave_id()] = inc;
  __syncthreads();
  u64 off = 0;
  for (u32 q = 0; q < wave_id(); q++) off += part[q];
  gstart[(u64)blockIdx.x * NB + d] = off + inc - v;
}

// 512-thread blocks are held to 128 VGPRs (four waves per SIMD), so that two
// blocks fit a CU whenever their LDS does.
template <class K, bool HASW, int D, int IT, int NT>
__global__ __launch_bounds__(NT, NT == 512 ? 4 : 1) void k_os_pass(const K* __restrict__ kin, K* __restrict__ kout,
                                                const i64* __restrict__ win, i64* __restrict__ wout,
                                                const u8* __restrict__ fin, u8* __restrict__ fout, u64 n, int shift,
                                                OsPass ps) {
  constexpr u32 NB = 1u << D;
  constexpr int NW = NT / 64;
  constexpr u32 TILE = (u32)NT * IT;
  static_assert(NB <= (u32)NT && NB <= 256 && NB % 64 == 0, "one thread per bucket");
  static_assert(sizeof(K) >= sizeof(i64), "weights reuse the key stage");
  static_assert(TILE <= 65535, "16-bit per-wave offsets");
  __shared__ K s_buf[TILE];
  __shared__ u8 s_dig[HASW ? TILE : 1];
  __shared__ u16 s_wcnt[NW][NB];  // per-wave digit counts, then offsets (< TILE <= 2^16)
  __shared__ u32 s_start[NB];
  __shared__ u64 s_gbase[NB];
  __shared__ u64 s_part[4];
  __shared__ u32 s_tile;

  // per-wave peer masks of the ranking, in the staging buffer (idle until
  // the keys are staged): entry d holds the lanes of the current step whose
  // digit is d
  u64 (*s_peer)[NB] = reinterpret_cast<u64 (*)[NB]>(s_buf);
  static_assert((u64)NW * NB * 8 <= (u64)TILE * sizeof(K), "peer masks fit the stage");
  const u32 tid = threadIdx.x, lane = lane_id(), wv = wave_id();
  if (tid == 0) s_tile = atomicAdd(ps.tile_ctr, 1u);
  for (u32 i = tid; i < NW * NB; i += NT) {
    (&s_wcnt[0][0])[i] = 0;
    (&s_peer[0][0])[i] = 0;
  }
  __syncthreads();
  const u32 tile = s_tile;
  const u64 tb = (u64)tile * TILE;
  const u64 hi = min(n, tb + TILE);  // > tb: the grid has exactly ceil(n / TILE) tiles
  const u64 wb = tb + (u64)wv * 64 * IT;
  // valid positions of the wave's stretch (wave-uniform): item it of this
  // lane is position it * 64 + lane of the stretch
  const u32 nvw = hi > wb ? (u32)min(hi - wb, (u64)IT * 64) : 0u;
  const bool has_f = HASW && fin != nullptr;
  const K* kw = kin + wb;
  K key[IT];
  u32 rk[IT];
  // every key load of the wave's stretch first, at clamped addresses
  if (nvw == IT * 64) {
#pragma unroll
    for (int it = 0; it < IT; it++) key[it] = kw[it * 64 + lane];
  } else {
    const u32 last = nvw ? nvw - 1 : 0;  // a wave past the tile's end reads a valid record it never uses
    const K* kc = nvw ? kw : kin + (hi - 1);
#pragma unroll
    for (int it = 0; it < IT; it++) key[it] = kc[min((u32)(it * 64) + lane, last)];
  }
  // stable rank within the wave's digit group; per-wave digit counts in LDS.
  // Per step every valid lane ORs its bit into its digit's peer mask, and
  // reads the mask and the digit's count so far; the lowest lane of each
  // mask then adds the mask's size to the count and clears the mask.  A
  // wave's LDS operations complete in issue order; the fences keep the
  // compiler from reordering them.
  const u64 mybit = 1ull << lane;
  const u64 below = lanemask_lt();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const bool valid = (u32)(it * 64) + lane < nvw;
    const u32 d = digit_of<K, D>(key[it], shift);
    if (valid) atomicOr((unsigned long long*)&s_peer[wv][d], (unsigned long long)mybit);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    u64 m = 0;
    u32 old = 0;
    if (valid) {
      m = s_peer[wv][d];
      old = s_wcnt[wv][d];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (valid && (m & below) == 0) {  // the lowest lane of its digit
      s_peer[wv][d] = 0;
      s_wcnt[wv][d] = (u16)(old + (u32)__popcll(m));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    rk[it] = old + (u32)__popcll(m & below);
  }
  __syncthreads();
  // the tile's digit totals, published at once; per-wave offsets
  u32 tot = 0;
  if (tid < NB) {
    u32 run = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) {
      const u32 c = s_wcnt[q][tid];
      s_wcnt[q][tid] = (u16)run;
      run += c;
    }
    tot = run;
    __hip_atomic_store(&ps.flags[(u64)tile * NB + tid], (tile ? kOsAgg : kOsPre) | (u64)tot, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    ps.flags_next[(u64)tile * NB + tid] = 0;
  }
  // exclusive scan of the totals over the digits (waves 0 .. NB/64 - 1)
  u64 inc = 0;
  if (tid < NB) {
    inc = wave_incl_scan((u64)tot);
    if (lane == 63) s_part[wv] = inc;
  }
  __syncthreads();
  if (tid < NB) {
    u64 off = 0;
    for (u32 q = 0; q < wv; q++) off += s_part[q];
    s_start[tid] = (u32)(off + inc - tot);
    // look back over the earlier tiles for this digit
    u64 excl = 0;
    if (tile) {
      long j = (long)tile - 1;
      u32 spins = 0;
      for (;;) {
        u64 v[kOsWindow];
#pragma unroll
        for (int q = 0; q < kOsWindow; q++)
          v[q] = j - q >= 0 ? __hip_atomic_load(&ps.flags[(u64)(j - q) * NB + tid], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : kOsPre;
        int q = 0;
        bool done = false;
        for (; q < kOsWindow; q++) {
          const u64 st = v[q] >> 62;
          if (!st) break;
          excl += v[q] & kOsVal;
          if (st == 2) {
            done = true;
            break;
          }
        }
        if (done) break;
        j -= q;
        if (q < kOsWindow) {
          if (++spins > kOsSpinBound) {
            __hip_atomic_store(ps.fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __hip_atomic_store(&ps.flags[(u64)tile * NB + tid], kOsPre | (excl + tot), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[tid] = ps.gstart[tid] + excl;
  }
  __syncthreads();
  // stage the keys in digit order, then write each digit's run
#pragma unroll
  for (int it = 0; it < IT; it++) {
    if ((u32)(it * 64) + lane < nvw) {
      const u32 d = digit_of<K, D>(key[it], shift);
      s_buf[s_start[d] + s_wcnt[wv][d] + rk[it]] = key[it];
    }
  }
  __syncthreads();
  const u32 tn = (u32)(hi - tb);
  if (tn == TILE) {
    // a full tile: every thread writes IT records, four at a time with their
    // LDS reads in flight together
#pragma unroll 4
    for (int it = 0; it < IT; it++) {
      const u32 p = tid + (u32)it * NT;
      const K k = s_buf[p];
      const u32 d = digit_of<K, D>(k, shift);
      kout[s_gbase[d] + (p - s_start[d])] = k;
      if constexpr (HASW) s_dig[p] = (u8)d;
    }
  } else {
    for (u32 p = tid; p < tn; p += NT) {
      const K k = s_buf[p];
      const u32 d = digit_of<K, D>(k, shift);
      kout[s_gbase[d] + (p - s_start[d])] = k;
      if constexpr (HASW) s_dig[p] = (u8)d;
    }
  }
  if constexpr (HASW) {
    // the weights (and flags) are read only now, straight into their ranked
    // stage positions: no registers hold them across the ranking and the
    // look-back, which leaves room for a second tile per CU
    __syncthreads();
    i64* s_w = reinterpret_cast<i64*>(s_buf);
    const i64* ww = win + wb;
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const u32 li = (u32)(it * 64) + lane;
      if (li < nvw) {
        const u32 d = digit_of<K, D>(key[it], shift);
        s_w[s_start[d] + s_wcnt[wv][d] + rk[it]] = ww[li];
      }
    }
    __syncthreads();
    if (tn == TILE) {
#pragma unroll 4
      for (int it = 0; it < IT; it++) {
        const u32 p = tid + (u32)it * NT;
        const u32 d = s_dig[p];
        wout[s_gbase[d] + (p - s_start[d])] = s_w[p];
      }
    } else {
      for (u32 p = tid; p < tn; p += NT) {
        const u32 d = s_dig[p];
        wout[s_gbase[d] + (p - s_start[d])] = s_w[p];
      }
    }
    if (has_f) {
      __syncthreads();
      u8* s_f = reinterpret_cast<u8*>(s_buf);
      const u8* fw = fin + wb;
#pragma unroll
      for (int it = 0; it < IT; it++) {
        const u32 li = (u32)(it * 64) + lane;
        if (li < nvw) {
          const u32 d = digit_of<K, D>(key[it], shift);
          s_f[s_start[d] + s_wcnt[wv][d] + rk[it]] = fw[li];
        }
      }
      __syncthreads();
      for (u32 p = tid; p < tn; p += NT) {
        const u32 d = s_dig[p];
        fout[s_gbase[d] + (p - s_start[d])] = s_f[p];
      }
    }
  }
}

// The candidate configurations the tuner times: digit bits x items per
// thread x threads per block, by reduce-then-scan passes (a histogram launch
// and a scatter launch per pass, kBlock threads) or single-sweep passes
// (u128 keys take half the items, so that a tile with weights fits LDS).
struct RadixCand {
  int digit, items, block;
  bool sweep;
};
constexpr RadixCand kRadixCands[] = {
    {8, 16, kBlock, false}, {8, 8, kBlock, false}, {7, 16, kBlock, false}, {7, 8, kBlock, false},
    {6, 16, kBlock, false}, {6, 8, kBlock, false}, {8, 32, kBlock, false}, {6, 32, kBlock, false},
    {8, 16, 256, true},     {8, 32, 256, true},    {8, 16, 512, true},     {8, 8, 512, true},
    {8, 8, 1024, true},     {7, 16, 512, true},    {6, 16, 512, true},     {8, 32, 512, true},
    {8, 16, 1024, true}};
constexpr int kNCands = sizeof(kRadixCands) / sizeof(kRadixCands[0]);
static_assert(kNCands <= RadixTune::kMaxCands, "tuner table width");
// The reduce-then-scan candidate a sweep candidate falls back to when a sort
// has more tiles than the sweep's published words allow.
constexpr int kOsFallback[] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 6, 0, 1, 1, 2, 4, 6, 0};
static_assert(sizeof(kOsFallback) / sizeof(int) == kNCands, "fallback per candidate");

template <class K>
constexpr int items_for(int c) {
  return sizeof(K) == 16 ? kRadixCands[c].items / 2 : kRadixCands[c].items;
}

// Algorithmic bytes: the histogram reads every key; the scatter reads and
// writes every record (rec bytes).
template <class K, bool HASW, int D, int IT>
static void radix_pass(const Exec& x, const K* a, K* b, const i64* wa, i64* wb, const u8* fa, u8* fb, u64 n, u64 G,
                       u64 chunk, int shift, u32* counts, u64* offs, double rec) {
  constexpr u64 NB = 1ull << D;
  {
    Scope sc(x.timer, "radix_hist", KC_SORT, (double)n * sizeof(K), n);
    k_radix_hist<K, D><<<(unsigned)G, kBlock, 0, x.stream>>>(a, n, chunk, shift, counts, (u32)G);
    SK_HIP(hipGetLastError());
  }
  scan_u32(x, counts, offs, NB * G, nullptr);
  Scope sc(x.timer, "radix_scatter", KC_SORT, 2.0 * rec * (double)n, n);
  k_radix_scatter<K, HASW, D, IT><<<(unsigned)G, kBlock, 0, x.stream>>>(a, b, wa, wb, fa, fb, n, chunk, shift, offs,
                                                                        (u32)G);
  SK_HIP(hipGetLastError());
}

template <class K, bool HASW>
static void radix_pass_c(int c, const Exec& x, const K* a, K* b, const i64* wa, i64* wb, const u8* fa, u8* fb, u64 n,
                         u64 G, u64 chunk, int shift, u32* counts, u64* offs, double rec) {
  switch (c) {
#define SK_RADIX_CASE(I)                                                                                       \
  case I:                                                                                                      \
    radix_pass<K, HASW, kRadixCands[I].digit, items_for<K>(I)>(x, a, b, wa, wb, fa, fb, n, G, chunk, shift,  \
                                                               counts, offs, rec);                             \
    break;
    SK_RADIX_CASE(0)
    SK_RADIX_CASE(1)
    SK_RADIX_CASE(2)
    SK_RADIX_CASE(3)
    SK_RADIX_CASE(4)
    SK_RADIX_CASE(5)
    SK_RADIX_CASE(6)
    SK_RADIX_CASE(7)
#undef SK_RADIX_CASE
    default:
      throw std::runtime_error("radix: configuration");
  }
}

static int size_class(u64 n) {
  int c = 0;
  while (c < RadixTune::kClasses - 1 && (n >> (RadixTune::kMinLog2 + 2 * c + 1)) > 0) c++;
  return c;
}

// Grid of a pass: chunks of whole tiles, at most `gcap` blocks.
static void radix_grid(u64 n, u64 tile, u64 gcap, u64* G, u64* chunk) {
  u64 g = (n + tile - 1) / tile;
  if (g > gcap) g = gcap;
  u64 ch = (n + g - 1) / g;
  ch = (ch + tile - 1) / tile * tile;
  *chunk = ch;
  *G = (n + ch - 1) / ch;
}

// One single-sweep pass of candidate I: one tile per ceil(n / tile) blocks.
template <class K, bool HASW, int I>
static void os_pass(const Exec& x, const K* a, K* b, const i64* wa, i64* wb, const u8* fa, u8* fb, u64 n, int shift,
                    const OsPass& ps, double rec) {
  constexpr int D = kRadixCands[I].digit, IT = items_for<K>(I), NT = kRadixCands[I].block;
  const u64 tiles = (n + (u64)NT * IT - 1) / ((u64)NT * IT);
  Scope sc(x.timer, "radix_sweep", KC_SORT, 2.0 * rec * (double)n, n);
  k_os_pass<K, HASW, D, IT, NT><<<(unsigned)tiles, NT, 0, x.stream>>>(a, b, wa, wb, fa, fb, n, shift, ps);
  SK_HIP(hipGetLastError());
}

template <class K, bool HASW>
static void os_pass_c(int c, const Exec& x, const K* a, K* b, const i64* wa, i64* wb, const u8* fa, u8* fb, u64 n,
                      int shift, const OsPass& ps, double rec) {
  switch (c) {
#define SK_OS_CASE(I)                                                           \
  case I:                                                                       \
    os_pass<K, HASW, I>(x, a, b, wa, wb, fa, fb, n, shift, ps, rec);            \
    break;
    SK_OS_CASE(8)
    SK_OS_CASE(9)
    SK_OS_CASE(10)
    SK_OS_CASE(11)
    SK_OS_CASE(12)
    SK_OS_CASE(13)
    SK_OS_CASE(14)
    SK_OS_CASE(15)
    SK_OS_CASE(16)
#undef SK_OS_CASE
    default:
      throw std::runtime_error("radix: sweep configuration");
  }
}

template <class K>
static void os_hist_c(int D, const Exec& x, const K* keys, u64 n, int lo_bit, int P, u64* counts) {
  const u64 per = (u64)kOsHistBlock * 4;
  const unsigned g = (unsigned)std::max<u64>(1, std::min<u64>(256, (n + per - 1) / per));
  Scope sc(x.timer, "radix_sweep_hist", KC_SORT, (double)n * sizeof(K), n);
  if (D == 8) k_os_hist<K, 8><<<g, kOsHistBlock, 0, x.stream>>>(keys, n, lo_bit, P, (unsigned long long*)counts);
  else if (D == 7) k_os_hist<K, 7><<<g, kOsHistBlock, 0, x.stream>>>(keys, n, lo_bit, P, (unsigned long long*)counts);
  else k_os_hist<K, 6><<<g, kOsHistBlock, 0, x.stream>>>(keys, n, lo_bit, P, (unsigned long long*)counts);
  SK_HIP(hipGetLastError());
}

static void os_scan(int D, const Exec& x, const u64* counts, u64* gstart, int P) {
  if (D == 8) k_os_scan<8><<<(unsigned)P, 256, 0, x.stream>>>(counts, gstart);
  else if (D == 7) k_os_scan<7><<<(unsigned)P, 128, 0, x.stream>>>(counts, gstart);
  else k_os_scan<6><<<(unsigned)P, 64, 0, x.stream>>>(counts, gstart);
  SK_HIP(hipGetLastError());
}

// Single-sweep sort of candidate c: false (nothing launched) when the sort
// has more tiles than the published words allow, or the context has no
// fault word.
template <class K>
static bool radix_run_sweep(const Exec& x, int c, K* keys, K* keys_alt, i64* w, i64* w_alt, u8* inf, u8* inf_alt,
                            u64 n, int key_bits, int lo_bit, K** out_keys, i64** out_w, u8** out_inf) {
  const int D = kRadixCands[c].digit;
  const u64 NB = 1ull << D;
  const u64 tile = (u64)kRadixCands[c].block * items_for<K>(c);
  const u64 tiles = (n + tile - 1) / tile;
  const int P = (key_bits + D - 1) / D;
  if (tiles > kOsMaxTiles || (u64)P * NB > kOsMaxSlots || !x.fault_d) return false;
  const bool hasw = (w != nullptr);
  const double rec = sizeof(K) + (hasw ? 8.0 : 0) + (inf ? 1.0 : 0);
  // [counts P*NB][tile counters P][flags 0: tiles*NB] are cleared here;
  // [flags 1: tiles*NB][bucket starts P*NB] are written before they are read
  const u64 wc = (u64)P * NB, wt = ((u64)P + 1) / 2, wf = tiles * NB;
  DBuf<u64> st(x.pool, wc + wt + 2 * wf + wc);
  u64* counts = st.get();
  u32* ctr = (u32*)(counts + wc);
  u64* fl0 = counts + wc + wt;
  u64* fl1 = fl0 + wf;
  u64* gstart = fl1 + wf;
  fill_u32(x, (u32*)counts, 0, 2 * (wc + wt + wf));
  os_hist_c<K>(D, x, keys, n, lo_bit, P, counts);
  os_scan(D, x, counts, gstart, P);
  K* a = keys;
  K* b = keys_alt;
  i64* wa = w;
  i64* wb = w_alt;
  u8* fa = inf;
  u8* fb = inf_alt;
  for (int p = 0; p < P; p++) {
    OsPass ps{(p & 1) ? fl1 : fl0, (p & 1) ? fl0 : fl1, ctr + p, gstart + (u64)p * NB, x.fault_d};
    const int shift = lo_bit + D * p;
    if (hasw) os_pass_c<K, true>(c, x, a, b, wa, wb, fa, fb, n, shift, ps, rec);
    else os_pass_c<K, false>(c, x, a, b, nullptr, nullptr, nullptr, nullptr, n, shift, ps, rec);
    std::swap(a, b);
    std::swap(wa, wb);
    std::swap(fa, fb);
  }
  *out_keys = a;
  if (out_w) *out_w = wa;
  if (out_inf) *out_inf = fa;
  return true;
}

template <class K>
static void radix_run(const Exec& x, int c, K* keys, K* keys_alt, i64* w, i64* w_alt, u8* inf, u8* inf_alt, u64 n,
                      int key_bits, int lo_bit, K** out_keys, i64** out_w, u8** out_inf) {
  if (kRadixCands[c].sweep) {
    if (radix_run_sweep<K>(x, c, keys, keys_alt, w, w_alt, inf, inf_alt, n, key_bits, lo_bit, out_keys, out_w,
                           out_inf))
      return;
    c = kOsFallback[c];
  }
  const int D = kRadixCands[c].digit;
  const u64 tile = (u64)kBlock * items_for<K>(c);
  const bool hasw = (w != nullptr);
  const int passes = (key_bits + D - 1) / D;
  const double rec = sizeof(K) + (hasw ? 8.0 : 0) + (inf ? 1.0 : 0);
  u64 G, chunk;
  radix_grid(n, tile, 2048, &G, &chunk);
  DBuf<u32> counts(x.pool, ((u64)1 << D) * G);
  DBuf<u64> offs(x.pool, ((u64)1 << D) * G);
  K* a = keys;
  K* b = keys_alt;
  i64* wa = w;
  i64* wb = w_alt;
  u8* fa = inf;
  u8* fb = inf_alt;
  for (int p = 0; p < passes; p++) {
    const int shift = lo_bit + D * p;
    if (hasw) radix_pass_c<K, true>(c, x, a, b, wa, wb, fa, fb, n, G, chunk, shift, counts.get(), offs.get(), rec);
    else radix_pass_c<K, false>(c, x, a, b, nullptr, nullptr, nullptr, nullptr, n, G, chunk, shift, counts.get(),
                                offs.get(), rec);
    std::swap(a, b);
    std::swap(wa, wb);
    std::swap(fa, fb);
  }
  *out_keys = a;
  if (out_w) *out_w = wa;
  if (out_inf) *out_inf = fa;
}

__global__ void k_tune_keys(u64* k, u64 n, u64 seed) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64 z = seed + i * 0x9e3779b97f4a7c15ull;  // splitmix64
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  k[i] = z ^ (z >> 31);
}

// Times a two-pass sort under every candidate on random keys of the class's
// size (median of three, HIP events on the context's stream) and records
// half of it as the candidate's per-pass time; a sort then takes the
// candidate with the least passes x time for its key width.
template <class K>
static void radix_tune_class(const Exec& x, RadixTune* t, bool hasw, int cls) {
  const u64 n = 1ull << (RadixTune::kMinLog2 + 2 * cls);
  const u64 words = n * sizeof(K) / 8;
  DBuf<u64> k1(x.pool, words), k2(x.pool, words);
  DBuf<i64> w1(x.pool, hasw ? n : 1), w2(x.pool, hasw ? n : 1);
  k_tune_keys<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, x.stream>>>(k1.get(), words, 0x5eed + cls);
  if (hasw) k_tune_keys<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, x.stream>>>((u64*)w1.get(), n, 77);
  SK_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  SK_HIP(hipEventCreate(&e0));
  SK_HIP(hipEventCreate(&e1));
  Exec xq = x;
  xq.timer = nullptr;
  for (int c = 0; c < kNCands; c++) {
    float ms[3];
    for (int r = 0; r < 3; r++) {
      K* ok;
      i64* ow;
      u8* of;
      // fresh random keys for every sort: a sort leaves k1 ordered on the
      // timed digits, which would favour some candidates over others
      k_tune_keys<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, x.stream>>>(k1.get(), words,
                                                                                      0x5eed + 64 * cls + 3 * c + r);
      SK_HIP(hipGetLastError());
      SK_HIP(hipEventRecord(e0, x.stream));
      radix_run<K>(xq, c, (K*)k1.get(), (K*)k2.get(), hasw ? w1.get() : nullptr, hasw ? w2.get() : nullptr, nullptr,
                   nullptr, n, 2 * kRadixCands[c].digit, 0, &ok, &ow, &of);
      SK_HIP(hipEventRecord(e1, x.stream));
      SK_HIP(hipEventSynchronize(e1));
      SK_HIP(hipEventElapsedTime(&ms[r], e0, e1));
      ms[r] *= 0.5f;  // per pass (a sweep's histogram shared by both)
    }
    check_fault(x);
    std::sort(ms, ms + 3);
    t->pass_ms[sizeof(K) == 16][hasw][cls][c] = ms[1];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t->tuned[sizeof(K) == 16][hasw][cls] = true;
}

// Per-pass timings are a property of the device, so every context of the
// process shares them: a class timed by one context (or by
// sk_ctx_set_autotune) is copied into the next context that needs it
// instead of being timed again.
namespace {
struct TuneRow {
  float ms[RadixTune::kMaxCands];
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, TuneRow> g_tune;  // (device, wide, weighted, class)

int current_device() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return dev;
}

void tune_publish(const RadixTune* t, int kw, int hw, int cls) {
  TuneRow r;
  std::memcpy(r.ms, t->pass_ms[kw][hw][cls], sizeof(r.ms));
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[std::make_tuple(current_device(), kw, hw, cls)] = r;
}

bool tune_fetch(RadixTune* t, int kw, int hw, int cls) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune.find(std::make_tuple(current_device(), kw, hw, cls));
  if (it == g_tune.end()) return false;
  std::memcpy(t->pass_ms[kw][hw][cls], it->second.ms, sizeof(it->second.ms));
  t->tuned[kw][hw][cls] = true;
  return true;
}
}  // namespace

template <class K>
static int radix_choose(const Exec& x, u64 n, int key_bits, bool hasw) {
  RadixTune* t = x.tune;
  if (!t || n < (1ull << RadixTune::kMinLog2)) return t && t->forced >= 0 ? t->forced : 0;
  if (t->forced >= 0) return t->forced;
  const int cls = size_class(n);
  const int kw = sizeof(K) == 16;
  if (!t->tuned[kw][hasw][cls] && !tune_fetch(t, kw, hasw, cls)) {
    if (!t->lazy) return 0;
    radix_tune_class<K>(x, t, hasw, cls);
    tune_publish(t, kw, hasw, cls);
  }
  int best = 0;
  double bt = 1e30;
  for (int c = 0; c < kNCands; c++) {
    const int D = kRadixCands[c].digit;
    const double tt = ((key_bits + D - 1) / D) * (double)t->pass_ms[kw][hasw][cls][c];
    if (tt < bt) {
      bt = tt;
      best = c;
    }
  }
  t->chosen[kw][hasw][cls] = best;
  return best;
}

template <class K>
void radix_sort(const Exec& x, K* keys, K* keys_alt, i64* w, i64* w_alt, u8* inf, u8* inf_alt, u64 n,
                int key_bits, K** out_keys, i64** out_w, u8** out_inf, int lo_bit) {
  *out_keys = keys;
  if (out_w) *out_w = w;
  if (out_inf) *out_inf = inf;
  if (n <= 1 || key_bits <= 0) return;
  const int c = radix_choose<K>(x, n, key_bits, w != nullptr);
  radix_run<K>(x, c, keys, keys_alt, w, w_alt, inf, inf_alt, n, key_bits, lo_bit, out_keys, out_w, out_inf);
}

void radix_autotune(const Exec& x, RadixTune* t) {
  for (int cls = 0; cls < RadixTune::kClasses; cls++)
    for (int hasw = 0; hasw < 2; hasw++) {
      radix_tune_class<u64>(x, t, hasw != 0, cls);
      tune_publish(t, 0, hasw, cls);
      radix_tune_class<u128>(x, t, hasw != 0, cls);
      tune_publish(t, 1, hasw, cls);
    }
}

int radix_candidate_digit(int c) { return c >= 0 && c < kNCands ? kRadixCands[c].digit : 0; }
int radix_candidate_items(int c, bool wide) {
  if (c < 0 || c >= kNCands) return 0;
  return wide ? items_for<u128>(c) : items_for<u64>(c);
}
int radix_candidates() { return kNCands; }
int radix_candidate_block(int c) { return c >= 0 && c < kNCands ? kRadixCands[c].block : 0; }
bool radix_candidate_sweep(int c) { return c >= 0 && c < kNCands && kRadixCands[c].sweep; }

// ============================================================ run collapse

template <class K>
__global__ __launch_bounds__(kBlock) void k_runs_count(const K* __restrict__ keys, u64 n, u64 per,
                                                       u64* __restrict__ cnt) {
  __shared__ u64 tmp[kWaves];
  u64 lo = (u64)blockIdx.x * per, hi = min(n, lo + per);
  u64 c = 0;
  for (u64 i = lo + threadIdx.x; i < hi; i += kBlock) c += (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
  u64 tot;
  block_excl_scan(c, tmp, &tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

template <class K, bool HASW>
__global__ __launch_bounds__(kBlock) void k_runs_write(const K* __restrict__ keys, const i64* __restrict__ w,
                                                       const u8* __restrict__ f, u64 n, u64 per,
                                                       const u64* __restrict__ offs, K* __restrict__ ok,
                                                       i64* __restrict__ ow, u8* __restrict__ of) {
  __shared__ u64 tmp[kWaves];
  u64 lo = (u64)blockIdx.x * per, hi = min(n, lo + per);
  u64 carry = offs[blockIdx.x];
  for (u64 base = lo; base < hi; base += kBlock) {
    u64 i = base + threadIdx.x;
    bool first = false;
    K k = 0;
    if (i < hi) {
      k = keys[i];
      first = (i == 0 || keys[i - 1] != k);
    }
    u64 tot;
    u64 ex = block_excl_scan(first ? 1 : 0, tmp, &tot);
    if (first) {
      u64 o = carry + ex;
      ok[o] = k;
      if constexpr (HASW) {
        i64 bw = w[i];
        u8 bf = f ? f[i] : (u8)0;
        for (u64 j = i + 1; j < n && keys[j] == k; j++) {
          i64 cw = w[j];
          u8 cf = f ? f[j] : (u8)0;
          if (w_less(cw, cf, bw, bf)) {
            bw = cw;
            bf = cf;
          }
        }
        ow[o] = bw;
        if (of) of[o] = bf;
      }
    }
    carry += tot;
  }
}

// ---- weighted runs: segmented minimum without serial run walks
//
// Pass 1 (k_runs_sum_w): per chunk, the number of run heads, the minimum of
// the elements before the chunk's first head (the tail of a run that started
// earlier) and the minimum of the run still open at the chunk's end.
// Carry (k_runs_carry): one thread walks the chunks in order and produces,
// per chunk, the minimum carried into it by a run that started before it,
// and the number of heads before it.
// Pass 2 (k_runs_write_w): each 256-element tile does a segmented inclusive
// min-scan (wave shuffles + LDS across waves) seeded with the carry; the last
// element of every run writes (key, run minimum) at the run's rank.

struct WMin {
  i64 w;
  u8 f;  // +inf
};

__device__ __forceinline__ WMin wmin(WMin a, WMin b) { return w_less(b.w, b.f, a.w, a.f) ? b : a; }

template <class K>
__global__ __launch_bounds__(kBlock) void k_runs_sum_w(const K* __restrict__ keys, const i64* __restrict__ w,
                                                       const u8* __restrict__ f, u64 n, u64 per,
                                                       u64* __restrict__ heads, i64* __restrict__ pre_w,
                                                       u8* __restrict__ pre_f, i64* __restrict__ tail_w,
                                                       u8* __restrict__ tail_f, u64* __restrict__ first_head) {
  __shared__ u64 tmp[kWaves];
  __shared__ u64 s_first[kWaves], s_last[kWaves];
  __shared__ i64 s_pw[kWaves], s_tw[kWaves];
  __shared__ u8 s_pf[kWaves], s_tf[kWaves];
  const u64 lo = (u64)blockIdx.x * per, hi = min(n, lo + per);
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u64 nh = 0;
  u64 firsth = hi;              // position of the chunk's first head
  WMin pre{0, 1}, tail{0, 1};   // +inf identities
  for (u64 base = lo; base < hi; base += kBlock) {
    u64 i = base + threadIdx.x;
    bool valid = i < hi;
    bool head = valid && (i == 0 || keys[i] != keys[i - 1]);
    WMin v{valid ? w[i] : 0, (u8)(valid ? (f ? f[i] : 0) : 1)};
    u64 hb = __ballot(head);
    if (lane == 0) {
      s_first[wv] = hb ? base + wv * 64 + __builtin_ctzll(hb) : ~0ull;
      s_last[wv] = hb ? base + wv * 64 + 63 - __builtin_clzll(hb) : ~0ull;
    }
    u64 t;
    block_excl_scan(head ? 1 : 0, tmp, &t);  // its barriers publish s_first / s_last
    u64 tf = ~0ull, tl = ~0ull;
    for (int q = 0; q < kWaves; q++) {
      if (s_first[q] != ~0ull && tf == ~0ull) tf = s_first[q];
      if (s_last[q] != ~0ull) tl = s_last[q];
    }
    // min of the elements before the tile's first head, and from its last head on
    WMin pv = (valid && i < tf) ? v : WMin{0, 1};
    WMin tv = (valid && tl != ~0ull && i >= tl) ? v : WMin{0, 1};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      WMin a{__shfl_xor(pv.w, o, 64), (u8)__shfl_xor((int)pv.f, o, 64)};
      WMin b{__shfl_xor(tv.w, o, 64), (u8)__shfl_xor((int)tv.f, o, 64)};
      pv = wmin(pv, a);
      tv = wmin(tv, b);
    }
    if (lane == 0) {
      s_pw[wv] = pv.w;
      s_pf[wv] = pv.f;
      s_tw[wv] = tv.w;
      s_tf[wv] = tv.f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      WMin pt{0, 1}, tt{0, 1};
      for (int q = 0; q < kWaves; q++) {
        pt = wmin(pt, WMin{s_pw[q], s_pf[q]});
        tt = wmin(tt, WMin{s_tw[q], s_tf[q]});
      }
      // elements before the tile's first head extend the run open so far
      if (nh == 0) pre = wmin(pre, pt);
      else tail = wmin(tail, pt);
      if (tf != ~0ull) {
        if (nh == 0) firsth = tf;
        tail = tt;
      }
    }
    nh += t;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    heads[blockIdx.x] = nh;
    pre_w[blockIdx.x] = pre.w;
    pre_f[blockIdx.x] = pre.f;
    tail_w[blockIdx.x] = nh ? tail.w : pre.w;
    tail_f[blockIdx.x] = nh ? tail.f : pre.f;
    first_head[blockIdx.x] = firsth;
  }
}

// Sequential over chunks: carry-in minimum and head offsets.  The grand
// total of heads lands in heads[nb].
__global__ void k_runs_carry(u64* heads, const i64* pre_w, const u8* pre_f, const i64* tail_w, const u8* tail_f,
                             i64* cin_w, u8* cin_f, u32 nb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  WMin carry{0, 1};
  u64 acc = 0;
  for (u32 b = 0; b < nb; b++) {
    cin_w[b] = carry.w;
    cin_f[b] = carry.f;
    u64 h = heads[b];
    heads[b] = acc;
    acc += h;
    if (h) carry = WMin{tail_w[b], tail_f[b]};
    else carry = wmin(carry, WMin{pre_w[b], pre_f[b]});
  }
  heads[nb] = acc;
}

template <class K>
__global__ __launch_bounds__(kBlock) void k_runs_write_w(const K* __restrict__ keys, const i64* __restrict__ w,
                                                         const u8* __restrict__ f, u64 n, u64 per,
                                                         const u64* __restrict__ offs, const i64* __restrict__ cin_w,
                                                         const u8* __restrict__ cin_f, K* __restrict__ ok,
                                                         i64* __restrict__ ow, u8* __restrict__ of) {
  __shared__ u64 tmp[kWaves];
  __shared__ i64 s_ww[kWaves];
  __shared__ u8 s_wf[kWaves], s_wh[kWaves];
  const u64 lo = (u64)blockIdx.x * per, hi = min(n, lo + per);
  u64 carry_h = offs[blockIdx.x];  // heads before this tile
  WMin carry{cin_w[blockIdx.x], cin_f[blockIdx.x]};
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (u64 base = lo; base < hi; base += kBlock) {
    u64 i = base + threadIdx.x;
    bool valid = i < hi;
    K k = valid ? keys[i] : (K)0;
    bool head = valid && (i == 0 || keys[i - 1] != k);
    bool last = valid && (i + 1 == n || keys[i + 1] != k);
    i64 vw = valid ? w[i] : 0;
    u8 vf = valid ? (f ? f[i] : (u8)0) : (u8)1;
    u8 fl = head ? 1 : 0;
    // wave segmented inclusive min-scan
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      i64 pw = __shfl_up(vw, o, 64);
      int pf = __shfl_up((int)vf, o, 64);
      int ph = __shfl_up((int)fl, o, 64);
      if (lane >= (u32)o) {
        if (!fl && w_less(pw, (u8)pf, vw, vf)) {
          vw = pw;
          vf = (u8)pf;
        }
        fl |= (u8)ph;
      }
    }
    if (lane == 63) {
      s_ww[wv] = vw;
      s_wf[wv] = vf;
      s_wh[wv] = fl;
    }
    u64 tot;
    u64 ex = block_excl_scan(head ? 1 : 0, tmp, &tot);  // barriers publish s_ww / s_wf / s_wh
    // carry into this wave: the tile carry combined with earlier waves
    WMin c = carry;
    for (u32 q = 0; q < wv; q++) {
      if (s_wh[q]) c = WMin{s_ww[q], s_wf[q]};
      else c = wmin(c, WMin{s_ww[q], s_wf[q]});
    }
    WMin mine{vw, vf};
    if (!fl) mine = wmin(c, mine);
    if (last) {
      // run rank = heads before this element, inclusive of its own head
      u64 o = carry_h + ex + (head ? 1 : 0) - 1;
      ok[o] = k;
      ow[o] = mine.w;
      if (of) of[o] = mine.f;
    }
    // tile carry: the state after the tile's last element
    WMin nc = carry;
    for (u32 q = 0; q < (u32)kWaves; q++) {
      if (s_wh[q]) nc = WMin{s_ww[q], s_wf[q]};
      else nc = wmin(nc, WMin{s_ww[q], s_wf[q]});
    }
    __syncthreads();
    carry = nc;
    carry_h += tot;
  }
}

template <class K>
u64 unique_runs(const Exec& x, const K* keys, const i64* w, const u8* inf, u64 n, K* ok, i64* ow, u8* of) {
  if (n == 0) return 0;
  if (w) {
    u64 nb = (n + kScanTile - 1) / kScanTile;
    if (nb > kScanMaxBlocks) nb = kScanMaxBlocks;
    u64 per = (n + nb - 1) / nb;
    per = (per + kBlock - 1) / kBlock * kBlock;
    nb = (n + per - 1) / per;
    DBuf<u64> heads(x.pool, nb + 1), firsth(x.pool, nb);
    DBuf<i64> pw(x.pool, nb), tw(x.pool, nb), cw(x.pool, nb);
    DBuf<u8> pf(x.pool, nb), tf(x.pool, nb), cf(x.pool, nb);
    const double rec = sizeof(K) + 8.0 + (inf ? 1.0 : 0);
    {
      Scope sc(x.timer, "runs_count", KC_DEDUP, (double)n * rec, n);
      k_runs_sum_w<K><<<(unsigned)nb, kBlock, 0, x.stream>>>(keys, w, inf, n, per, heads.get(), pw.get(), pf.get(),
                                                             tw.get(), tf.get(), firsth.get());
    }
    k_runs_carry<<<1, 64, 0, x.stream>>>(heads.get(), pw.get(), pf.get(), tw.get(), tf.get(), cw.get(), cf.get(),
                                         (u32)nb);
    {
      Scope sc(x.timer, "runs_write", KC_DEDUP, (double)n * rec, n);
      k_runs_write_w<K><<<(unsigned)nb, kBlock, 0, x.stream>>>(keys, w, inf, n, per, heads.get(), cw.get(), cf.get(),
                                                               ok, ow, of);
    }
    SK_HIP(hipGetLastError());
    return read_u64(x, heads.get() + nb);
  }
  u64 nb = (n + kScanTile - 1) / kScanTile;
  if (nb > kScanMaxBlocks) nb = kScanMaxBlocks;
  u64 per = (n + nb - 1) / nb;
  per = (per + kBlock - 1) / kBlock * kBlock;
  nb = (n + per - 1) / per;
  DBuf<u64> cnt(x.pool, nb + 1);
  {
    Scope sc(x.timer, "runs_count", KC_DEDUP, (double)n * sizeof(K), n);
    k_runs_count<K><<<(unsigned)nb, kBlock, 0, x.stream>>>(keys, n, per, cnt.get());
  }
  k_scan_sums<<<1, kBlock, 0, x.stream>>>(cnt.get(), (u32)nb, cnt.get() + nb);
  {
  const double rec = sizeof(K) + (w ? 8.0 : 0) + (inf ? 1.0 : 0);
  Scope sc(x.timer, "runs_write", KC_DEDUP, (double)n * rec, n);
  if (w)
    k_runs_write<K, true><<<(unsigned)nb, kBlock, 0, x.stream>>>(keys, w, inf, n, per, cnt.get(), ok, ow, of);
  else
    k_runs_write<K, false><<<(unsigned)nb, kBlock, 0, x.stream>>>(keys, nullptr, nullptr, n, per, cnt.get(), ok,
                                                                  nullptr, nullptr);
  }
  SK_HIP(hipGetLastError());
  return read_u64(x, cnt.get() + nb);
}

// ============================================================== merge path
//
// The merged order takes full (B) before cand (A) on equal keys, so a match
// is an adjacent (B, A) pair.  Each tile of kMergeTile merged positions
// loads its A range and its B range into one LDS buffer, each range with a
// one-element halo on both sides (the global neighbours of the tile).  Every
// thread finds its split of the tile by merge path, reads the element just
// before its first position from the split itself, and then streams its
// positions in order: an element's union row is decided when the element
// after it is known, its fresh row as soon as it arrives.  Nothing crosses
// threads but the block scan of the row counts.

constexpr int kMergeItems = 8;
constexpr u64 kMergeTile = (u64)kBlock * kMergeItems;  // 2048
constexpr u32 kMergeSlots = (u32)kMergeTile + 4;       // A and B ranges plus four halo slots

template <class K>
__device__ __forceinline__ u64 merge_path(const K* a, u64 m, const K* b, u64 n, u64 d) {
  u64 lo = d > n ? d - n : 0, hi = d < m ? d : m;
  while (lo < hi) {
    u64 mid = (lo + hi) >> 1;
    if (a[mid] < b[d - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <class K>
__global__ void k_merge_partition(const K* __restrict__ a, u64 m, const K* __restrict__ b, u64 n, u64 ntiles,
                                  u64* __restrict__ part) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  u64 d = min(t * kMergeTile, m + n);
  part[t] = merge_path(a, m, b, n, d);
}

// One thread's walk over `cnt` merged positions starting at tile-local split
// (li, lj).  Slot layout of the LDS buffer: A[x] is slot 1 + x and B[x] is
// slot boff + 1 + x, for x in [-1, na] resp. [-1, nb]; a halo slot is valid
// only when its global index lies inside the input.
template <class K, bool HASW, class FU, class FF>
__device__ __forceinline__ void merge_walk(const K* sk, const i64* sw, const u8* sf, u32 boff, u32 na, u32 nb,
                                           u64 i0, u64 j0, u64 m, u64 n, u32 li, u32 lj, u32 cnt, FU&& on_union,
                                           FF&& on_fresh) {
  if (cnt == 0) return;
  // the merged element just before the first position: the later of
  // A[li-1] and B[lj-1] (A is later on equal keys)
  bool haslast = false;
  K lk{};
  i64 lw = 0;
  u8 lf = 0;
  {
    bool va = i0 + li > 0, vb = j0 + lj > 0;
    if (va || vb) {
      bool pick_a = va && (!vb || !(sk[li] < sk[boff + lj]));
      u32 s = pick_a ? li : boff + lj;
      lk = sk[s];
      if constexpr (HASW) { lw = sw[s]; lf = sf[s]; }
      haslast = true;
    }
  }
  bool held = false;
  K hk{};
  i64 hw = 0;
  u8 hf = 0;
#pragma unroll
  for (int k = 0; k < kMergeItems; k++) {
    if ((u32)k < cnt) {
      bool take_b = (lj < nb) && (li >= na || !(sk[1 + li] < sk[boff + 1 + lj]));
      u32 s = take_b ? boff + 1 + lj : 1 + li;
      K ck = sk[s];
      i64 cw = 0;
      u8 cf = 0;
      if constexpr (HASW) { cw = sw[s]; cf = sf[s]; }
      if (take_b) lj++;
      else li++;
      bool dup = haslast && lk == ck;
      if (held && !dup) on_union(hk, hw, hf);
      i64 uw = cw;
      u8 uf = cf;
      bool fresh = !dup;
      if constexpr (HASW) {
        if (dup && w_less(lw, lf, cw, cf)) { uw = lw; uf = lf; }
        if (dup && w_less(cw, cf, lw, lf)) fresh = true;
      }
      if (!take_b && fresh) on_fresh(ck, cw, cf);
      held = true;
      hk = ck; hw = uw; hf = uf;
      lk = ck; lw = cw; lf = cf;
      haslast = true;
    }
  }
  // the merged element just after the last position: the earlier of A[li]
  // and B[lj] (B is earlier on equal keys)
  bool va = i0 + li < m, vb = j0 + lj < n;
  bool dup_next = false;
  if (va || vb) {
    bool pick_b = vb && (!va || !(sk[1 + li] < sk[boff + 1 + lj]));
    dup_next = (pick_b ? sk[boff + 1 + lj] : sk[1 + li]) == hk;
  }
  if (!dup_next) on_union(hk, hw, hf);
}

// PASS 0: count union / fresh rows per tile.  PASS 1: write them.
template <class K, bool HASW, int PASS>
__global__ __launch_bounds__(kBlock) void k_merge(
    const K* __restrict__ a, const i64* __restrict__ aw, const u8* __restrict__ af, u64 m,
    const K* __restrict__ b, const i64* __restrict__ bw, const u8* __restrict__ bf, u64 n,
    const u64* __restrict__ part, u64* __restrict__ ucnt, u64* __restrict__ fcnt,
    K* __restrict__ uk, i64* __restrict__ uw, u8* __restrict__ uf,
    K* __restrict__ fk, i64* __restrict__ fw, u8* __restrict__ ff) {
  __shared__ K sk[kMergeSlots];
  __shared__ i64 sw[HASW ? kMergeSlots : 1];
  __shared__ u8 sf[HASW ? kMergeSlots : 1];
  __shared__ u64 tmp[kWaves];

  const u64 t = blockIdx.x;
  const u64 d0 = t * kMergeTile, d1 = min(d0 + kMergeTile, m + n);
  const u64 i0 = part[t], i1 = part[t + 1];
  const u64 j0 = d0 - i0, j1 = d1 - i1;
  const u32 na = (u32)(i1 - i0), nb = (u32)(j1 - j0);
  const u32 boff = na + 2;
  const bool has_f = HASW && (af != nullptr);
  // slot s of the A range holds a[i0 - 1 + s]; of the B range b[j0 - 1 + s]
  for (u32 s = threadIdx.x; s < na + 2; s += kBlock) {
    u64 g = i0 + s - 1;  // wraps past m for s == 0 when i0 == 0
    if (g < m) {
      sk[s] = a[g];
      if constexpr (HASW) {
        sw[s] = aw[g];
        sf[s] = has_f ? af[g] : (u8)0;
      }
    }
  }
  for (u32 s = threadIdx.x; s < nb + 2; s += kBlock) {
    u64 g = j0 + s - 1;
    if (g < n) {
      sk[boff + s] = b[g];
      if constexpr (HASW) {
        sw[boff + s] = bw[g];
        sf[boff + s] = has_f ? bf[g] : (u8)0;
      }
    }
  }
  __syncthreads();

  const u32 tn = (u32)(d1 - d0);
  const u32 ld = min((u32)threadIdx.x * kMergeItems, tn);
  const u32 cnt = min(ld + (u32)kMergeItems, tn) - ld;
  const u32 li = (u32)merge_path<K>(sk + 1, na, sk + boff + 1, nb, ld);
  const u32 lj = ld - li;

  u32 ucount = 0, fcount = 0;
  merge_walk<K, HASW>(sk, sw, sf, boff, na, nb, i0, j0, m, n, li, lj, cnt,
                      [&](K, i64, u8) { ucount++; }, [&](K, i64, u8) { fcount++; });
  u64 ut, ft;
  u64 uo = block_excl_scan(ucount, tmp, &ut);
  u64 fo = block_excl_scan(fcount, tmp, &ft);
  if constexpr (PASS == 0) {
    if (threadIdx.x == 0) {
      ucnt[t] = ut;
      fcnt[t] = ft;
    }
  } else {
    uo += ucnt[t];
    fo += fcnt[t];
    merge_walk<K, HASW>(
        sk, sw, sf, boff, na, nb, i0, j0, m, n, li, lj, cnt,
        [&](K key, i64 w, u8 f) {
          uk[uo] = key;
          if constexpr (HASW) {
            uw[uo] = w;
            if (uf) uf[uo] = f;
          }
          uo++;
        },
        [&](K key, i64 w, u8 f) {
          fk[fo] = key;
          if constexpr (HASW) {
            fw[fo] = w;
            if (ff) ff[fo] = f;
          }
          fo++;
        });
  }
}

template <class K>
void merge_diff(const Exec& x, const K* a, const i64* aw, const u8* ainf, u64 m, const K* b, const i64* bw,
                const u8* binf, u64 n, DBuf<K>* u_keys, DBuf<i64>* u_w, DBuf<u8>* u_inf, u64* n_union,
                DBuf<K>* f_keys, DBuf<i64>* f_w, DBuf<u8>* f_inf, u64* n_fresh, bool weighted, bool has_inf) {
  *n_union = n;
  *n_fresh = 0;
  if (m == 0) return;
  double rb = sizeof(K) + (weighted ? 8.0 + (has_inf ? 1.0 : 0) : 0);
  u64 tot = m + n;
  u64 nt = (tot + kMergeTile - 1) / kMergeTile;
  DBuf<u64> part(x.pool, nt + 1);
  DBuf<u64> uc(x.pool, nt + 1), fc(x.pool, nt + 1);
  k_merge_partition<K><<<grid_for(nt + 1), kBlock, 0, x.stream>>>(a, m, b, n, nt, part.get());
  const u8* af = has_inf ? ainf : nullptr;
  const u8* bf = has_inf ? binf : nullptr;
  {
  Scope sc(x.timer, "merge_count", KC_MERGE, (double)(m + n) * rb, m + n);
  if (weighted)
    k_merge<K, true, 0><<<(unsigned)nt, kBlock, 0, x.stream>>>(a, aw, af, m, b, bw, bf, n, part.get(), uc.get(),
                                                               fc.get(), nullptr, nullptr, nullptr, nullptr,
                                                               nullptr, nullptr);
  else
    k_merge<K, false, 0><<<(unsigned)nt, kBlock, 0, x.stream>>>(a, nullptr, nullptr, m, b, nullptr, nullptr, n,
                                                                part.get(), uc.get(), fc.get(), nullptr, nullptr,
                                                                nullptr, nullptr, nullptr, nullptr);
  }
  k_scan_sums<<<1, kBlock, 0, x.stream>>>(uc.get(), (u32)nt, uc.get() + nt);
  k_scan_sums<<<1, kBlock, 0, x.stream>>>(fc.get(), (u32)nt, fc.get() + nt);
  SK_HIP(hipGetLastError());
  u64 nu = read_u64(x, uc.get() + nt);
  u64 nf = read_u64(x, fc.get() + nt);
  *n_fresh = nf;
  if (nf == 0) return;
  *n_union = nu;
  u_keys->ensure(x.pool, nu);
  f_keys->ensure(x.pool, nf);
  // algorithmic bytes: read both inputs once, write union and fresh once
  Scope sc(x.timer, "merge_write", KC_MERGE, (double)(m + n + nu + nf) * rb, m + n);
  if (weighted) {
    u_w->ensure(x.pool, nu);
    f_w->ensure(x.pool, nf);
    if (has_inf) {
      u_inf->ensure(x.pool, nu);
      f_inf->ensure(x.pool, nf);
    }
    k_merge<K, true, 1><<<(unsigned)nt, kBlock, 0, x.stream>>>(
        a, aw, af, m, b, bw, bf, n, part.get(), uc.get(), fc.get(), u_keys->get(), u_w->get(),
        has_inf ? u_inf->get() : nullptr, f_keys->get(), f_w->get(), has_inf ? f_inf->get() : nullptr);
  } else {
    k_merge<K, false, 1><<<(unsigned)nt, kBlock, 0, x.stream>>>(a, nullptr, nullptr, m, b, nullptr, nullptr, n,
                                                                part.get(), uc.get(), fc.get(), u_keys->get(),
                                                                nullptr, nullptr, f_keys->get(), nullptr, nullptr);
  }
  SK_HIP(hipGetLastError());
}

// ===================================================================== CSR

template <class K>
__global__ void k_csr(const K* __restrict__ keys, u64 n, int shift, u64 nvals, u64* __restrict__ csr) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  // rows [i-1, i) boundary: values (prev, cur] start at row i
  long long cur = i < n ? (long long)(u64)(keys[i] >> shift) : (long long)nvals;
  long long prev = i > 0 ? (long long)(u64)(keys[i - 1] >> shift) : -1;
  if (cur > (long long)nvals) cur = (long long)nvals;
  for (long long v = prev + 1; v <= cur; v++) csr[v] = i;
}

// Few values over many rows: csr[v] = the first row whose leading value is
// >= v, by one binary search per value.
template <class K>
__global__ void k_csr_search(const K* __restrict__ keys, u64 n, int shift, u64 nvals, u64* __restrict__ csr) {
  const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nvals) return;
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if ((u64)(keys[mid] >> shift) < v) lo = mid + 1;
    else hi = mid;
  }
  csr[v] = lo;
}

template <class K>
void build_csr(const Exec& x, const K* keys, u64 n, int shift, u64 nvals, u64* csr) {
  if ((nvals + 1) * 32 < n) {
    Scope sc(x.timer, "csr", KC_MISC, (nvals + 1) * 8.0, nvals + 1);
    k_csr_search<K><<<grid_for(nvals + 1), kBlock, 0, x.stream>>>(keys, n, shift, nvals, csr);
    SK_HIP(hipGetLastError());
    return;
  }
  Scope sc(x.timer, "csr", KC_MISC, (double)n * sizeof(K) + (nvals + 1) * 8.0, n);
  k_csr<K><<<grid_for(n + 1), kBlock, 0, x.stream>>>(keys, n, shift, nvals, csr);
  SK_HIP(hipGetLastError());
}

// ========================================================== pack / unpack

template <class K>
__global__ void k_unpack(const K* __restrict__ keys, u64 n, int arity, int bits, Cols cols) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = keys[i];
  K mask = (((K)1) << bits) - 1;
  for (int j = 0; j < arity; j++) {
    int sh = bits * (arity - 1 - j);
    cols.c[j][i] = (u32)((k >> sh) & mask);
  }
}

template <class K>
void unpack_cols(const Exec& x, const K* keys, u64 n, int arity, int bits, u32* const* cols) {
  if (!n || !arity) return;
  Scope sc(x.timer, "unpack", KC_MISC, (double)n * (sizeof(K) + 4.0 * arity), n);
  Cols c{};
  for (int j = 0; j < arity && j < kMaxCols; j++) c.c[j] = cols[j];
  k_unpack<K><<<grid_for(n), kBlock, 0, x.stream>>>(keys, n, arity, bits, c);
  SK_HIP(hipGetLastError());
}

struct Perm8 {
  int p[8];
};

template <class K, class KO>
__global__ void k_repack(const K* __restrict__ keys, u64 n, int arity, int bits, Perm8 perm, KO* __restrict__ out) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = keys[i];
  K mask = (((K)1) << bits) - 1;
  KO o = 0;
  for (int j = 0; j < arity; j++) {
    int src = perm.p[j];
    K c = (k >> (bits * (arity - 1 - src))) & mask;
    o = (o << bits) | (KO)c;
  }
  out[i] = o;
}

template <class K, class KO>
void repack(const Exec& x, const K* keys, u64 n, int arity, int bits, const int* perm, KO* out) {
  if (!n) return;
  Scope sc(x.timer, "repack", KC_MISC, (double)n * (sizeof(K) + sizeof(KO)), n);
  Perm8 p8;
  for (int j = 0; j < 8; j++) p8.p[j] = j < arity ? perm[j] : 0;
  k_repack<K, KO><<<grid_for(n), kBlock, 0, x.stream>>>(keys, n, arity, bits, p8, out);
  SK_HIP(hipGetLastError());
}

// ============================================================ membership

template <class K>
__global__ void k_member(const K* __restrict__ s, u64 ns, const K* __restrict__ q, u64 nq, u8* __restrict__ flags) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  K k = q[i];
  u64 lo = 0, hi = ns;
  while (lo < hi) {
    u64 mid = (lo + hi) >> 1;
    if (s[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  flags[i] = (lo < ns && s[lo] == k) ? 1 : 0;
}

template <class K>
void member(const Exec& x, const K* s, u64 ns, const K* q, u64 nq, u8* flags) {
  if (!nq) return;
  Scope sc(x.timer, "member", KC_MISC, (double)nq * (sizeof(K) + 1.0), nq);
  k_member<K><<<grid_for(nq), kBlock, 0, x.stream>>>(s, ns, q, nq, flags);
  SK_HIP(hipGetLastError());
}

// ======================================================= instantiations

#define SK_INST_K(K)                                                                                     \
  template void radix_sort<K>(const Exec&, K*, K*, i64*, i64*, u8*, u8*, u64, int, K**, i64**, u8**, int);  \
  template u64 unique_runs<K>(const Exec&, const K*, const i64*, const u8*, u64, K*, i64*, u8*);        \
  template void merge_diff<K>(const Exec&, const K*, const i64*, const u8*, u64, const K*, const i64*, \
                              const u8*, u64, DBuf<K>*, DBuf<i64>*, DBuf<u8>*, u64*, DBuf<K>*,          \
                              DBuf<i64>*, DBuf<u8>*, u64*, bool, bool);                                 \
  template void build_csr<K>(const Exec&, const K*, u64, int, u64, u64*);                               \
  template void unpack_cols<K>(const Exec&, const K*, u64, int, int, u32* const*);                      \
  template void member<K>(const Exec&, const K*, u64, const K*, u64, u8*);

SK_INST_K(u64)
SK_INST_K(u128)
template void repack<u64, u64>(const Exec&, const u64*, u64, int, int, const int*, u64*);
template void repack<u64, u128>(const Exec&, const u64*, u64, int, int, const int*, u128*);
template void repack<u128, u64>(const Exec&, const u128*, u64, int, int, const int*, u64*);
template void repack<u128, u128>(const Exec&, const u128*, u64, int, int, const int*, u128*);

// ======================================================= pool and timer

u64 Pool::size_class(u64 b) {
  if (b <= 256) return 256;
  // classes at 1/4 steps between powers of two
  int lg = 63 - __builtin_clzll(b - 1);
  u64 base = 1ull << lg;
  u64 step = base / 4;
  return (b + step - 1) / step * step;
}

void* Pool::alloc(u64 bytes) {
  u64 c = size_class(bytes);
  auto it = free_.lower_bound(c);
  if (it != free_.end() && it->first <= c + c / 4) {
    void* p = it->second;
    u64 sz = it->first;
    free_.erase(it);
    live_[p] = sz;
    in_use_ += sz;
    return p;
  }
  if (held_ + c > budget_) trim();
  if (held_ + c > budget_) throw BudgetError{bytes};
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw BudgetError{bytes};
    }
  }
  live_[p] = c;
  held_ += c;
  in_use_ += c;
  if (held_ > peak_) peak_ = held_;
  return p;
}

void Pool::free(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  in_use_ -= it->second;
  free_.emplace(it->second, p);
  live_.erase(it);
}

void Pool::trim() {
  if (free_.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    held_ -= kv.first;
  }
  free_.clear();
}

Pool::~Pool() {
  (void)hipDeviceSynchronize();
  for (auto& kv : free_) (void)hipFree(kv.second);
  for (auto& kv : live_) (void)hipFree(kv.first);
}

hipEvent_t Timer::get_event() {
  if (!spare_.empty()) {
    hipEvent_t e = spare_.back();
    spare_.pop_back();
    return e;
  }
  hipEvent_t e;
  SK_HIP(hipEventCreate(&e));
  return e;
}

void Timer::begin(const char* name, int cls, double bytes, u64 items) {
  if (open_) end();
  cur_ = Pending{get_event(), nullptr, name, cls, bytes, items};
  SK_HIP(hipEventRecord(cur_.a, stream));
  open_ = true;
}

void Timer::end() {
  if (!open_) return;
  cur_.b = get_event();
  SK_HIP(hipEventRecord(cur_.b, stream));
  pending_.push_back(cur_);
  open_ = false;
  if (pending_.size() > 4096) collect();
}

void Timer::credit(const char* name, int cls, double bytes, u64 items) {
  if (!enabled) return;
  stats[cls].bytes += bytes;
  stats[cls].items += items;
  KRecord& r = kernels[name];
  r.name = name;
  r.cls = cls;
  r.bytes += bytes;
  r.items += items;
}

void Timer::collect() {
  if (pending_.empty()) return;
  SK_HIP(hipStreamSynchronize(stream));
  for (auto& p : pending_) {
    float ms = 0;
    SK_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    stats[p.cls].launches++;
    stats[p.cls].ms += ms;
    stats[p.cls].bytes += p.bytes;
    stats[p.cls].items += p.items;
    KRecord& r = kernels[p.name];
    r.name = p.name;
    r.cls = p.cls;
    r.launches++;
    r.ms += ms;
    r.bytes += p.bytes;
    r.items += p.items;
    spare_.push_back(p.a);
    spare_.push_back(p.b);
  }
  pending_.clear();
}

void Timer::reset() {
  collect();
  for (auto& st : stats) st = KStat();
  kernels.clear();
}

}  // namespace sk
