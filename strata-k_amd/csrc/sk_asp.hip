// sk_asp.hip — the grounding simplification filter over a ground program in
// CSR form (crates/strata-gpu/src/asp.rs:53-99, 148-355).
//
// A rule `h :- pos, not neg` dies when a positive body atom is impossible or
// a negative one is certain; in a kept rule, positive atoms that are certain
// and negative atoms that are impossible are dropped (trivially true).  Body
// literals are signed: +(atom+1) positive, -(atom+1) negative.
//
// Four launches and one host read:
//   0. the atom state table (below);
//   1. count: one thread per rule, 1024 rules per block; the block writes
//      its (kept rules, kept literals) totals.  Every offset and literal is
//      bounds-checked; a malformed one raises a flag.
//   2. a one-block scan of the block totals into block offsets, which also
//      writes the output's closing offset;
//   3. emit: every thread decides its rule again, the block scans its
//      per-thread totals, gathers its kept heads, offsets and literals in
//      LDS in input order and stores them at its block offset.  It does nothing when the flag
//      is raised (overlapping malformed ranges could otherwise overrun the
//      outputs).
// The host then reads the two totals and the flag with one copy.
//
// Per literal the work is one literal load and one gather of its atom's
// state (0 impossible, 1 possible, 2 certain) from a byte table built by a
// first small launch: one gather per literal instead of two (possible,
// certain).  A thread walks its body in order and stops where the rule dies,
// as in the reference.  (Tried and slower: staging the state table in LDS
// per block, and loading literals four at a time at clamped addresses.)
#include <map>
#include <mutex>

#include "sk_dev.h"

namespace sk {

namespace {

constexpr int kAspT = 1024;             // threads and rules per block (count, emit)
constexpr int kAspTW = kAspT / 64;      // its waves
constexpr int kAspAliveShift = 40;      // in-block scan: kept rules above, kept literals below
constexpr int kAspScanPer = 8;          // block totals per thread of the one-block scan, per round
constexpr int kAspEmitLds = 4096;       // kept literals of a block gathered in LDS before the stores
constexpr int kAspRegs = 8;             // leading literals of a body the emit pass keeps in registers

__device__ __forceinline__ bool asp_valid(int32_t l, u64 n_atoms, u64* a) {
  *a = (u64)(l > 0 ? (int64_t)l : -(int64_t)l) - 1;
  return l != 0 && l != INT32_MIN && *a < n_atoms;
}

// -1 the rule dies, 0 the literal is dropped, 1 it is kept.
__device__ __forceinline__ int asp_decide(int32_t l, u32 s) {
  if (l > 0) return s == 0 ? -1 : (s == 2 ? 0 : 1);
  return s == 2 ? -1 : (s == 0 ? 0 : 1);
}

// Rule body [s, e): its kept literal count, -1 when the rule dies, -2 when
// an offset or a literal is malformed.
__device__ __forceinline__ long asp_eval(u64 s, u64 e, const int32_t* __restrict__ lits, u64 n_lits,
                                         const u8* __restrict__ st, u64 n_atoms) {
  if (e < s || e > n_lits) return -2;
  long k = 0;
  for (u64 j = s; j < e; j++) {
    const int32_t l = lits[j];
    u64 a;
    if (!asp_valid(l, n_atoms, &a)) return -2;
    const int d = asp_decide(l, st[a]);
    if (d < 0) return -1;
    k += d;
  }
  return k;
}

// asp_eval that also keeps the first kAspRegs literal values and which of
// them are kept (bit q of *mask), so that the emit pass writes them without
// walking the body again.
__device__ __forceinline__ long asp_eval_regs(u64 s, u64 e, const int32_t* __restrict__ lits, u64 n_lits,
                                              const u8* __restrict__ st, u64 n_atoms, int32_t (&v)[kAspRegs],
                                              u32* mask) {
  *mask = 0;
  if (e < s || e > n_lits) return -2;
  long k = 0;
#pragma unroll
  for (int q = 0; q < kAspRegs; q++) {
    if (s + q >= e) return k;
    const int32_t l = lits[s + q];
    u64 a;
    if (!asp_valid(l, n_atoms, &a)) return -2;
    const int d = asp_decide(l, st[a]);
    if (d < 0) return -1;
    v[q] = l;
    if (d) {
      k++;
      *mask |= 1u << q;
    }
  }
  for (u64 j = s + kAspRegs; j < e; j++) {
    const int32_t l = lits[j];
    u64 a;
    if (!asp_valid(l, n_atoms, &a)) return -2;
    const int d = asp_decide(l, st[a]);
    if (d < 0) return -1;
    k += d;
  }
  return k;
}

// Block-wide inclusive sum of one u64 per thread (kAspT threads); *total the
// block's sum.
__device__ __forceinline__ u64 asp_block_incl(u64 v, u64* total) {
  __shared__ u64 wsum[kAspTW];
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(v, o, 64);
    if (lane >= (u32)o) v += y;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  u64 before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kAspTW; i++) {
    before += (u32)i < w ? wsum[i] : 0;
    all += wsum[i];
  }
  *total = all;
  return v + before;
}

}  // namespace

// Atom states: 0 impossible, 1 possible, 2 certain.  Certain atoms must also
// be possible (the two reference paths, asp.rs:53-66 and its device filter,
// disagree otherwise); an atom certain but not possible raises the malformed
// flag.
__global__ void k_asp_state(const u8* __restrict__ poss, const u8* __restrict__ cert, u64 n, u8* __restrict__ st,
                            u64* __restrict__ flag) {
  const u64 a = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) {
    const bool c = cert[a] != 0, p = poss[a] != 0;
    st[a] = c ? 2 : (p ? 1 : 0);
    if (c && !p) atomicOr((unsigned long long*)flag, 1ull);
  }
}

__global__ __launch_bounds__(kAspT) void k_asp_count(const u32* __restrict__ bs, const int32_t* __restrict__ lits,
                                                     u64 n_rules, u64 n_lits, const u8* __restrict__ st,
                                                     u64 n_atoms, u64* __restrict__ blk_alive,
                                                     u64* __restrict__ blk_kept, u64* __restrict__ flag) {
  const u64 i = (u64)blockIdx.x * kAspT + threadIdx.x;
  u64 packed = 0;
  if (i < n_rules) {
    const long k = asp_eval(bs[i], bs[i + 1], lits, n_lits, st, n_atoms);
    if (k == -2) atomicOr((unsigned long long*)flag, 1ull);
    if (k >= 0) packed = (1ull << kAspAliveShift) + (u64)k;
  }
  u64 total;
  asp_block_incl(packed, &total);
  if (threadIdx.x == 0) {
    blk_alive[blockIdx.x] = total >> kAspAliveShift;
    blk_kept[blockIdx.x] = total & ((1ull << kAspAliveShift) - 1);
  }
}

// One block of 1024: exclusive offsets of the block totals (in place), the
// totals into tot[0..1], and the output's closing offset out_bs[kept rules].
// A round loads 8 x 1024 totals with coalesced loads, all in flight
// together, then scans them 1024 at a time.
__global__ __launch_bounds__(1024) void k_asp_scan(u64* __restrict__ blk_alive, u64* __restrict__ blk_kept, u64 nblk,
                                                   u64* __restrict__ tot, u32* __restrict__ out_bs) {
  __shared__ u64 wa[16], wk[16];
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  u64 ca = 0, ck = 0;  // running totals, the same in every thread
  for (u64 g = 0; g < nblk; g += 1024ull * kAspScanPer) {
    u64 va[kAspScanPer], vk[kAspScanPer];
#pragma unroll
    for (int q = 0; q < kAspScanPer; q++) {
      const u64 j = g + (u64)q * 1024 + threadIdx.x;
      const u64 jc = j < nblk ? j : nblk - 1;
      va[q] = blk_alive[jc];
      vk[q] = blk_kept[jc];
    }
#pragma unroll
    for (int q = 0; q < kAspScanPer; q++) {
      const u64 j0 = g + (u64)q * 1024;
      if (j0 >= nblk) continue;  // uniform
      const u64 j = j0 + threadIdx.x;
      u64 a = j < nblk ? va[q] : 0, k = j < nblk ? vk[q] : 0;
      const u64 a0 = a, k0 = k;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const u64 ya = __shfl_up(a, o, 64), yk = __shfl_up(k, o, 64);
        if (lane >= (u32)o) {
          a += ya;
          k += yk;
        }
      }
      if (lane == 63) {
        wa[w] = a;
        wk[w] = k;
      }
      __syncthreads();
      u64 ba = 0, bk = 0, sa = 0, sk = 0;
#pragma unroll
      for (u32 x = 0; x < 16; x++) {
        if (x < w) {
          ba += wa[x];
          bk += wk[x];
        }
        sa += wa[x];
        sk += wk[x];
      }
      if (j < nblk) {
        blk_alive[j] = ca + ba + a - a0;
        blk_kept[j] = ck + bk + k - k0;
      }
      ca += sa;
      ck += sk;
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    tot[0] = ca;
    tot[1] = ck;
    out_bs[ca] = (u32)ck;
  }
}

__global__ __launch_bounds__(kAspT) void k_asp_emit(const int32_t* __restrict__ head, const u32* __restrict__ bs,
                                                    const int32_t* __restrict__ lits, u64 n_rules, u64 n_lits,
                                                    const u8* __restrict__ st, u64 n_atoms,
                                                    const u64* __restrict__ blk_alive,
                                                    const u64* __restrict__ blk_kept, const u64* __restrict__ flag,
                                                    int32_t* __restrict__ out_head, u32* __restrict__ out_bs,
                                                    int32_t* __restrict__ out_lits) {
  if (*flag) return;
  const u64 i = (u64)blockIdx.x * kAspT + threadIdx.x;
  long k = -1;
  u64 s = 0, e = 0;
  int32_t v[kAspRegs];
  u32 vmask = 0;
  if (i < n_rules) {
    s = bs[i];
    e = bs[i + 1];
    k = asp_eval_regs(s, e, lits, n_lits, st, n_atoms, v, &vmask);
  }
  const u64 packed = k >= 0 ? (1ull << kAspAliveShift) + (u64)k : 0;
  u64 total;
  const u64 excl = asp_block_incl(packed, &total) - packed;
  // The block's kept rules and literals are contiguous in the output: they
  // are gathered in LDS and written with coalesced stores (the literals
  // directly when more than fit).
  __shared__ int32_t s_head[kAspT];
  __shared__ u32 s_bs[kAspT];
  __shared__ int32_t s_lit[kAspEmitLds];
  const u64 mask = (1ull << kAspAliveShift) - 1;
  const u64 n_alive = total >> kAspAliveShift, n_kept = total & mask;
  const u64 rbase = blk_alive[blockIdx.x], lbase = blk_kept[blockIdx.x];
  const bool staged = n_kept <= (u64)kAspEmitLds;
  if (k >= 0) {
    const u64 r = excl >> kAspAliveShift;
    u64 lp = excl & mask;
    s_head[r] = head[i];
    s_bs[r] = (u32)(lbase + lp);
#pragma unroll
    for (int q = 0; q < kAspRegs; q++) {
      if (vmask >> q & 1u) {
        if (staged)
          s_lit[lp] = v[q];
        else
          out_lits[lbase + lp] = v[q];
        lp++;
      }
    }
    for (u64 j = s + kAspRegs; j < e; j++) {  // a long body's tail; every literal is known valid
      const int32_t l = lits[j];
      const u64 a = (u64)(l > 0 ? (int64_t)l : -(int64_t)l) - 1;
      if (asp_decide(l, st[a]) == 1) {
        if (staged)
          s_lit[lp] = l;
        else
          out_lits[lbase + lp] = l;
        lp++;
      }
    }
  }
  __syncthreads();
  for (u64 t = threadIdx.x; t < n_alive; t += kAspT) {
    out_head[rbase + t] = s_head[t];
    out_bs[rbase + t] = s_bs[t];
  }
  if (staged)
    for (u64 t = threadIdx.x; t < n_kept; t += kAspT) out_lits[lbase + t] = s_lit[t];
}

// Clears the two counts and the flag; with no rules, also writes the
// output's single offset (0).
__global__ void k_asp_clear(u64* __restrict__ tot, u32* __restrict__ out_bs0) {
  if (threadIdx.x < 3) tot[threadIdx.x] = 0;
  if (threadIdx.x == 0 && out_bs0) *out_bs0 = 0;
}

// ------------------------------------------------ allocation-free variant
//
// The filter's scratch for callers that cannot allocate device memory: a
// file-scope device array holding the two counts and the malformed flag,
// the per-block totals of up to kAspMaxRules rules, and the atom state table
// of up to kAspMaxAtoms atoms.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): each device's
// address is resolved once with hipGetSymbolAddress and kept per device.
// Every call on one device shares that device's copy; calls that overlap on
// one device (two streams, two host threads) race on it.

struct AspScratch {
  u64* ba;
  u64* bk;
  u64* tot;
  u8* st;
};

constexpr u64 kAspMaxBlocks = kAspMaxRules / kAspT;
constexpr u64 kAspScratchBytes = 256 + 2 * kAspMaxBlocks * sizeof(u64) + kAspMaxAtoms;
static_assert(kAspMaxRules % kAspT == 0, "whole blocks");

__device__ __align__(256) unsigned char scratch[kAspScratchBytes];

static AspScratch asp_scratch_parts(u8* base) {
  u64* ba = reinterpret_cast<u64*>(base + 256);
  return AspScratch{ba, ba + kAspMaxBlocks, reinterpret_cast<u64*>(base), base + 256 + 2 * kAspMaxBlocks * 8};
}

namespace {
std::mutex g_scratch_mu;
std::map<int, u8*> g_scratch_ptr;  // device ordinal -> that device's copy of `scratch`
}  // namespace

u8* asp_scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) return it->second;
  int cur = -1;
  SK_HIP(hipGetDevice(&cur));
  if (cur != device) throw DeviceError{hipErrorInvalidDevice, "asp_scratch_address: device not current"};
  void* p = nullptr;
  SK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<u8*>(p);
  return static_cast<u8*>(p);
}

// Device-resident inputs and outputs (caller-sized: n_rules heads, n_rules + 1
// offsets, n_lits literals).  Returns false when a literal or an offset is
// malformed (outputs then undefined).
bool asp_simplify(const Exec& x, const int32_t* head, const u32* bs, const int32_t* lits, u64 n_rules, u64 n_lits,
                  const u8* poss, const u8* cert, u64 n_atoms, int32_t* out_head, u32* out_bs, int32_t* out_lits,
                  u64* out_rules, u64* out_lits_n, u8* scratch_ptr) {
  const u64 nblk = (n_rules + kAspT - 1) / kAspT;
  if (nblk > 0x7fffffffull) throw BudgetError{n_rules};
  // block totals, the two counts and the malformed flag, the atom states:
  // from the pool, or carved from the caller-resolved static scratch
  DBuf<u64> ba_b, bk_b, tot_b;
  DBuf<u8> st_b;
  AspScratch sc_ = scratch_ptr ? asp_scratch_parts(scratch_ptr) : AspScratch{};
  if (!scratch_ptr) {
    ba_b.reset(x.pool, nblk ? nblk : 1);
    bk_b.reset(x.pool, nblk ? nblk : 1);
    tot_b.reset(x.pool, 3);
    st_b.reset(x.pool, n_atoms ? n_atoms : 1);
    sc_ = AspScratch{ba_b.get(), bk_b.get(), tot_b.get(), st_b.get()};
  }
  u64* const ba = sc_.ba;
  u64* const bk = sc_.bk;
  u64* const tot = sc_.tot;
  u8* const st = sc_.st;
  // a fill kernel, not a memset, so the sequence can be captured in a graph
  k_asp_clear<<<1, 64, 0, x.stream>>>(tot, n_rules ? nullptr : out_bs);
  SK_HIP(hipGetLastError());
  if (n_atoms) {
    Scope sc(x.timer, "asp_state", KC_MISC, (double)n_atoms * 3.0, n_atoms);
    k_asp_state<<<(unsigned)((n_atoms + kBlock - 1) / kBlock), kBlock, 0, x.stream>>>(poss, cert, n_atoms, st,
                                                                                    tot + 2);
    SK_HIP(hipGetLastError());
  }
  if (n_rules) {
    {
      // per rule its end offset, per literal the literal (the atom states
      // stay in cache), per block two totals written
      Scope sc(x.timer, "asp_count", KC_JOIN, (double)n_rules * 4.0 + (double)n_lits * 4.0 + (double)nblk * 16.0,
               n_rules);
      k_asp_count<<<(unsigned)nblk, kAspT, 0, x.stream>>>(bs, lits, n_rules, n_lits, st, n_atoms, ba,
                                                          bk, tot + 2);
      SK_HIP(hipGetLastError());
    }
    {
      Scope sc(x.timer, "asp_scan", KC_MISC, (double)nblk * 32.0, nblk);
      k_asp_scan<<<1, 1024, 0, x.stream>>>(ba, bk, nblk, tot, out_bs);
      SK_HIP(hipGetLastError());
    }
    {
      // bytes credited once the totals are known
      Scope sc(x.timer, "asp_emit", KC_JOIN, 0.0, n_rules);
      k_asp_emit<<<(unsigned)nblk, kAspT, 0, x.stream>>>(head, bs, lits, n_rules, n_lits, st, n_atoms,
                                                         ba, bk, tot + 2, out_head, out_bs,
                                                         out_lits);
      SK_HIP(hipGetLastError());
    }
  }
  u64 hv[3] = {0, 0, 0};
  read_u64s(x, tot, hv, 3);
  if (hv[2]) return false;
  // emit: per rule its end offset, per literal the literal, per kept rule
  // its head read and head + offset written, per kept literal one write
  if (n_rules && x.timer)
    x.timer->credit("asp_emit", KC_JOIN,
                    (double)n_rules * 4.0 + (double)n_lits * 4.0 + (double)hv[0] * 12.0 + (double)hv[1] * 4.0, 0);
  *out_rules = hv[0];
  *out_lits_n = hv[1];
  return true;
}

}  // namespace sk
