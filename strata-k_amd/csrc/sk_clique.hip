// sk_clique.hip — 4-clique (K4) counting over an undirected edge list, the
// counterpart of count_4cliques (crates/strata-gpu/src/lib.rs:109-114,
// wcoj.rs:162-184: same input, same count).
//
// Design:
//   1. The edge list is made simple on the device: each pair is canonicalised
//      to (min, max), self-loops become a sentinel key that sorts last, and a
//      radix sort plus run collapse removes duplicates.
//   2. Edges are oriented from the endpoint of lower (degree, id) to the
//      higher one and re-sorted into a CSR.  In this acyclic orientation an
//      out-list holds at most O(sqrt(|E|)) nodes.
//   3. Every K4 {a, b, c, d}, listed in orientation order, has exactly one
//      bottom edge (a, b): c and d both lie in T(a, b) = N+(a) ∩ N+(b), and
//      (c, d) is an edge inside T.  So
//          K4 = Σ over oriented edges (u, v) of |{(w, y) : w ∈ T, y ∈ N+(w) ∩ T}|.
//   4. Light edges (d+u + d+v ≤ 64, so |T| ≤ 32) take one lane each: the lane
//      merges the two out-lists into a 32-entry slice of LDS, then intersects
//      each member's out-list with the slice (a merge, or binary searches into
//      a much longer list).
//   5. Heavy edges take a wave each.  Lanes test the shorter out-list's
//      values against the longer one by binary search and append the hits,
//      in order, to an LDS table of T.  Then, 64 members of T at a time, the
//      (member, out-list position) pairs are dealt over the lanes through a
//      wave prefix sum of the members' out-degrees, and each lane tests its
//      pair's node for membership in T by binary search in LDS.  When T does
//      not fit the table, membership is tested as "in both out-lists" by two
//      binary searches in global memory, and T's members are re-derived chunk
//      by chunk instead of read from LDS.
#include <algorithm>
#include <climits>

#include "sk_dev.h"

namespace sk {

namespace {

constexpr int kCqSlice = 32;    // light path: LDS entries per lane
constexpr int kCqLight = 64;    // light iff d+u + d+v <= kCqLight (so |T| <= 32)
constexpr int kCqCap = 1024;    // heavy path: LDS entries of T per wave

__device__ __forceinline__ u32 cq_lane() { return threadIdx.x & 63; }
__device__ __forceinline__ u32 cq_wave() { return threadIdx.x >> 6; }

__device__ __forceinline__ u64 cq_incl_scan(u64 v) {
  const u32 l = cq_lane();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(v, o, 64);
    if (l >= (u32)o) v += y;
  }
  return v;
}

__device__ __forceinline__ u64 cq_sum(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// First position in a[lo, hi) holding a value >= x.
__device__ __forceinline__ u64 cq_lower(const u32* __restrict__ a, u64 lo, u64 hi, u32 x) {
  while (lo < hi) {
    const u64 m = (lo + hi) >> 1;
    if (a[m] < x) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__device__ __forceinline__ bool cq_has(const u32* __restrict__ a, u64 lo, u64 hi, u32 x) {
  const u64 p = cq_lower(a, lo, hi, x);
  return p < hi && a[p] == x;
}

static unsigned cq_grid(u64 n, u64 per = kBlock) {
  u64 g = (n + per - 1) / per;
  if (g == 0) g = 1;
  if (g > 0x7fffffffull) g = 0x7fffffffull;
  return (unsigned)g;
}

}  // namespace

// Largest node id of the input (atomicMax of per-wave maxima).
__global__ void k_cq_maxid(const u32* __restrict__ a, const u32* __restrict__ b, u64 n, u32* __restrict__ mx) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u32 m = 0;
  if (i < n) m = max(a[i], b[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (u32)__shfl_xor((int)m, o, 64));
  if (cq_lane() == 0 && m) atomicMax(mx, m);
}

// (min << bits) | max per pair; a self-loop becomes the all-ones key of
// 2*bits bits, which no pair with min < max can equal.
__global__ void k_cq_canon(const u32* __restrict__ a, const u32* __restrict__ b, u64 n, int bits,
                           u64* __restrict__ out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 x = a[i], y = b[i];
  const u64 sentinel = bits == 32 ? ~0ull : ((1ull << (2 * bits)) - 1);
  out[i] = x == y ? sentinel : (((u64)min(x, y) << bits) | max(x, y));
}

__global__ void k_cq_degree(const u64* __restrict__ keys, u64 n, int bits, u32* __restrict__ deg) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 k = keys[i];
  atomicAdd(&deg[k >> bits], 1u);
  atomicAdd(&deg[k & ((1ull << bits) - 1)], 1u);
}

// Orient each edge from the endpoint of lower (degree, id) to the higher.
__global__ void k_cq_orient(const u64* __restrict__ keys, u64 n, int bits, const u32* __restrict__ deg,
                            u64* __restrict__ out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 k = keys[i];
  const u32 x = (u32)(k >> bits), y = (u32)(k & ((1ull << bits) - 1));  // x < y
  const u32 dx = deg[x], dy = deg[y];
  const bool fwd = dx <= dy;  // equal degrees go by id, and x < y
  out[i] = fwd ? k : (((u64)y << bits) | x);
}

// heavy[e] = 1 when the edge's two out-lists hold more than kCqLight values
// (or force_heavy); light[e] = the complement.
__global__ void k_cq_classify(const u64* __restrict__ okeys, u64 m, int bits, const u64* __restrict__ csr,
                              int force_heavy, u8* __restrict__ heavy, u8* __restrict__ light) {
  const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const u64 k = okeys[e];
  const u32 u = (u32)(k >> bits), v = (u32)(k & ((1ull << bits) - 1));
  const u64 work = (csr[u + 1] - csr[u]) + (csr[v + 1] - csr[v]);
  const bool h = force_heavy || work > (u64)kCqLight;
  heavy[e] = h ? 1 : 0;
  light[e] = h ? 0 : 1;
}

__global__ void k_cq_gather(const u8* __restrict__ flag, const u64* __restrict__ pos, u64 m, u32* __restrict__ idx) {
  const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m && flag[e]) idx[pos[e]] = (u32)e;
}

// One lane per light edge.  T lives in a 32-entry LDS slice per lane,
// interleaved (entry j of thread t at j * kBlock + t) so that lanes reading
// the same entry index hit distinct banks.
__global__ __launch_bounds__(kBlock) void k_cq_light(const u64* __restrict__ okeys, const u32* __restrict__ idx,
                                                     u64 nl, int bits, const u64* __restrict__ csr,
                                                     const u32* __restrict__ adj,
                                                     unsigned long long* __restrict__ total) {
  __shared__ u32 T[kCqSlice * kBlock];
  const u64 i = (u64)blockIdx.x * kBlock + threadIdx.x;
  u64 c = 0;
  if (i < nl) {
    const u64 k = okeys[idx[i]];
    const u32 u = (u32)(k >> bits), v = (u32)(k & ((1ull << bits) - 1));
    u64 p = csr[u], pe = csr[u + 1], q = csr[v], qe = csr[v + 1];
    u32 nt = 0;
    while (p < pe && q < qe) {  // |T| <= min(d+u, d+v) <= 32
      const u32 x = adj[p], y = adj[q];
      if (x == y) T[nt++ * kBlock + threadIdx.x] = x;
      p += x <= y;
      q += y <= x;
    }
    for (u32 j = 0; j < nt; j++) {
      const u32 w = T[j * kBlock + threadIdx.x];
      const u64 s = csr[w], se = csr[w + 1];
      const u64 dw = se - s;
      if (dw <= 4u * nt) {
        // merge N+(w) with the slice
        u64 r = s;
        u32 t = 0;
        while (r < se && t < nt) {
          const u32 x = adj[r], y = T[t * kBlock + threadIdx.x];
          c += x == y;
          r += x <= y;
          t += y <= x;
        }
      } else {
        for (u32 t = 0; t < nt; t++) c += cq_has(adj, s, se, T[t * kBlock + threadIdx.x]);
      }
    }
  }
  c = cq_sum(c);
  if (cq_lane() == 0 && c) atomicAdd(total, (unsigned long long)c);
}

// One wave per heavy edge (grid-stride over the heavy list).
__global__ __launch_bounds__(kBlock) void k_cq_heavy(const u64* __restrict__ okeys, const u32* __restrict__ idx,
                                                     u64 nh, int bits, const u64* __restrict__ csr,
                                                     const u32* __restrict__ adj, u32 tcap,
                                                     unsigned long long* __restrict__ total) {
  __shared__ u32 s_t[kWaves][kCqCap];
  __shared__ u64 s_incl[kWaves][64];  // inclusive prefix of the batch's out-degrees
  __shared__ u64 s_start[kWaves][64];
  const u32 lane = cq_lane(), wv = cq_wave();
  u32* T = s_t[wv];
  const u64 nwaves = (u64)gridDim.x * kWaves;
  const u64 lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  u64 c = 0;
  for (u64 h = (u64)blockIdx.x * kWaves + wv; h < nh; h += nwaves) {
    const u64 k = okeys[idx[h]];
    const u32 u = (u32)(k >> bits), v = (u32)(k & ((1ull << bits) - 1));
    u64 s0 = csr[u], s1 = csr[u + 1], l0 = csr[v], l1 = csr[v + 1];
    if (s1 - s0 > l1 - l0) {
      u64 t0 = s0, t1 = s1;
      s0 = l0; s1 = l1; l0 = t0; l1 = t1;
    }
    const u64 ns = s1 - s0;
    // T = the shorter list's values found in the longer one, in order
    u64 nt = 0;
    for (u64 b = 0; b < ns; b += 64) {
      const u64 i = b + lane;
      const bool hit = i < ns && cq_has(adj, l0, l1, adj[s0 + i]);
      const u64 m = __ballot(hit);
      const u64 at = nt + __popcll(m & lt);
      if (hit && at < tcap) T[at] = adj[s0 + i];
      nt += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const bool in_lds = nt <= tcap;
    // the pairs (w, y), w in T, y in N+(w), 64 members of T at a time
    const u64 nsrc = in_lds ? nt : ns;
    for (u64 b = 0; b < nsrc; b += 64) {
      const u64 i = b + lane;
      bool valid = false;
      u32 w = 0;
      if (in_lds) {
        valid = i < nt;
        if (valid) w = T[i];
      } else if (i < ns) {
        w = adj[s0 + i];
        valid = cq_has(adj, l0, l1, w);
      }
      const u64 ws = valid ? csr[w] : 0;
      const u64 dw = valid ? csr[w + 1] - ws : 0;
      const u64 incl = cq_incl_scan(dw);
      s_incl[wv][lane] = incl;
      s_start[wv][lane] = ws;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const u64 npairs = __shfl(incl, 63, 64);
      for (u64 p = lane; p < npairs; p += 64) {
        // owner: the first lane whose inclusive prefix exceeds p
        u32 lo = 0, hi = 63;
        while (lo < hi) {
          const u32 mid = (lo + hi) >> 1;
          if (s_incl[wv][mid] > p) hi = mid;
          else lo = mid + 1;
        }
        const u64 before = lo ? s_incl[wv][lo - 1] : 0;
        const u32 y = adj[s_start[wv][lo] + (p - before)];
        bool mem;
        if (in_lds) {
          u32 a = 0, e = (u32)nt;
          while (a < e) {
            const u32 mid = (a + e) >> 1;
            if (T[mid] < y) a = mid + 1;
            else e = mid;
          }
          mem = a < nt && T[a] == y;
        } else {
          mem = cq_has(adj, s0, s1, y) && cq_has(adj, l0, l1, y);
        }
        c += mem;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
  c = cq_sum(c);
  if (lane == 0 && c) atomicAdd(total, (unsigned long long)c);
}

// Count the 4-cliques of the undirected graph (a[i], b[i]), i < n, given in
// device memory (duplicates, both directions and self-loops allowed).
// force_heavy sends every edge to the wave path; tcap (1 .. 1024) bounds the
// wave path's LDS table of T (tests use both to reach every path).
u64 clique4_count(const Exec& x, const u32* a, const u32* b, u64 n, bool force_heavy, u32 tcap) {
  if (!n) return 0;
  if (n >= (1ull << 32)) throw BudgetError{n};  // edge indices are u32
  if (tcap < 1 || tcap > (u32)kCqCap) tcap = kCqCap;
  // ids < nv, packed with `bits` bits per endpoint
  DBuf<u64> mx(x.pool, 1);
  memset_async(x, mx.get(), 0, 8);
  k_cq_maxid<<<cq_grid(n), kBlock, 0, x.stream>>>(a, b, n, (u32*)mx.get());
  SK_HIP(hipGetLastError());
  const u64 nv = (read_u64(x, mx.get()) & 0xffffffffull) + 1;
  int bits = 1;
  while ((1ull << bits) < nv) bits++;
  const u64 sentinel = bits == 32 ? ~0ull : ((1ull << (2 * bits)) - 1);
  // simple graph: canonical pairs, sorted, duplicates and self-loops dropped
  DBuf<u64> k1(x.pool, n), k2(x.pool, n);
  u64 m = 0;
  u64* keys = nullptr;  // the simple graph's canonical keys
  u64* spare = nullptr;
  {
    Scope sc(x.timer, "k4_canon", KC_MISC, (double)n * 16.0, n);
    k_cq_canon<<<cq_grid(n), kBlock, 0, x.stream>>>(a, b, n, bits, k1.get());
    SK_HIP(hipGetLastError());
  }
  u64* sk = nullptr;
  radix_sort<u64>(x, k1.get(), k2.get(), nullptr, nullptr, nullptr, nullptr, n, 2 * bits, &sk, nullptr, nullptr);
  spare = sk == k1.get() ? k2.get() : k1.get();
  m = unique_runs<u64>(x, sk, nullptr, nullptr, n, spare, nullptr, nullptr);
  keys = spare;
  spare = sk;
  if (m && read_u64(x, keys + m - 1) == sentinel) m--;
  if (!m) return 0;
  // orientation by (degree, id), then a CSR of the oriented edges
  DBuf<u32> deg(x.pool, nv);
  memset_async(x, deg.get(), 0, nv * 4);
  {
    Scope sc(x.timer, "k4_orient", KC_MISC, (double)m * 24.0, m);
    k_cq_degree<<<cq_grid(m), kBlock, 0, x.stream>>>(keys, m, bits, deg.get());
    SK_HIP(hipGetLastError());
    k_cq_orient<<<cq_grid(m), kBlock, 0, x.stream>>>(keys, m, bits, deg.get(), spare);
    SK_HIP(hipGetLastError());
  }
  u64* okeys = nullptr;
  radix_sort<u64>(x, spare, keys, nullptr, nullptr, nullptr, nullptr, m, 2 * bits, &okeys, nullptr, nullptr);
  DBuf<u64> csr(x.pool, nv + 1);
  build_csr<u64>(x, okeys, m, bits, nv, csr.get());
  DBuf<u32> from(x.pool, m), adj(x.pool, m);
  u32* cols[2] = {from.get(), adj.get()};
  unpack_cols<u64>(x, okeys, m, 2, bits, cols);
  // light and heavy edge lists
  DBuf<u8> hv(x.pool, m), lt(x.pool, m);
  DBuf<u64> hpos(x.pool, m), lpos(x.pool, m), tot(x.pool, 2);
  k_cq_classify<<<cq_grid(m), kBlock, 0, x.stream>>>(okeys, m, bits, csr.get(), force_heavy ? 1 : 0, hv.get(),
                                                     lt.get());
  SK_HIP(hipGetLastError());
  scan_u8(x, hv.get(), hpos.get(), m, tot.get());
  scan_u8(x, lt.get(), lpos.get(), m, tot.get() + 1);
  u64 cnt[2];
  read_u64s(x, tot.get(), cnt, 2);
  const u64 nh = cnt[0], nl = cnt[1];
  DBuf<u32> hidx(x.pool, nh ? nh : 1), lidx(x.pool, nl ? nl : 1);
  if (nh) {
    k_cq_gather<<<cq_grid(m), kBlock, 0, x.stream>>>(hv.get(), hpos.get(), m, hidx.get());
    SK_HIP(hipGetLastError());
  }
  if (nl) {
    k_cq_gather<<<cq_grid(m), kBlock, 0, x.stream>>>(lt.get(), lpos.get(), m, lidx.get());
    SK_HIP(hipGetLastError());
  }
  DBuf<u64> acc(x.pool, 1);
  memset_async(x, acc.get(), 0, 8);
  if (nl) {
    // per edge its key, index and four CSR words
    Scope sc(x.timer, "k4_light", KC_LEAPFROG, (double)nl * 44.0, nl);
    k_cq_light<<<cq_grid(nl), kBlock, 0, x.stream>>>(okeys, lidx.get(), nl, bits, csr.get(), adj.get(),
                                                     (unsigned long long*)acc.get());
    SK_HIP(hipGetLastError());
  }
  if (nh) {
    Scope sc(x.timer, "k4_heavy", KC_LEAPFROG, (double)nh * 44.0, nh);
    const u64 blocks = std::min<u64>((nh + kWaves - 1) / kWaves, 8192);
    k_cq_heavy<<<(unsigned)blocks, kBlock, 0, x.stream>>>(okeys, hidx.get(), nh, bits, csr.get(), adj.get(), tcap,
                                                          (unsigned long long*)acc.get());
    SK_HIP(hipGetLastError());
  }
  return read_u64(x, acc.get());
}

}  // namespace sk
