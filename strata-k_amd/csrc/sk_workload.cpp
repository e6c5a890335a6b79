// sk_workload.cpp — deterministic synthetic graph generators for the
// benchmark configurations (host code, part of the library so the CPU and GPU
// runs read the same EDB bytes).
//
// Randomness is counter-based: draw k of stream s is mix(seed, s, k), a
// splitmix64 finaliser over a Weyl sequence, so every generator gives the same
// output for any thread count.  Bounded integers use Lemire's multiply-shift
// with rejection (unbiased).
//
// The reference has no generator for these workloads; its examples build
// chains and hub graphs by hand (crates/strata-gpu/examples/bench.rs:35-42,
// tri.rs:29-60).  RMAT follows the usual recursive quadrant draw with
// probabilities (a, b, c, d = 1-a-b-c), most significant bit first.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

inline uint64_t mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Draw k of stream s under seed.
inline uint64_t draw(uint64_t seed, uint64_t s, uint64_t k) {
  return mix(mix(seed ^ (s * 0xd1342543de82ef95ULL)) + k * 0x9e3779b97f4a7c15ULL);
}

// Sequential stream used where rejection needs "the next" value.
struct Stream {
  uint64_t seed, s, k;
  uint64_t next() { return draw(seed, s, k++); }
  // Unbiased integer in [0, range) (Lemire 2019, "Fast random integer
  // generation in an interval").
  uint64_t below(uint64_t range) {
    if (range == 0) return 0;
    for (;;) {
      uint64_t x = next();
      unsigned __int128 m = (unsigned __int128)x * range;
      uint64_t l = (uint64_t)m;
      if (l < range) {
        uint64_t t = (0 - range) % range;
        if (l < t) continue;
      }
      return (uint64_t)(m >> 64);
    }
  }
};

unsigned nthreads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 4;
  return std::min(t, 16u);
}

template <class F>
void parallel_for(uint64_t n, F f) {
  unsigned T = nthreads();
  if (n < 100000 || T <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  uint64_t chunk = (n + T - 1) / T;
  for (unsigned t = 0; t < T; t++) {
    uint64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// Sort u64 keys: partition by the top significant bits into buckets, then
// sort the buckets in parallel.
void parallel_sort(uint64_t* keys, uint64_t n) {
  if (n < (1u << 20)) {
    std::sort(keys, keys + n);
    return;
  }
  uint64_t mx = 0;
  for (uint64_t i = 0; i < n; i++) mx = std::max(mx, keys[i]);
  int bits = mx ? 64 - __builtin_clzll(mx) : 1;
  const int B = 10;  // 1024 buckets
  int shift = bits > B ? bits - B : 0;
  const uint64_t NB = 1u << B;
  unsigned T = nthreads();
  std::vector<std::vector<uint64_t>> cnt(T, std::vector<uint64_t>(NB, 0));
  uint64_t chunk = (n + T - 1) / T;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; t++) {
    th.emplace_back([&, t] {
      uint64_t lo = t * chunk, hi = std::min(n, lo + chunk);
      for (uint64_t i = lo; i < hi; i++) cnt[t][keys[i] >> shift]++;
    });
  }
  for (auto& x : th) x.join();
  th.clear();
  std::vector<uint64_t> bstart(NB + 1, 0);
  std::vector<std::vector<uint64_t>> off(T, std::vector<uint64_t>(NB, 0));
  uint64_t acc = 0;
  for (uint64_t b = 0; b < NB; b++) {
    bstart[b] = acc;
    for (unsigned t = 0; t < T; t++) {
      off[t][b] = acc;
      acc += cnt[t][b];
    }
  }
  bstart[NB] = acc;
  std::vector<uint64_t> tmp(n);
  for (unsigned t = 0; t < T; t++) {
    th.emplace_back([&, t] {
      uint64_t lo = t * chunk, hi = std::min(n, lo + chunk);
      std::vector<uint64_t> o = off[t];
      for (uint64_t i = lo; i < hi; i++) tmp[o[keys[i] >> shift]++] = keys[i];
    });
  }
  for (auto& x : th) x.join();
  th.clear();
  std::atomic<uint64_t> next{0};
  for (unsigned t = 0; t < T; t++) {
    th.emplace_back([&] {
      for (;;) {
        uint64_t b = next.fetch_add(1);
        if (b >= NB) break;
        std::sort(tmp.begin() + bstart[b], tmp.begin() + bstart[b + 1]);
      }
    });
  }
  for (auto& x : th) x.join();
  std::memcpy(keys, tmp.data(), n * sizeof(uint64_t));
}

inline uint32_t thresh(double p) {
  double x = std::ldexp(p, 32);
  if (x >= 4294967295.0) return 0xffffffffu;
  if (x <= 0) return 0;
  return (uint32_t)x;
}

}  // namespace

extern "C" {

// Uniform directed graph: draws (src, dst) uniformly from [0, n_nodes)^2,
// rejecting self-loops and already-drawn pairs, until n_edges distinct edges
// exist (or the complete graph is reached).  Edges are written in acceptance
// order.  Returns the number written.
uint64_t sk_wl_uniform_graph(uint64_t seed, uint32_t n_nodes, uint64_t n_edges, uint32_t* src,
                             uint32_t* dst) {
  if (n_nodes < 2) return 0;
  uint64_t maxe = (uint64_t)n_nodes * (n_nodes - 1);
  if (n_edges > maxe) n_edges = maxe;
  std::unordered_set<uint64_t> seen;
  seen.reserve(n_edges * 2);
  Stream st{seed, 1, 0};
  uint64_t m = 0;
  while (m < n_edges) {
    uint32_t a = (uint32_t)st.below(n_nodes);
    uint32_t b = (uint32_t)st.below(n_nodes);
    if (a == b) continue;
    uint64_t k = ((uint64_t)a << 32) | b;
    if (!seen.insert(k).second) continue;
    src[m] = a;
    dst[m] = b;
    m++;
  }
  return m;
}

// RMAT draws: n_draws edges over 2^scale ids (quadrant probabilities a, b, c,
// d = 1-a-b-c), keeping only draws with both ids < n_limit (n_limit = 0 keeps
// all).  Draw i uses stream (2, i), so the output is thread-count
// independent.  Self-loops and duplicates are kept; sk_wl_normalize removes
// them.  Returns the number of draws written.
uint64_t sk_wl_rmat(uint64_t seed, uint32_t scale, uint64_t n_draws, double a, double b, double c,
                    uint32_t n_limit, uint32_t* src, uint32_t* dst) {
  const uint32_t ta = thresh(a), tab = thresh(a + b), tabc = thresh(a + b + c);
  std::vector<uint8_t> keep(n_limit ? n_draws : 0);
  parallel_for(n_draws, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; i++) {
      uint32_t u = 0, v = 0;
      uint64_t r = 0;
      for (uint32_t l = 0; l < scale; l++) {
        if ((l & 1) == 0) r = draw(seed, 2, i * 32 + (l >> 1));
        uint32_t x = (l & 1) ? (uint32_t)(r >> 32) : (uint32_t)r;
        uint32_t bu = 0, bv = 0;
        if (x < ta) {
        } else if (x < tab) {
          bv = 1;
        } else if (x < tabc) {
          bu = 1;
        } else {
          bu = 1;
          bv = 1;
        }
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      src[i] = u;
      dst[i] = v;
      if (n_limit) keep[i] = (u < n_limit && v < n_limit);
    }
  });
  if (!n_limit) return n_draws;
  uint64_t m = 0;
  for (uint64_t i = 0; i < n_draws; i++)
    if (keep[i]) {
      src[m] = src[i];
      dst[m] = dst[i];
      m++;
    }
  return m;
}

// Sort by (src, dst) and drop duplicates; optionally drop self-loops and,
// with undirected != 0, first map each pair to (min, max).  In place;
// returns the new length.
uint64_t sk_wl_normalize(uint32_t* src, uint32_t* dst, uint64_t n, int drop_self_loops,
                         int undirected) {
  std::vector<uint64_t> k(n);
  parallel_for(n, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; i++) {
      uint32_t a = src[i], b = dst[i];
      if (undirected && b < a) std::swap(a, b);
      k[i] = ((uint64_t)a << 32) | b;
    }
  });
  if (drop_self_loops) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++)
      if ((uint32_t)(k[i] >> 32) != (uint32_t)k[i]) k[m++] = k[i];
    n = m;
  }
  parallel_sort(k.data(), n);
  uint64_t m = 0;
  for (uint64_t i = 0; i < n; i++)
    if (i == 0 || k[i] != k[i - 1]) k[m++] = k[i];
  parallel_for(m, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; i++) {
      src[i] = (uint32_t)(k[i] >> 32);
      dst[i] = (uint32_t)k[i];
    }
  });
  return m;
}

// Weights uniform in [lo, hi]: weight i uses stream (3, i) with rejection.
void sk_wl_weights(uint64_t seed, uint64_t n, int64_t lo, int64_t hi, int64_t* out) {
  uint64_t range = (uint64_t)(hi - lo) + 1;
  parallel_for(n, [&](uint64_t a, uint64_t b) {
    for (uint64_t i = a; i < b; i++) {
      Stream st{seed ^ 0x5bd1e995u, 3 + (i << 8), 0};
      out[i] = lo + (int64_t)st.below(range);
    }
  });
}

// k distinct ids uniform in [0, n_nodes), in draw order.
uint32_t sk_wl_sample_distinct(uint64_t seed, uint32_t n_nodes, uint32_t k, uint32_t* out) {
  if (k > n_nodes) k = n_nodes;
  std::unordered_set<uint32_t> seen;
  Stream st{seed, 4, 0};
  uint32_t m = 0;
  while (m < k) {
    uint32_t x = (uint32_t)st.below(n_nodes);
    if (seen.insert(x).second) out[m++] = x;
  }
  return m;
}

// Orient a simple undirected edge list (each pair once) from lower to higher
// (degree, id), then sort by (src, dst).  In place.
void sk_wl_orient_by_degree(uint32_t* a, uint32_t* b, uint64_t n, uint32_t n_nodes) {
  std::vector<uint32_t> deg(n_nodes, 0);
  for (uint64_t i = 0; i < n; i++) {
    deg[a[i]]++;
    deg[b[i]]++;
  }
  parallel_for(n, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; i++) {
      uint32_t u = a[i], v = b[i];
      bool swap = deg[u] > deg[v] || (deg[u] == deg[v] && u > v);
      if (swap) {
        a[i] = v;
        b[i] = u;
      }
    }
  });
  sk_wl_normalize(a, b, n, 0, 0);
}

// Log-uniform integer in [lo, hi]: exp(uniform(log lo, log hi)), rounded
// down, from stream (5, i).
uint64_t sk_wl_log_uniform(uint64_t seed, uint64_t i, uint64_t lo, uint64_t hi) {
  uint64_t x = draw(seed, 5, i);
  double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
  double v = std::exp(std::log((double)lo) + u * (std::log((double)hi + 1.0) - std::log((double)lo)));
  uint64_t r = (uint64_t)v;
  if (r < lo) r = lo;
  if (r > hi) r = hi;
  return r;
}

// The reference's skewed triangle workload (crates/strata-gpu/examples/tri.rs:
// 27-56): a clique on h hubs, b bridge nodes joined to every hub, and for
// each hub d random neighbours drawn from its own block of the remaining ids
// below n by the example's LCG (state * 6364136223846793005 +
// 1442695040888963407, value (state >> 33) % block, from
// 0x123456789abcdef0).  src/dst hold h(h-1)/2 + h*b + h*min(d, block)
// entries; returns the count written.
uint64_t sk_wl_tri_hubs(uint32_t h, uint32_t b, uint32_t d, uint32_t n, uint32_t* src, uint32_t* dst) {
  uint64_t m = 0;
  for (uint32_t i = 0; i < h; i++)
    for (uint32_t j = i + 1; j < h; j++) {
      src[m] = i;
      dst[m++] = j;
    }
  for (uint32_t k = 0; k < b; k++)
    for (uint32_t hub = 0; hub < h; hub++) {
      src[m] = hub;
      dst[m++] = h + k;
    }
  uint64_t state = 0x123456789abcdef0ull;
  const uint32_t lo = h + b;
  const uint32_t block = std::max<uint32_t>(h ? (n - lo) / h : (n - lo), 1u);
  for (uint32_t hub = 0; hub < h; hub++) {
    const uint32_t base = lo + hub * block;
    for (uint32_t q = 0; q < std::min(d, block); q++) {
      state = state * 6364136223846793005ull + 1442695040888963407ull;
      src[m] = hub;
      dst[m++] = base + (uint32_t)((state >> 33) % block);
    }
  }
  return m;
}

// Uniform double in [0, 1) from stream (6, i) (Poisson arrival gaps).
double sk_wl_uniform01(uint64_t seed, uint64_t i) {
  return (double)(draw(seed, 6, i) >> 11) * (1.0 / 9007199254740992.0);
}

}  // extern "C"
