// sk_dev.h — shared definitions for the gfx950 engine: key types, the device
// memory pool, stream/event helpers and the launch wrappers of every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace sk {

using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;
using u128 = unsigned __int128;

constexpr int kBlock = 256;  // 4 wavefronts of 64
constexpr int kWaves = kBlock / 64;

// ------------------------------------------------------------------ errors

struct DeviceError {
  hipError_t err;
  std::string where;
};

#define SK_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) throw ::sk::DeviceError{e_, std::string(#call) + " at " __FILE__ ":" + std::to_string(__LINE__)}; \
  } while (0)

struct BudgetError {
  u64 want;
};

// Device error record written by kernels: the first overflow of a checked
// Trop addition wins (atomicCAS on `code`).  `widen` is set (atomicOr) by a
// derivation that does not fit a narrow cell; keeping it in the same record
// lets the host read both with one copy per round.
struct DevErr {
  int32_t code;  // 0 none, 2 overflow
  uint32_t widen;
  i64 lhs, rhs;
};

// ----------------------------------------------------------- memory pool

// Caching device allocator.  Blocks are rounded up to a size class and kept
// on free lists; hipFree (which synchronises the device) only happens when
// the budget is exceeded or the pool is destroyed.
class Pool {
 public:
  explicit Pool(u64 budget) : budget_(budget) {}
  ~Pool();
  void* alloc(u64 bytes);
  void free(void* p);
  void trim();  // return cached blocks to the runtime
  u64 in_use() const { return in_use_; }
  u64 held() const { return held_; }
  u64 peak() const { return peak_; }
  u64 budget() const { return budget_; }
  void set_budget(u64 b) { budget_ = b; }

 private:
  static u64 size_class(u64 b);
  u64 budget_;
  u64 in_use_ = 0, held_ = 0, peak_ = 0;
  std::multimap<u64, void*> free_;
  std::map<void*, u64> live_;
};

// Typed RAII buffer on a pool.
template <class T>
struct DBuf {
  Pool* pool = nullptr;
  T* p = nullptr;
  u64 n = 0;  // capacity in elements
  DBuf() = default;
  DBuf(Pool* pl, u64 count) { reset(pl, count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      pool = o.pool; p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void reset(Pool* pl, u64 count) {
    release();
    pool = pl;
    n = count;
    p = count ? static_cast<T*>(pl->alloc(count * sizeof(T))) : nullptr;
  }
  // Grow capacity (contents are not preserved).
  void ensure(Pool* pl, u64 count) {
    if (count > n || pool != pl) reset(pl, count);
  }
  void release() {
    if (p && pool) pool->free(p);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

// Set bit v of a device bitmap, for a value that may repeat heavily across
// the launch (a sorted column's runs, hub ids).  A lane whose wave neighbour
// below holds the same value leaves it to that lane; the others read the
// word at device scope (not a stale L1 copy) and skip the atomic when the
// bit is already set.  Must be reached by every lane of the wave.
__device__ __forceinline__ void mark_bit_once(u32* bitmap, u32 v, bool on) {
  const u32 lane = threadIdx.x & 63;
  const u32 pv = __shfl_up(v, 1, 64);
  const bool pon = __shfl_up(on ? 1 : 0, 1, 64) != 0;
  if (!on || (lane > 0 && pon && pv == v)) return;
  const u32 m = 1u << (v & 31);
  if (__hip_atomic_load(&bitmap[v >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m) return;
  atomicOr(&bitmap[v >> 5], m);
}

// ------------------------------------------------------------- profiling

enum KClass { KC_SORT = 0, KC_DEDUP = 1, KC_MERGE = 2, KC_JOIN = 3, KC_LEAPFROG = 4, KC_MISC = 5, KC_N = 6 };

struct KStat {
  u64 launches = 0;
  double ms = 0, bytes = 0;
  u64 items = 0;
};

// Per-stream launch timer: when enabled, brackets single kernel launches
// with HIP events on the launch stream and accumulates per kernel name and
// per kernel class.  A scope never spans a host synchronisation.
struct KRecord {
  std::string name;
  int cls = KC_MISC;
  u64 launches = 0;
  double ms = 0, bytes = 0;
  u64 items = 0;
};

class Timer {
 public:
  bool enabled = false;
  hipStream_t stream = nullptr;
  void begin(const char* name, int cls, double bytes, u64 items);
  void end();
  // Credit bytes / items to a kernel record without a launch (work whose
  // size is known only after the launch was timed).
  void credit(const char* name, int cls, double bytes, u64 items);
  void collect();  // synchronise and fold pending event pairs into stats
  void reset();
  KStat stats[KC_N];
  std::map<std::string, KRecord> kernels;

 private:
  struct Pending {
    hipEvent_t a, b;
    const char* name;
    int cls;
    double bytes;
    u64 items;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> spare_;
  bool open_ = false;
  Pending cur_{};
  hipEvent_t get_event();
};

struct Scope {
  Timer* t;
  bool on;
  Scope(Timer* tm, const char* name, int cls, double bytes, u64 items) : t(tm), on(tm && tm->enabled) {
    if (on) t->begin(name, cls, bytes, items);
  }
  ~Scope() {
    if (on) t->end();
  }
};

// Radix sort autotuner state of a context: per key type (u64 / u128), with
// or without weights, and size class (2^16, 2^18, ..., 2^26 records), the
// measured time of one pass of every candidate configuration (digit bits x
// items per thread x threads per block, reduce-then-scan or single-sweep
// passes).  A sort takes the candidate with the least passes x time for its
// key width; a class is timed on first use (`lazy`) or all at once by
// radix_autotune, and the timings are shared by every context of the
// process on the same device.  `forced` >= 0 pins one candidate.
struct RadixTune {
  static constexpr int kClasses = 6, kMinLog2 = 16, kMaxCands = 20;
  bool lazy = true;
  int forced = -1;
  bool tuned[2][2][kClasses] = {};
  float pass_ms[2][2][kClasses][kMaxCands] = {};
  int chosen[2][2][kClasses] = {};
};

// Execution context handed to every primitive.
struct Exec {
  Pool* pool;
  hipStream_t stream;
  Timer* timer;
  DevErr* err;  // device error record
  u64* host = nullptr;          // pinned words for small device-to-host reads
  hipEvent_t read_ev = nullptr;  // polled behind those reads
  RadixTune* tune = nullptr;    // radix configuration choices (null: the default)
  // Pinned host word (device-visible pointer and host pointer) that a
  // single-sweep radix pass sets when a look-back exceeds its spin bound;
  // read_u64s and the engine's status reads raise DeviceError when it is set.
  u32* fault_d = nullptr;
  const volatile u32* fault_h = nullptr;
};

// Throws DeviceError when a kernel has set the context's fault word.
void check_fault(const Exec& x);

// ------------------------------------------------------ primitive wrappers

// Exclusive scan of n values; out[n] receives nothing, *total (device) the sum.
void scan_u32(const Exec& x, const u32* in, u64* out, u64 n, u64* d_total);
void scan_u64(const Exec& x, const u64* in, u64* out, u64 n, u64* d_total);
void scan_u8(const Exec& x, const u8* in, u64* out, u64 n, u64* d_total);
// Scan of n <= kScanOneMax values in one launch: total[0] = the sum and, when
// `side` is given, total[1] = the sum of side[0..n); total[2 + i] =
// extra[i] for i < nextra (a snapshot taken after every earlier launch on
// the stream, so one host read can fetch it with the sums).
constexpr u64 kScanOneMax = (u64)kBlock * 64;
void scan_u32_one(const Exec& x, const u32* in, u64* out, u64 n, u64* total, const u64* side,
                  const u64* extra = nullptr, int nextra = 0);

// LSD radix sort of keys (stable) on the low `key_bits` bits, optionally
// carrying i64 weights and u8 +inf flags.  Result in (keys, w, inf); the
// alt buffers are scratch of the same length.  With lo_bit > 0 the sort
// uses bits [lo_bit, lo_bit + key_bits) only; the keys must have no set bit
// at or above lo_bit + key_bits (the sort reads whole 8-bit digits).
template <class K>
void radix_sort(const Exec& x, K* keys, K* keys_alt, i64* w, i64* w_alt, u8* inf, u8* inf_alt, u64 n,
                int key_bits, K** out_keys, i64** out_w, u8** out_inf, int lo_bit = 0);

// Time every radix candidate for every size class now (see RadixTune).
void radix_autotune(const Exec& x, RadixTune* t);
int radix_candidates();
int radix_candidate_digit(int c);
int radix_candidate_items(int c, bool wide);
int radix_candidate_block(int c);
bool radix_candidate_sweep(int c);

// Collapse runs of equal keys of a sorted array.  Bool: keep one.  With
// weights: keep the minimum (+inf above every finite).  Returns the count.
template <class K>
u64 unique_runs(const Exec& x, const K* keys, const i64* w, const u8* inf, u64 n, K* out_keys,
                i64* out_w, u8* out_inf);

// Merge cand (sorted unique) into full (sorted unique).  fresh = cand rows
// absent from full, or (weights) strictly below full's weight; union = full ∪
// cand with the minimum weight.  Counts are returned; when *n_fresh == 0 the
// union is not written (it equals full).
template <class K>
void merge_diff(const Exec& x, const K* a, const i64* aw, const u8* ainf, u64 m, const K* b,
                const i64* bw, const u8* binf, u64 n, DBuf<K>* u_keys, DBuf<i64>* u_w,
                DBuf<u8>* u_inf, u64* n_union, DBuf<K>* f_keys, DBuf<i64>* f_w, DBuf<u8>* f_inf,
                u64* n_fresh, bool weighted, bool has_inf);

// First-column CSR over sorted packed keys: csr[v] = first row whose leading
// column (key >> shift) is >= v, for v in [0, nvals].
template <class K>
void build_csr(const Exec& x, const K* keys, u64 n, int shift, u64 nvals, u64* csr);

// Unpack packed keys into u32 column arrays.
template <class K>
void unpack_cols(const Exec& x, const K* keys, u64 n, int arity, int bits, u32* const* cols);

constexpr int kMaxCols = 8;
struct Cols {
  u32* c[kMaxCols];
};
// Re-pack rows of a relation in a permuted column order: out column j = in column perm[j].
template <class K, class KO>
void repack(const Exec& x, const K* keys, u64 n, int arity, int bits, const int* perm, KO* out);

// Membership of packed keys (sorted unique) : flags[i] = key present.
template <class K>
void member(const Exec& x, const K* sorted, u64 ns, const K* q, u64 nq, u8* flags);

// Limits of the ASP filter's allocation-free variant (its static scratch),
// and that scratch's address on `device`, which must be current (resolved
// once per device with hipGetSymbolAddress, then cached).
constexpr u64 kAspMaxRules = 1ull << 27;
constexpr u64 kAspMaxAtoms = 1ull << 22;
u8* asp_scratch_address(int device);

void memset_async(const Exec& x, void* p, int v, u64 bytes);
void fill_u8(const Exec& x, u8* p, u8 v, u64 n);
void fill_u32(const Exec& x, u32* p, u32 v, u64 n);

u64 read_u64(const Exec& x, const u64* d);
// n <= 8 device words to the host with one copy and one synchronisation.
void read_u64s(const Exec& x, const u64* d, u64* h, int n);

}  // namespace sk
