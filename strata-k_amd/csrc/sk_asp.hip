// sk_asp.hip — the grounding simplification filter over a ground program in
// CSR form (crates/strata-gpu/src/asp.rs:53-99, 148-355).
//
// A rule `h :- pos, not neg` dies when a positive body atom is impossible or
// a negative one is certain; in a kept rule, positive atoms that are certain
// and negative atoms that are impossible are dropped (trivially true).  Body
// literals are signed: +(atom+1) positive, -(atom+1) negative.
//
// Count pass (one thread per rule: alive flag and kept-literal count, with
// every literal and offset bounds-checked), two device scans (output rule
// positions, output literal offsets), emit pass.  Kept rules and literals
// keep their input order.
#include "sk_dev.h"

namespace sk {

// Per-atom state: 0 impossible, 1 possible and not certain, 2 certain.
__global__ void k_asp_state(const u8* __restrict__ poss, const u8* __restrict__ cert, u64 n, u8* __restrict__ st) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st[i] = cert[i] ? 2 : (poss[i] ? 1 : 0);
}

// -1: the rule dies; 0: drop the literal; 1: keep it; 2: malformed literal.
__device__ __forceinline__ int asp_decide(int32_t lit, const u8* st, u64 n_atoms) {
  if (lit == 0 || lit == INT32_MIN) return 2;
  const u64 a = (u64)(lit > 0 ? lit : -lit) - 1;
  if (a >= n_atoms) return 2;
  const u8 s = st[a];
  if (lit > 0) return s == 0 ? -1 : (s == 2 ? 0 : 1);
  return s == 2 ? -1 : (s == 0 ? 0 : 1);
}

__global__ void k_asp_count(const u32* __restrict__ bs, const int32_t* __restrict__ lits, u64 n_rules, u64 n_lits,
                            const u8* __restrict__ st, u64 n_atoms, u8* __restrict__ alive, u32* __restrict__ kept,
                            u32* __restrict__ bad) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rules) return;
  const u64 s = bs[i], e = bs[i + 1];
  u8 al = 1;
  u32 k = 0;
  if (e < s || e > n_lits) {
    atomicOr(bad, 1u);
    al = 0;
  } else {
    for (u64 j = s; j < e; j++) {
      const int d = asp_decide(lits[j], st, n_atoms);
      if (d == 2) {
        atomicOr(bad, 1u);
        al = 0;
        break;
      }
      if (d < 0) {
        al = 0;
        break;
      }
      k += (u32)d;
    }
  }
  alive[i] = al;
  kept[i] = al ? k : 0;
}

__global__ void k_asp_emit(const int32_t* __restrict__ head, const u32* __restrict__ bs,
                           const int32_t* __restrict__ lits, u64 n_rules, const u8* __restrict__ st, u64 n_atoms,
                           const u8* __restrict__ alive, const u64* __restrict__ rpos, const u64* __restrict__ lpos,
                           int32_t* __restrict__ out_head, u32* __restrict__ out_bs, int32_t* __restrict__ out_lits) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rules || !alive[i]) return;
  const u64 r = rpos[i];
  u64 o = lpos[i];
  out_head[r] = head[i];
  out_bs[r] = (u32)o;
  for (u64 j = bs[i], e = bs[i + 1]; j < e; j++)
    if (asp_decide(lits[j], st, n_atoms) == 1) out_lits[o++] = lits[j];
}

static unsigned grid_a(u64 n) {
  const u64 g = (n + kBlock - 1) / kBlock;
  if (g > 0x7fffffffull) throw BudgetError{n};
  return (unsigned)(g ? g : 1);
}

// Device-resident inputs and outputs (caller-sized: n_rules heads, n_rules + 1
// offsets, n_lits literals).  Returns false when a literal or an offset is
// malformed (outputs then undefined).
bool asp_simplify(const Exec& x, const int32_t* head, const u32* bs, const int32_t* lits, u64 n_rules, u64 n_lits,
                  const u8* poss, const u8* cert, u64 n_atoms, int32_t* out_head, u32* out_bs, int32_t* out_lits,
                  u64* out_rules, u64* out_lits_n) {
  DBuf<u8> st(x.pool, n_atoms ? n_atoms : 1), alive(x.pool, n_rules ? n_rules : 1);
  DBuf<u32> kept(x.pool, n_rules ? n_rules : 1), bad(x.pool, 1);
  DBuf<u64> rpos(x.pool, n_rules ? n_rules : 1), lpos(x.pool, n_rules ? n_rules : 1), tot(x.pool, 2);
  memset_async(x, bad.get(), 0, 4);
  if (n_atoms) k_asp_state<<<grid_a(n_atoms), kBlock, 0, x.stream>>>(poss, cert, n_atoms, st.get());
  u64 hv[2] = {0, 0};
  if (n_rules) {
    {
      // per rule: its end offset read, alive flag and kept count written;
      // per literal: the literal and its atom's state read
      Scope sc(x.timer, "asp_count", KC_JOIN, (double)n_rules * 9.0 + (double)n_lits * 5.0, n_rules);
      k_asp_count<<<grid_a(n_rules), kBlock, 0, x.stream>>>(bs, lits, n_rules, n_lits, st.get(), n_atoms, alive.get(),
                                                            kept.get(), bad.get());
      SK_HIP(hipGetLastError());
    }
    scan_u8(x, alive.get(), rpos.get(), n_rules, tot.get());
    scan_u32(x, kept.get(), lpos.get(), n_rules, tot.get() + 1);
    u32 b4 = 0;
    read_u64s(x, tot.get(), hv, 2);
    SK_HIP(hipMemcpyAsync(&b4, bad.get(), 4, hipMemcpyDeviceToHost, x.stream));
    SK_HIP(hipStreamSynchronize(x.stream));
    if (b4) return false;
    // per rule: the alive flag; per kept rule: both positions, both offsets,
    // the head read and head + offset written; per kept literal: literal and
    // state read, literal written (dropped literals of kept rules, also
    // read, are not charged)
    Scope sc(x.timer, "asp_emit", KC_JOIN, (double)n_rules + (double)hv[0] * 36.0 + (double)hv[1] * 9.0, n_rules);
    k_asp_emit<<<grid_a(n_rules), kBlock, 0, x.stream>>>(head, bs, lits, n_rules, st.get(), n_atoms, alive.get(),
                                                         rpos.get(), lpos.get(), out_head, out_bs, out_lits);
    SK_HIP(hipGetLastError());
  }
  const u32 last = (u32)hv[1];
  SK_HIP(hipMemcpyAsync(out_bs + hv[0], &last, 4, hipMemcpyHostToDevice, x.stream));
  SK_HIP(hipStreamSynchronize(x.stream));
  *out_rules = hv[0];
  *out_lits_n = hv[1];
  return true;
}

}  // namespace sk
