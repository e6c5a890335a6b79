// sk_cells.h — dense cell state of a relation while its stratum runs.
//
// A relation whose column domains multiply out to a modest number of cells
// is kept as one cell per possible tuple instead of a sorted key list:
//   Bool  one presence bit per cell;
//   Trop  one u64 per cell holding the order-preserving code of the least
//         weight (finite w -> w ^ 2^63; kCellTop = absent), plus two rare-case
//         bitmaps: `maxf` (present with the finite weight INT64_MAX, whose
//         code collides with kCellTop) and `infb` (present with +inf).
// Derivations are ⊕-applied in place (atomic min / or); a derivation that
// strictly improves its cell sets the cell's bit in `touched`.  Every round
// is one Jacobi application: the rules read last round's delta list, never
// the cells, so updating in place during the round changes nothing they
// see; the next delta is the touched set with the cells' final values, which
// compacts in cell order, i.e. sorted by packed key.
//
// Column c of a cell index holds the column value's rank in its domain
// (identity when dom_bits[c] is null, else the number of domain codes below
// it: dom_pre[word] + popcount of the lower bits in the word); ranks are
// monotone in the code, so cell order equals key order.
#pragma once

#include <vector>

#include "sk_dev.h"

namespace sk {

constexpr u64 kCellTop = ~0ull;
constexpr u32 kCellTop32 = ~0u;
// Narrow weights below this fit a packed delta word beside a 6-bit lane.
constexpr u64 kPackLim = 1ull << 26;

// State of a device-driven run of delta rounds (a hipGraph of rounds that
// the host relaunches until `done`): each round's kernels do nothing once
// `done` is set; `stamp` is the last compaction's row stamp; `iters` counts
// the rounds that ran.
struct LoopState {
  u32 done;
  u32 stamp;  // Bool rounds: the compaction stamp; Trop rounds: the next round's kind (0 pull, 1 push)
  u64 iters;
  u32 any;    // the round's compaction found a touched cell
  u32 leave;  // Trop rounds: 1 the push join is cheaper next round, 2 a device
              // status (overflow, widen) to handle, 3 the round cap exceeded
  u64 n;      // Trop rounds: the last delta's size and its push cost
  u64 pc;
  DevErr st;  // the status that made the rounds leave
};
static_assert(sizeof(LoopState) == 64, "read back as eight words");

struct CellDesc {
  int on = 0;
  int trop = 0;
  int ncols = 0;
  u64 ncodes = 0;            // dictionary size: every valid code is below it
  int shift[kMaxCols] = {};  // cell = sum of rank_c << shift[c]
  int ebits[kMaxCols] = {};  // rank_c < 2^ebits[c]
  const u32* dom_bits[kMaxCols] = {};  // domain bitmap over codes, or null = identity
  const u64* dom_pre[kMaxCols] = {};   // per-word count of domain codes in earlier words
  const u32* dom_list[kMaxCols] = {};  // rank -> code (null = identity)
  // Batched runs (column 0 holds the program id p, whose code is p): a
  // column may instead be ranked within its program — its codes for program
  // p are pdom_codes[c][pdom_off[c][p] .. pdom_off[c][p+1]), sorted, and a
  // code's rank is its position in that segment.  Program ids lead the cell
  // order, so cell order is still key order.
  const u64* pdom_off[kMaxCols] = {};
  const u32* pdom_codes[kMaxCols] = {};
  u64 nparts = 0;
  u64* cell = nullptr;     // Trop weight codes
  // Narrow Trop cells: when every weight the run can hold is a non-negative
  // value below 2^32-1, cells hold the weight itself as u32 (kCellTop32 =
  // absent), halving cell traffic.  A derivation outside that range sets
  // *widen and is not applied; the engine then widens the cells and runs the
  // round again (⊕ is idempotent, so re-applying the round changes nothing
  // already applied).
  int narrow = 0;
  u32* cell32 = nullptr;
  u32* widen = nullptr;
  // Narrow derivations of weight >= narrow_lim set *widen (kCellTop32, or
  // 2^26 when the delta is also written packed, below).
  u32 narrow_lim = kCellTop32;
  u32* present = nullptr;  // Bool presence bits
  u32* maxf = nullptr;     // Trop: present with finite INT64_MAX
  u32* infb = nullptr;     // Trop: present with +inf
  u32* touched = nullptr;  // strictly improved in the current round
  // When set, a compaction of the touched cells (from cell 0) also writes one
  // record per row of 2^rlbits <= 64 cells: {the row's touched lanes, the
  // position of the row's first touched cell in the compacted list}.
  u64* rowrec = nullptr;
  u64 rnrows = 0;
  int rlbits = 0;
  // With rowrec: that compaction also writes each listed cell's lane
  // (cell mod 2^rlbits), in list order.
  u8* dlane = nullptr;
  // Narrow cells below 2^26 (narrow_lim): with rowrec, that compaction also
  // writes each listed cell as one word, (weight << 6) | lane, and a second
  // row record of one word per row, rowrec8[row] = position | count << 40,
  // so a pull reads one array per delta run and 8 bytes per row.
  u32* dpack = nullptr;
  u64* rowrec8 = nullptr;
  // When set with rowrec, the counting pass of that compaction also sums,
  // over the touched cells, pc_csr[row+1] - pc_csr[row] (row < pc_nv): the
  // output count of the push join of the new delta against that CSR.
  const u64* pc_csr = nullptr;
  u64 pc_nv = 0;
  // Groups of 2^grp_shift rows of 64 cells (rlbits == 6, groups of at least
  // 256 rows): with grp_mask set, a touched compaction from cell 0 also ORs
  // each group's touched lanes into grp_mask[g], adds its push cost (per
  // touched cell, the degree in pc_csr of its row's index within the group)
  // to grp_cost[g] (both arrays zeroed by the caller), and writes the list
  // position of the group's first cell to grp_start[g].
  u64* grp_mask = nullptr;
  u64* grp_cost = nullptr;
  u64* grp_start = nullptr;
  int grp_shift = 0;
  // Bool relations with a word-parallel pull rule: a touched compaction
  // (from cell 0) also copies every touched word into `dbits` (the delta as
  // a bitmap) and stamps each row of 2^(5 + dwshift) cells that holds a
  // delta cell with `stamp` in `rstamp` (rows are 1 << dwshift words).
  u32* dbits = nullptr;
  u32* rstamp = nullptr;
  u32 stamp = 0;
  int dwshift = 0;
  // Device-driven rounds (cells_loop_round): with `pact` (batched runs) the
  // round's compaction also marks the program of every row holding a delta
  // cell: program rank = row >> pact_shift, id = pact_list[rank] (or rank).
  u8* pact = nullptr;
  int pact_shift = 0;
  const u32* pact_list = nullptr;
  // A touched compaction's write pass (k_cells_write) marks in `pact` the
  // program of every word holding a delta cell when pact_wshift >= 0:
  // program rank = word index >> pact_wshift (a word never spans programs).
  int pact_wshift = -1;
  // Device-driven Trop rounds: the compaction kernels do nothing once the
  // rounds are done or left.
  const LoopState* gate = nullptr;
  u32* bad = nullptr;      // set when a value falls outside its column domain
};

__device__ __forceinline__ u64 cell_enc(i64 w) { return (u64)w ^ (1ull << 63); }
__device__ __forceinline__ i64 cell_dec(u64 e) { return (i64)(e ^ (1ull << 63)); }

// Rank of `code` in column c's domain; `pid` is the row's column-0 code
// (read only for a column ranked within its program).
__device__ __forceinline__ bool cell_rank(const CellDesc& cd, int c, u32 code, u32 pid, u64* r) {
  if (code >= cd.ncodes) return false;
  if (cd.pdom_off[c]) {
    if (pid >= cd.nparts) return false;
    const u64 a0 = cd.pdom_off[c][pid];
    u64 a = a0, b = cd.pdom_off[c][pid + 1];
    const u32* pc = cd.pdom_codes[c];
    while (a < b) {
      const u64 m = (a + b) >> 1;
      if (pc[m] < code) a = m + 1;
      else b = m;
    }
    if (a == cd.pdom_off[c][pid + 1] || pc[a] != code) return false;
    *r = a - a0;
    return true;
  }
  const u32* b = cd.dom_bits[c];
  if (!b) {
    *r = code;
    return true;
  }
  u32 w = b[code >> 5], m = 1u << (code & 31);
  if (!(w & m)) return false;
  *r = cd.dom_pre[c][code >> 5] + (u64)__popc(w & (m - 1));
  return true;
}

__device__ __forceinline__ void cell_touch(const CellDesc& cd, u64 c) {
  atomicOr(&cd.touched[c >> 5], 1u << (c & 31));
}

// ⊕-application of one derivation to cell c, split so that callers can issue
// many cell reads before acting on any: `cell_peek` is a plain read (the Trop
// code, or the Bool presence word), `cell_apply` decides with it.  Plain
// reads may be stale within the kernel, but cell states only decrease, so a
// stale value can only send a derivation to the atomic, which decides.
__device__ __forceinline__ u64 cell_peek(const CellDesc& cd, u64 c) {
  if (!cd.trop) return (u64)cd.present[c >> 5];
  return cd.narrow ? (u64)cd.cell32[c] : cd.cell[c];
}

// Raw Trop cell code -> (absent?, weight); maxf / infb decide the absent case.
__device__ __forceinline__ bool cell_code_absent(const CellDesc& cd, u64 e) {
  return cd.narrow ? e == (u64)kCellTop32 : e == kCellTop;
}
__device__ __forceinline__ i64 cell_code_value(const CellDesc& cd, u64 e) {
  return cd.narrow ? (i64)e : cell_dec(e);
}

template <bool TROP>
__device__ __forceinline__ void cell_apply_t(const CellDesc& cd, u64 c, i64 w, u8 inf, u64 seen) {
  const u32 m = 1u << (c & 31);
  if constexpr (!TROP) {
    if ((u32)seen & m) return;
    if (!(atomicOr(&cd.present[c >> 5], m) & m)) cell_touch(cd, c);
    return;
  }
  if (cd.narrow) {
    if (!inf) {
      if ((u64)w >= (u64)cd.narrow_lim) {  // negative, or too large for a narrow cell
        atomicOr(cd.widen, 1u);
        return;
      }
      const u32 e = (u32)w;
      if ((u64)e >= seen) return;
      if (e < atomicMin(&cd.cell32[c], e)) cell_touch(cd, c);
      return;
    }
    if (seen != (u64)kCellTop32) return;
    if (!(atomicOr(&cd.infb[c >> 5], m) & m)) cell_touch(cd, c);
    return;
  }
  unsigned long long* p = (unsigned long long*)&cd.cell[c];
  if (!inf) {
    const u64 e = cell_enc(w);
    if (e != kCellTop) {
      if (e >= seen) return;
      if (e < (u64)atomicMin(p, (unsigned long long)e)) cell_touch(cd, c);
      return;
    }
    // finite INT64_MAX: better only than absent or +inf
    if (seen != kCellTop) return;
    if (!(atomicOr(&cd.maxf[c >> 5], m) & m)) cell_touch(cd, c);
    return;
  }
  // +inf: better only than absent
  if (seen != kCellTop) return;
  if (cd.maxf[c >> 5] & m) return;
  if (!(atomicOr(&cd.infb[c >> 5], m) & m)) cell_touch(cd, c);
}

__device__ __forceinline__ void cell_insert(const CellDesc& cd, u64 c, i64 w, u8 inf) {
  if (cd.trop) cell_apply_t<true>(cd, c, w, inf, cell_peek(cd, c));
  else cell_apply_t<false>(cd, c, w, inf, cell_peek(cd, c));
}

// Cell index of a tuple given its column codes; false (and `bad` set) when a
// code lies outside its column's domain.
template <class GetCode>
__device__ __forceinline__ bool cell_of(const CellDesc& cd, GetCode code_of, u64* cell) {
  u64 c = 0;
  const u32 pid = cd.nparts ? code_of(0) : 0u;
  for (int j = 0; j < cd.ncols; j++) {
    u64 r;
    if (!cell_rank(cd, j, code_of(j), pid, &r) || (r >> cd.ebits[j])) {
      atomicOr(cd.bad, 1u);
      return false;
    }
    c |= r << cd.shift[j];
  }
  *cell = c;
  return true;
}

// ----------------------------------------------------------- host wrappers

// Bitmap of the codes appearing in column j of a packed relation.
template <class K>
void cells_mark_col(const Exec& x, const K* keys, u64 n, int arity, int bits, int j, u32* bitmap);
void cells_mark_code(const Exec& x, u32* bitmap, u32 code);
void cells_or(const Exec& x, u32* dst, const u32* src, u64 nwords);
// Domain tables of a bitmap: per-word prefix counts and the rank -> code list.
// Returns the number of codes in the domain.
u64 cells_domain(const Exec& x, const u32* bitmap, u64 nwords, u64* pre, DBuf<u32>* list);
// Per-program domain of a column in a batched run: the distinct (program,
// code) pairs of the given relation columns (column 0 of each relation is the
// program id, whose code is below nparts).  off (nparts + 1) and codes are
// (re)allocated; returns the largest per-program count.
struct PartCol {
  const void* keys;
  bool wide;
  u64 n;
  int arity, bits, col;
};
u64 cells_part_domain(const Exec& x, const std::vector<PartCol>& parts, u64 nparts, DBuf<u64>* off, DBuf<u32>* codes);
// Load a relation's tuples into empty cells (no touched bits).
template <class K>
void cells_load(const Exec& x, const CellDesc& cd, const K* keys, const i64* w, const u8* inf, u64 n, int arity,
                int bits);
// ⊕-apply packed derivations (marks touched cells).
template <class K>
void cells_insert(const Exec& x, const CellDesc& cd, const K* keys, const i64* w, const u8* inf, u64 n, int arity,
                  int bits);
// Compact the touched cells (clearing them) or all present cells of
// [cell0, cell0 + ncells) into a packed relation in cell order; returns its
// size.  cell0 is a multiple of 32.  The outputs are (re)allocated.  With
// `push_cost` given and cd.pc_csr set, *push_cost receives the sum described
// at CellDesc::pc_csr, or ~0 when it was not computed.
template <class K>
u64 cells_compact(const Exec& x, const CellDesc& cd, u64 ncells, bool touched_only, int bits, DBuf<K>* keys,
                  DBuf<i64>* w, DBuf<u8>* inf, bool want_inf, u64 cell0 = 0, u64* push_cost = nullptr);

// Touched compaction (from cell 0) into caller buffers whose capacity is at
// least the number of touched cells (ncells bounds it), with no host read
// between the passes: one read at the end returns the count, the push cost
// (as cells_compact, ~0 when not computed) and, when `status` is given, a
// snapshot of that device record taken after every earlier launch.  Needs
// ncells / (kBlock * 32) <= kScanOneMax blocks (cells_into_fits).
bool cells_into_fits(u64 ncells);
// One device-driven round's compaction (see LoopState): the touched cells
// become the delta bitmap and row stamps (no delta list), then the round's
// bookkeeping — iters += 1, stamp += 1, done when no cell was touched;
// batched runs (nparts > 0) count the round for each program whose previous
// delta was non-empty (active_prev) and move active_next (marked by the
// compaction) into active_prev.  Two launches, no host read; capturable in
// a graph.
void cells_loop_round(const Exec& x, const CellDesc& cd, u64 ncells, LoopState* ls, u8* active_prev,
                      u8* active_next, u64* piters, u32 nparts);
// A device-driven Trop round's compaction (cd.gate set): the touched
// compaction into the capacity buffers with the push cost and a snapshot of
// the device status (as cells_compact_into, without the host read), then
// the round's bookkeeping: a status sets leave = 2; otherwise iters += 1,
// n and pc recorded, leave = 3 past cap_left rounds, done on an empty delta;
// when pc < push_limit the next round pushes on the device (stamp = 1) if
// the delta has at most push_cap rows, else leave = 1; otherwise it pulls
// (stamp = 0).  cnt / off / pcb hold ncells / (kBlock * 32) entries, tot 8.
template <class K>
void cells_loop_round_trop(const Exec& x, const CellDesc& cd, u64 ncells, int bits, K* keys, i64* w, u8* inf,
                           u32* cnt, u64* off, u64* pcb, u64* tot, const DevErr* status, LoopState* ls,
                           u64 cap_left, u64 push_limit, u64 push_cap);
template <class K>
u64 cells_compact_into(const Exec& x, const CellDesc& cd, u64 ncells, int bits, K* keys, i64* w, u8* inf,
                       u64* push_cost, const DevErr* status, DevErr* status_out);
// A device-driven batched Trop round's compaction (cd.gate, cd.pact =
// active_next, cd.pact_wshift set): the touched compaction into the
// capacity buffers with a snapshot of the device status, marking each
// program holding a delta cell.  A status (widen, overflow) or any
// program's overflow record leaves (2) before the write pass, so the
// touched cells and the previous delta stay for the host to finish the
// round; otherwise the round's per-program bookkeeping follows: every
// program active in the previous round gets one round (piters[p] += 1),
// leaving (3) once some piters[p] > capleft[p], active_prev takes
// active_next (cleared), iters += 1, ls->n = the new delta's size, and done
// on an empty delta.  cnt / off hold ncells / (kBlock * 32) entries, tot 8.
template <class K>
void cells_loop_round_trop_parts(const Exec& x, const CellDesc& cd, u64 ncells, int bits, K* keys, i64* w, u32* cnt,
                                 u64* off, u64* tot, const DevErr* status, LoopState* ls, u8* active_prev,
                                 u8* active_next, u64* piters, const u64* capleft, const DevErr* part_derr,
                                 u32 nparts);
// Touched compaction (from cell 0) writing row records (cd.rowrec8 and/or
// cd.rowrec, rows of 2^rlbits <= 64 cells) and, per listed cell, the packed
// word (*dpack; narrow cells) or the weight and the lane (*w, *lane), no keys.
// Outputs are (re)allocated to the returned count.
u64 cells_compact_rows(const Exec& x, CellDesc cd, u64 ncells, DBuf<u32>* dpack, DBuf<i64>* w, DBuf<u8>* lane);
// Mark the cells of n packed keys touched (the keys' cells are in the layout).
template <class K>
void cells_mark_touched(const Exec& x, const CellDesc& cd, const K* keys, u64 n, int arity, int bits);

// State of arbitrary cells: present[i] and the weight of cell idx[i]
// (+inf reads as present with weight 0; check infb where it matters).
void cells_gather(const Exec& x, const CellDesc& cd, const u64* idx, u64 n, i64* w, u8* present);
// Narrow cell codes -> wide codes (same states).
void cells_widen(const Exec& x, const u32* c32, u64* c64, u64 n);

}  // namespace sk
