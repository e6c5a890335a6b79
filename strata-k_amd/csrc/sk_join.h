// sk_join.h — descriptors and launch wrappers of the join-side kernels.
#pragma once

#include <vector>

#include "sk_cells.h"
#include "sk_dev.h"

namespace sk {

// Binding table: one u32 column per live variable, plus the Trop
// accumulator (w, inf) when the rule's head is Trop.
struct Table {
  u64 n = 0;
  int ncols = 0;
  u32* col[kMaxCols] = {};
  i64* w = nullptr;
  u8* inf = nullptr;
  int srccol[kMaxCols] = {};  // compact_table: output column c copies input column srccol[c]
  // batched runs: the column holding the program id and the per-program
  // error records (an overflow is charged to that program only)
  int part_col = -1;
  DevErr* perr = nullptr;
};

using TableAlloc = void (*)(void* ctx, Table* t, u64 n);

// Scan of a relation into a table.
struct ScanDesc {
  int arity = 0, bits = 0;
  int nconst = 0;
  int const_col[kMaxCols] = {};
  u32 const_val[kMaxCols] = {};
  int neq = 0;
  int eq_a[kMaxCols] = {}, eq_b[kMaxCols] = {};
  int out_col[kMaxCols] = {};  // table column c = relation column out_col[c]
};

// Probe of a relation index whose key order puts the bound columns first.
struct ProbeDesc {
  int arity = 0, bits = 0;
  u64 sn = 0;
  int nprobe = 0;
  int psrc[kMaxCols] = {};  // prefix column j: table column psrc[j], or -1 = constant pconst[j]
  u32 pconst[kMaxCols] = {};
  int nfilt = 0;  // index columns that must be equal (repeated unbound variable)
  int fa[kMaxCols] = {}, fb[kMaxCols] = {};
};

// Output of a join / projection: either the next table's columns or a
// packed head key.  kind 0 = table column idx, 1 = probed relation column
// idx (index order), 2 = constant cnst.
struct OutDesc {
  int ncols = 0;
  int kind[kMaxCols] = {};
  int idx[kMaxCols] = {};
  u32 cnst[kMaxCols] = {};
  u32* col[kMaxCols] = {};
  void* keys = nullptr;
  int bits = 0;
  i64* w = nullptr;
  u8* inf = nullptr;
  Table* table = nullptr;  // host side: destination table of a non-final join
  CellDesc cd;             // cd.on: head rows are ⊕-applied to the head's cells, not written
};

using HeadAlloc = void (*)(void* ctx, OutDesc* o, u64 n);

// Pull evaluation of a linear recursive rule over dense cells (see
// k_pull): head H(r, x) ⊕= Δ(j, x) ⊗ E(j → r), rows r of the head taken from
// E's index on its r column, lanes x passed through from the delta.
// One wave's share of a pull round: index rows [lo, lo + n) of E's R-index,
// which belong to nr <= kPullRows consecutive non-empty destination rows;
// destination row s owns index rows from lo + off(s) (off(0) = 0).  A row
// with more than kPullPack index rows is split into chunks of at most 256,
// each with nr = 1.
constexpr u32 kPullRows = 4;
constexpr u32 kPullPack = 64;
struct PullChunk {
  u64 lo;
  u32 n;
  u32 nr_off;  // bits 0-7: nr; bits 8s..8s+7: off(s) for s = 1..3
  u32 row[kPullRows];
};
static_assert(sizeof(PullChunk) == 32, "PullChunk is two 16-byte loads");

struct PullDesc {
  u64 nchunks = 0;
  const PullChunk* chunks = nullptr;
  const u32* adj = nullptr;        // per index row: the j value (code = rank)
  const i64* w = nullptr;          // per index row: the weight (Trop)
  const u64* rowrec = nullptr;     // per row: {last round's delta lanes, position of the row's first
                                   //           delta cell in the delta list}
  const i64* dw = nullptr;         // delta weights, in cell order
  const u8* dlane = nullptr;       // delta cells' lanes, in cell order
  const u32* dpack = nullptr;      // packed narrow delta ((weight << 6) | lane), in cell order, or null
  const u64* rowrec8 = nullptr;    // with dpack: per row, position | count << 40
  int lbits = 0;                   // lanes per row = 2^lbits
  int delta_first = 1;             // ⊗ operand order: the delta's weight on the left
  u64 nrows = 0;                   // rows of the cell layout (j values at or above are absent)
  const LoopState* gate = nullptr; // device-driven rounds: nothing to do once done or left
};

// Word-parallel pull of a Bool linear recursive rule over dense cells:
// head row d ⊕= OR of the delta rows s with (d, s) in the rule's row graph
// (one pair per tuple of the fixed relation E: d = the head's row, s = the
// delta's row, both cell-row indices).  Rows are 2^wshift words of 32
// lanes; the lane column passes through unchanged.  One wave takes a chunk:
// up to 256 graph entries of one destination row.
struct BitChunk {
  u64 lo;
  u32 n;
  u32 row;
};
static_assert(sizeof(BitChunk) == 16, "BitChunk is one 16-byte load");

struct BitPullDesc {
  u64 nchunks = 0;
  const BitChunk* chunks = nullptr;
  const u64* graph = nullptr;  // (d << 32 | s), sorted by d
  const u32* dbits = nullptr;  // last round's delta as a bitmap
  const u32* rstamp = nullptr; // per row: stamp of the last compaction that found a delta cell in it
  u32 stamp = 0;
  int wshift = 0;
  const LoopState* loop = nullptr;  // device-driven rounds: stamp and done flag from here
};

// Row graph of a bit-pull rule from E's packed tuples: for every tuple, the
// head's and the delta's cell rows (column c of H takes the value of E
// column eh[c] / ed[c]; -1 marks the lane column).  Invalid pairs (a value
// outside H's domain) sort last as (nrows << 32).  Returns the valid count;
// `graph` is sorted by destination row and `csr` (nrows + 1) indexes it.
template <class K>
u64 bitpull_graph(const Exec& x, const K* ekeys, u64 n, int arity, int bits, const CellDesc& cd, int hcols,
                  const int* eh, const int* ed, int lbits, u64 nrows, DBuf<u64>* graph, DBuf<u64>* csr);
u64 bitpull_chunks(const Exec& x, const u64* csr, u64 nrows, DBuf<BitChunk>* out);
void bitpull_apply(const Exec& x, const BitPullDesc& p, const CellDesc& cd, double bytes, u64 items);

struct AntiDesc {
  int arity = 0, bits = 0;
  int src[kMaxCols] = {};  // table column, or -1 = constant
  u32 cnst[kMaxCols] = {};
  int never_present = 0;   // arity mismatch: the negation always holds
};

struct AggDesc {
  int nkey = 0;
  int key_kind[kMaxCols] = {};  // 0 = table column, 2 = constant
  int key_idx[kMaxCols] = {};
  u32 key_cnst[kMaxCols] = {};
  int bits = 0;
  int op = 0;       // 0 count, 1 sum, 2 min, 3 max
  int val_col = -1;
  u32 first_int = 0, n_ints = 0;
};

struct TriDesc {
  int bits = 0;
  int u_col = 0;  // driver column holding u (the other holds v)
};

template <class K>
u64 scan_rel(const Exec& x, const K* keys, const i64* w, const u8* f, u64 n, const ScanDesc& d, Table* out,
             TableAlloc alloc, void* actx);

template <class K>
u64 join(const Exec& x, const Table& in, const ProbeDesc& p, const K* skeys, const i64* sw, const u8* sf,
         const u64* csr, OutDesc* o, bool final_step, bool k128_out, bool trop, TableAlloc alloc, void* actx,
         HeadAlloc halloc, void* hctx);

template <class K>
u64 anti_join(const Exec& x, const Table& in, const AntiDesc& a, const K* skeys, u64 sn, Table* out,
              TableAlloc alloc, void* actx);

void project(const Exec& x, const Table& in, OutDesc* o, bool k128_out, HeadAlloc halloc, void* hctx);

template <class K>
u64 aggregate(const Exec& x, const Table& in, const AggDesc& a, const i64* int_vals, int key_bits,
              std::vector<K>* keys_out, std::vector<i64>* vals_out, bool* bad_aggregand);

template <class K>
u64 triangles(const Exec& x, const K* d_keys, u64 nd, const TriDesc& t, const u64* csr2, const u32* adj2,
              const u64* csr3, const u32* adj3, OutDesc* o, bool k128_out, bool count_only, HeadAlloc halloc,
              void* hctx, double* sum_list_len);

template <class K>
double tri_list_len(const Exec& x, const K* d_keys, u64 nd, const TriDesc& t, const u64* csr2, const u64* csr3);

template <class K, class KO>
void remap_keys(const Exec& x, const K* in, u64 n, int arity, int bits_in, int bits_out, const u32* map, KO* out);

template <class K>
void pack_cols(const Exec& x, const Cols& cols, int arity, int bits, u64 n, K* out);

void mark_codes(const Exec& x, const u32* c, u64 n, u32 limit, u32* bitmap);

// Keep rows whose column-0 program id is not flagged dead; in place
// (stream-compaction into the given output buffers).  Returns the count.
template <class K>
u64 drop_partitions(const Exec& x, const K* keys, const i64* w, const u8* inf, u64 n, int arity, int bits,
                    const u8* dead, K* ok, i64* ow, u8* of);
u64 popcount_bits(const Exec& x, const u32* bitmap, u64 nwords);

// Chunks of at most `per` index rows per head row, from a CSR over the head
// row column (nrows + 1 offsets).  Returns the chunk count.
// Chunk list of a CSR over nrows destination rows (host-side greedy packing);
// *rows_used (when given) receives the number of rows with an entry.
u64 pull_chunks(const Exec& x, const u64* csr, u64 nrows, DBuf<PullChunk>* out, u64* rows_used = nullptr);
// Σ over the delta rows of the push join's output count (the push cost).
template <class K>
u64 push_cost(const Exec& x, const K* dkeys, u64 n, int arity, int bits, int col, const u64* csr);
void pull_apply(const Exec& x, const PullDesc& p, const CellDesc& cd, bool trop, double bytes, u64 items);

// One device-driven push round (graph capture) of a Trop rule that scans a
// delta and joins it to a fixed relation, the head applied to cells: the
// delta's ls->n rows (packed keys KD, weights) are scanned into table `t`
// (sd.out_col), joined (count, device-sized scan, load-balanced emit) and
// applied to o.cd.  t, lo, cnt and off hold `cap` rows (at least any delta),
// bsum 1024 words, jtot one.  Every kernel returns at once when the rounds
// are done or have left (ls).
constexpr u32 kDevScanBlocks = 1024;
// Rows of the push table of device-driven plain Trop rounds: a delta with
// more rows leaves its push round to the host.
constexpr u64 kDevPushRows = 1ull << 24;
template <class KD, class KE>
void push_round_dev(const Exec& x, const KD* dkeys, const i64* dw, int darity, int dbits, const ScanDesc& sd,
                    const Table& t, const ProbeDesc& p, const KE* skeys, const i64* sw, const u64* csr,
                    const OutDesc& o, u64* lo, u32* cnt, u64* off, u64* bsum, u64* jtot, u64 cap,
                    const LoopState* ls);

}  // namespace sk
