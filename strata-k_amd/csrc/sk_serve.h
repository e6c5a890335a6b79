// sk_serve.h — device side of the query server's lane groups.
//
// The server keeps cost(instance, Z) in groups of 64 instances: a group is
// a dense cell array of 2^zbits rows (one per node Z) of 64 lanes (one per
// instance), cell = Z << 6 | lane, the layout the pull round reads (a
// destination row's 64 lanes are one 256-byte line for narrow cells).  The
// groups lie back to back, so one touched compaction over groups [0, hi)
// gives every instance's next delta with per-row records (rowrec8).
#pragma once

#include "sk_cells.h"
#include "sk_dev.h"

namespace sk {

// Push table of group g's delta entries at positions [p0, p1) of the delta
// list: columns (group, lane, node, instance = g * 64 + lane) and the
// weight, entry p at row p - p0.  Weights come from dpack (narrow) or dw
// (wide).
void srv_table(const Exec& x, const u64* rowrec8, int zbits, u32 g, const u32* dpack, const i64* dw,
               const u8* dlane, u64 p0, u64 p1, u32* tg, u32* tl, u32* ty, u32* ti, i64* tw);

// Answers and lane resets of one group (gcd: the group's cells, a two-column
// view lane | node << 6).  For every lane in take_mask the present cells of
// rows [0, n_nodes) are written in node order to outz[lane] / outw[lane],
// allocated here to counts[lane]; every lane in clear_mask is then reset to
// absent, or to weight 0 (INT64_MIN for wide cells) when `poison` is set: a
// poisoned lane takes no derivation, so a delta already listed for it dies
// out in one round.
void srv_lanes(const Exec& x, const CellDesc& gcd, u64 n_nodes, u64 take_mask, u64 clear_mask, bool poison,
               DBuf<u32>* outz, DBuf<i64>* outw, u64* counts);

// Narrow cells (u32) of [0, n) widened in place to u64 codes: the array must
// have room for n u64 words; the narrow cells are its first 4n bytes.
void srv_widen_in_place(const Exec& x, void* cells, u64 n);

}  // namespace sk
