// hostalg.h -- host-side graph algorithms of the pressure-solve path:
// the reference's greedy agglomeration, multicolour orderings and the level
// schedules of the lexicographic smoothers.  Integer/sequential work that
// runs once per hierarchy; its outputs are uploaded to HBM.
#pragma once
#include <cstdint>
#include <vector>

namespace ufvm {

struct HostCSR {
  int32_t n = 0;
  std::vector<int32_t> ptr, col;  // reference (cconn) order
};

struct AggResult {
  int32_t nc = 0;
  std::vector<int32_t> parents;     // fine -> coarse (0-based)
  HostCSR coarse;                   // coarse cconn, first-encounter order
  std::vector<int32_t> child_ptr;   // coarse row -> children, increasing fine index
  std::vector<int32_t> child;
  std::vector<int32_t> entry_slot;  // fine nnz -> coarse nnz slot, -1 if same parent
  std::vector<int32_t> gal_ptr, gal_src;  // galerkin_lists of this level, when built
};

// cfdAgglomerateLevel.m:3-104 on one level (anb values in cconn order).
void agglomerate_level(const HostCSR &g, const double *anb, AggResult &out);

// The Galerkin coarse operator as gather lists, one per coarse value in
// the reference's accumulation order (assembleAgglomeratedLHS.m:16-30: fine
// rows by parent, children in increasing order, entries in cconn order).
// Value o < nc is AC[o], value nc + t is ANB slot t.  Sources: e >= 0 is
// fine anb[e], e < 0 is fine ac[-1 - e].
void galerkin_lists(const HostCSR &g, const AggResult &ag, std::vector<int32_t> &ptr,
                    std::vector<int32_t> &src);

// greedy first-fit colouring in index order; returns number of colours
int32_t greedy_colour(const HostCSR &g, std::vector<int32_t> &colour);
// the same restricted to each block of `rows_per_block` consecutive rows
// (only neighbours inside the row's own block constrain its colour)
int32_t greedy_colour_blocks(const HostCSR &g, int32_t rows_per_block,
                             std::vector<int32_t> &colour);

// Level schedule of a lexicographic sweep.  forward: row i waits for its
// neighbours j < i; backward: for neighbours j > i.  rows grouped by level,
// increasing row index inside a level.
struct Schedule {
  std::vector<int32_t> level_ptr;  // [nlevels+1]
  std::vector<int32_t> rows;       // [n]
};
void lex_schedule(const HostCSR &g, bool forward, Schedule &s);

// For DILU: for every row j, its lower neighbours i < j sorted by i, with the
// position of i in row j (a_ji) and of j in row i (a_ij).
struct LowerLists {
  std::vector<int32_t> ptr;       // [n+1]
  std::vector<int32_t> nb;        // i
  std::vector<int32_t> pos_ji;    // CSR slot of a_ji in row j
  std::vector<int32_t> pos_ij;    // CSR slot of a_ij in row i
};
bool lower_lists(const HostCSR &g, LowerLists &L);

// For ILU(0) in IKJ form over a numbering whose lower neighbours are the
// smaller indices: for every lower record o of row j (LowerLists order, k =
// nb[o]), the entries (k, c) of row k with c > k, in entry order, that land in
// row j -- on the diagonal (c == j) or on the first entry (j, c) of row j.
// Entries (k, c) with no such target drop out (ILU(0) keeps no fill).
struct Ilu0Lists {
  std::vector<int32_t> ptr;       // [lower records + 1]
  std::vector<int32_t> src;       // CSR slot of (k, c) in row k
  std::vector<int32_t> dst;       // CSR slot of (j, c) in row j, -1: the diagonal
};
void ilu0_lists(const HostCSR &g, const LowerLists &L, Ilu0Lists &U);

// 16-bit form of a sliced-ELL column array of n rows (64-row chunks, chunk q's
// slots at [soff[q], soff[q+1]) column-major; row t has ptr[t+1] - ptr[t]
// real entries, the rest of its slots are padding).  On success base[q] +
// d16[slot] is scol[slot] for every real entry, padding slots take the row's
// last real column (the row itself when it has none), and slots of rows past
// n take base[q].  False, leaving d16 and base unspecified, when a chunk's
// columns span 65536 or more.
bool pack_sell_columns(const std::vector<int32_t> &ptr, const std::vector<int32_t> &soff,
                       const std::vector<int32_t> &scol, int32_t n, std::vector<uint16_t> &d16,
                       std::vector<int32_t> &base);

}  // namespace ufvm
