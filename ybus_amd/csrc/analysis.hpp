// Host-side symbolic analysis of an uploaded factor: each triangular sweep read as rows to pull
// (every row subtracting its dependencies in the order ybus_sparse's CPU solve does), scheduled
// into level sets, with the solve's gather and scatter permutations folded into the rows a sweep
// reads and writes; the fundamental supernodes of a Cholesky factor's elimination tree; and the
// tile-width choice for a block of right-hand sides.
#pragma once

#include <climits>
#include <cstdint>
#include <string>
#include <vector>

namespace ybus_hip {

// A triangular matrix as rows to pull: row i is (b_i - sum over its entries of val * x[dep]),
// the entries in the order they are subtracted, then divided by diag[i] (no division when
// `unit`). `lower` rows depend only on earlier rows, upper rows only on later ones.
struct RowForm {
  int64_t n = 0;
  bool lower = true;
  bool unit = false;
  std::vector<int64_t> ptr;  // n + 1
  std::vector<int32_t> dep;
  std::vector<double> val;
  std::vector<double> diag;  // n, unless unit
  // imaginary parts of a complex factor's val and diag (empty for a real factor)
  std::vector<double> val_im, diag_im;
};

// A sweep scheduled for the device: rows in level order (position k), each reading its own
// right-hand side at src[k] of the source block, writing its result at dst[k] of the
// destination block, and reading dependencies at dep[p] of the destination block.
struct Sweep {
  int32_t n = 0;
  int32_t levels = 0;
  bool unit = false;
  std::vector<int32_t> level_ptr;  // levels + 1
  std::vector<int32_t> src, dst;   // n each
  std::vector<int32_t> row_ptr;    // n + 1
  std::vector<int32_t> dep;
  std::vector<double> val;
  std::vector<double> diag;  // n, unless unit
  std::vector<double> val_im, diag_im;  // a complex factor's imaginary parts, else empty
  int32_t widest_level = 0;
  // A partitioned sweep (parts > 0): its rows are grouped into `parts` independent parts and a
  // top part (index parts), in that order, each leveled on its own dependencies; part q's levels
  // are level_ptr[part_level[q]] .. level_ptr[part_level[q + 1]] (parts + 2 entries).
  int32_t parts = 0;
  std::vector<int32_t> part_level;
  // A split sweep (split_rows): a position is a segment of a row's terms rather than a whole
  // row, so there may be more positions than rows; a segment that continues a row reads its
  // right-hand side from the row's own slot of the destination block (src = ~dst).
  bool split = false;
  int32_t positions() const { return level_ptr.empty() ? 0 : level_ptr.back(); }
};

// Schedules `form`: level[i] is one more than the deepest of row i's dependencies; rows are
// ordered by (level, row). src_map / dst_map give each factor row's row in the source and
// destination blocks (dependencies are mapped through dst_map).
Sweep schedule(const RowForm& form, const std::vector<int64_t>& src_map, const std::vector<int64_t>& dst_map);

// Splits the rows of a lower triangular form into independent parts and a top part, when its
// dependency graph is a tree in this sense: every row's dependents other than the first are also
// dependents of the first (a Cholesky factor: the elimination tree). Subtrees are cut from the
// root down until none holds more than n / (2 target) rows, the cut rows forming the top part, and
// the subtrees are packed into at most `target` parts, largest first into the lightest. Returns
// each row's part (parts for the top part), or nothing when the graph is not such a tree or
// there would be fewer than two parts.
std::vector<int32_t> tree_parts(const RowForm& lower, int32_t target, int32_t& parts);

// schedule() with the rows grouped by part_of (0 .. parts - 1, then the top part `parts`) and
// each part leveled on its dependencies inside it. top_first: the top part is swept before the
// others (a part's rows may depend on its own rows and the top part's; the top part's only on
// its own); otherwise after them (a part's rows only on its own; the top part's on any). A
// partition the form does not satisfy gives a sweep with parts = 0 and no rows.
Sweep schedule_parts(const RowForm& form, const std::vector<int64_t>& src_map, const std::vector<int64_t>& dst_map,
                     const std::vector<int32_t>& part_of, int32_t parts, bool top_first);

// A scheduled sweep with every row of more than `cap` terms cut into chained segments, so that
// a long row's terms no longer wait, one chunk after another, at the row's own level. The terms
// keep their order: segment j of a row takes a run of its terms after segment j - 1's, starting
// from the row's right-hand side (the first segment) or from the partial value the previous one
// left in the row's own slot (src = ~dst), and only the last divides by the diagonal (the others
// carry continuation_diag: 1, an exact division, in a solve's sweep). The segments are placed
// from the last back: the last at the row's own level, each earlier one a level before the next,
// each of at most `cap` terms unless a term before it would then wait for a dependency that is
// not final by the level before (the segment then reaches back to that term), so no level and
// no dependent moves. The rows'
// bits do not change: every term is subtracted in its row's order, each result rounded as
// before.
Sweep split_rows(const Sweep& s, int32_t cap, double continuation_diag = 1.0);

// The top parts of a partitioned forward sweep and of its backward sweep (top first, so the
// backward top depends on top rows only), renumbered for a solve that holds them in LDS: each
// top row takes a slot, in the order its first forward segment comes in, and each row of the
// other parts that a forward top segment depends on takes a halo slot after them. Each segment
// then names its row's slot and each term its dependency's slot. A first forward segment starts
// from its row's right-hand side (rhs_row), a first backward one from its row's forward value
// (the same slot), and a continuing segment from the partial value in its slot, so a segment
// needs no source of its own. rows = 0 when the sweeps do not have this form.
struct TopSlots {
  struct Part {
    int32_t level0 = 0, levels = 0, seg0 = 0, segs = 0, term0 = 0, terms = 0;  // in the sweep's arrays
    std::vector<int32_t> slot;  // segs
    std::vector<int32_t> dep;   // terms
  };
  int32_t rows = 0, halo = 0;
  std::vector<int32_t> rhs_row;   // rows: each top row's forward right-hand side row
  std::vector<int32_t> halo_row;  // halo: the forward block's row each halo slot holds
  std::vector<int32_t> out_row;   // rows: each top row's row of the backward (solution) block
  Part fwd, bwd;
};
TopSlots top_slots(const Sweep& fwd, const Sweep& bwd);

// The independent parts of a partitioned forward sweep and of its backward sweep, each part
// renumbered for a solve that holds its values in LDS: the part's rows take slots in the order
// their first segments come in, and (backward only) each top row a part's terms read takes a
// halo slot after them. A first segment starts from its row's source row (in_row: the forward's
// right-hand side row, the backward's forward value), a continuing one from the partial value in
// its slot; each row's result goes to out_row. Per part q: segments seg_begin[q] ..
// seg_begin[q + 1] of the sweep (its own segment order), slots slot_begin[q] .. (in_row,
// out_row), halo slots halo_begin[q] .. (halo_row: the backward block's row each holds).
// parts = 0 when the sweeps do not have this form.
struct PartSlots {
  struct Dir {
    std::vector<int32_t> seg_begin, slot_begin, halo_begin;  // parts + 1 each
    std::vector<int32_t> slot;                               // per segment: its row's slot in its part
    std::vector<int32_t> dep;                                // per term of those segments: its slot
    std::vector<int32_t> in_row, out_row, halo_row;
  };
  int32_t parts = 0;
  Dir fwd, bwd;
};
PartSlots part_slots(const Sweep& fwd, const Sweep& bwd);

// A sweep's sources encoded for a solve that keeps its last `cap` positions' values in an LDS ring
// (position q in slot q mod cap) and loads other values WINDOW_AHEAD levels before their row's
// level. For a position at level l, a term's row d (written last at position q) is:
//   ~(q mod cap) (negative), when q is still in the ring as level l runs (q >= the level's end
//     - cap);
//   d, when q is at level l - WINDOW_AHEAD - 1 or earlier, or d is never written by the sweep
//     (the pad row): its value is final by the start of level l - WINDOW_AHEAD, so it can be
//     loaded then;
//   d + n + 1, otherwise: loaded when the row is swept.
// src: a first segment's right-hand side row; a continuing segment's own row by the same rule,
// as ~(q mod cap), or WINDOW_PARTIAL (its partial value read from the block at dst).
constexpr int32_t WINDOW_PARTIAL = INT32_MIN;
constexpr int32_t WINDOW_AHEAD = 2;
struct WindowCode {
  int32_t cap = 0;
  std::vector<int32_t> src;  // per position
  std::vector<int32_t> dep;  // per term
};
WindowCode window_code(const Sweep& s, int32_t cap);

// Each row's level: one more than the deepest of its dependencies (0 for none).
std::vector<int32_t> levels_of(const RowForm& form);

// The four ways a factor is read, built from its host arrays.
// Cholesky (cholesky.rs:468-501): L in CSR (each row ends with its diagonal) forward, and L^T
// (L in CSC) backward, both ascending.
RowForm cholesky_forward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val);
RowForm cholesky_backward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val);
// LU (lu.rs:640-649, 770-785): the CPU sweeps L and U by columns; pulled by rows, row i of L
// takes its entries in ascending column order and row i of U in descending column order, which
// is the order the column sweeps update y[i] in.
RowForm lu_forward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val);
RowForm lu_backward(int64_t n, const uint64_t* u_ptr, const uint64_t* u_idx, const double* u_val);
// LU transposed (lu.rs:506-533): U^T forward, each column of U read as a row (diagonal last),
// then unit L^T backward, each column of L below its diagonal read as a row.
RowForm lu_transpose_forward(int64_t n, const uint64_t* u_ptr, const uint64_t* u_idx, const double* u_val);
RowForm lu_transpose_backward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val);

// A complex factor's row form from two built with the same builder over the same pattern: `re`
// from the entries' real parts, `im` from their imaginary parts. `conjugate` negates the
// imaginary parts of the entries off the diagonal and leaves the diagonal's: the reference's
// L^H (cholesky.rs:283-297) conjugates L's columns below the diagonal and keeps each (real)
// diagonal as it is.
RowForm complex_form(RowForm re, const RowForm& im, bool conjugate);

// Fundamental supernodes of a Cholesky factor (multifrontal.rs:94-138): the elimination tree
// comes from L's pattern (the parent of column j is the first row below j in column j).
struct Supernodes {
  std::vector<int64_t> start;  // one per supernode, plus n at the end
  int64_t largest = 0;
  int64_t in_chains = 0;  // columns in supernodes of two or more columns
};
Supernodes cholesky_supernodes(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx);

// Validates a CSR/CSC pattern: pointers non-decreasing from 0, indices below `minor`.
bool valid_compressed(int64_t major, int64_t minor, const uint64_t* ptr, const uint64_t* idx, std::string* why);

// The RHS tile width (columns one workgroup owns) for `ncols` columns on `cus` compute units: the
// widest power of two up to 64 that still gives nine tiles for every ten compute units, else 1.
// (A tile's cost is set by its levels, not its width, so one round of wide tiles beats two
// rounds of narrow ones.)
int tile_width_for(int64_t ncols, int cus);

}  // namespace ybus_hip
