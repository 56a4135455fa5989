// Launchers of the device kernels (kernels.hip). Every launcher enqueues on `stream` and
// returns the launch's hipError_t; shapes are checked by the caller before any launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ybus_hip {

// A sweep row's work as a record: the row it writes (dst) and reads its own right-hand side from
// (src), its entries p0..p1 in the sweep's arrays, the first RECORD_ENTRIES of them inline
// (dependency and value), and its diagonal (stored as 1 for a unit-diagonal sweep, so a fetch is
// the same loads for every sweep). The records are stored field by field (one array
// of 16-byte items per field, indexed by position), so the lanes of a wavefront that fetch
// consecutive rows' records read consecutive 16-byte items: every fetch is coalesced, and a
// thread can fetch its next rows' records ahead with independent loads.
constexpr int RECORD_ENTRIES = 4;
struct RecordsDev {
  const int4* head = nullptr;      // dst, src, p0, p1
  const int4* dep = nullptr;       // the first four dependencies (unused ones: dst)
  const double2* val01 = nullptr;  // their FP64 values, two at a time (unused ones: 0)
  const double2* val23 = nullptr;
  const float4* val32 = nullptr;   // their FP32 values
  const double* diag64 = nullptr;  // the row's diagonal (1 for a unit diagonal)
  const float* diag32 = nullptr;
};

// One scheduled triangular sweep on the device (see analysis.hpp's Sweep).
struct SweepDev {
  int32_t n = 0, levels = 0;
  const int32_t* level_ptr = nullptr;
  const int32_t* src = nullptr;
  const int32_t* dst = nullptr;
  const int32_t* row_ptr = nullptr;
  const int32_t* dep = nullptr;
  const double* val64 = nullptr;
  const float* val32 = nullptr;
  const double* diag64 = nullptr;  // null: unit diagonal
  const float* diag32 = nullptr;
  RecordsDev rec;  // one record per position, in processing order
  bool unit = false;
  // positions are segments of rows (a split sweep): a segment with src < 0 continues row ~src
  // from the partial value the previous segment left in its slot of the destination block
  bool split = false;
  // a partitioned sweep: `parts` independent parts, then the top part (index parts); part q's
  // levels are part_level[q] .. part_level[q + 1] of level_ptr (parts + 2 entries)
  int32_t parts = 0;
  const int32_t* part_level = nullptr;
};

// One scheduled sweep of a complex factor: the arrays of SweepDev, each value a (re, im) pair.
struct ComplexSweepDev {
  int32_t n = 0, levels = 0;
  const int32_t* level_ptr = nullptr;
  const int4* head = nullptr;  // per position: dst, src, first term, end of terms
  const int32_t* src = nullptr;
  const int32_t* dst = nullptr;
  const int32_t* row_ptr = nullptr;
  const int32_t* dep = nullptr;
  const double2* val = nullptr;
  const double2* diag = nullptr;  // null: unit diagonal
  const float2* val32 = nullptr;   // the values rounded to FP32 (a refined solve's sweeps)
  const float2* diag32 = nullptr;
};

// A complex matrix in CSR form on the device ((re, im) values), for a refined solve's residual
struct ZCsrDev {
  int32_t rows = 0;
  const int32_t* ptr = nullptr;
  const int32_t* col = nullptr;
  const double2* val = nullptr;
};

// A CSR matrix on the device.
// The FP64 residual loads a row's entries this many at a time from any position up to nnz, so a
// matrix uploaded for it carries this many padding entries past its last (column 0, value 0)
constexpr int RESIDUAL_CHUNK = 8;

struct CsrDev {
  int32_t rows = 0, cols = 0;
  const int32_t* ptr = nullptr;
  const int32_t* col = nullptr;
  const double* val = nullptr;
};

// Right-hand sides of a forward sweep: a block (rows x ld, FP64), or pairs e_plus - e_minus
// built on the fly (plus/minus per column, -1 for none) when `block` is null.
struct RhsDev {
  const double* block = nullptr;
  const int32_t* plus = nullptr;
  const int32_t* minus = nullptr;
};

// Right-hand sides of a column-major block (column c at c * ld, by unknown), or pairs built on
// the fly when `block` is null.
struct ColumnsDev {
  const double* block = nullptr;
  int64_t ld = 0;
  const int32_t* plus = nullptr;
  const int32_t* minus = nullptr;
};

// One factor of a batch: its column-path sweeps and maps, its order, and where its first column
// starts in the batch's right-hand-side and solution arrays (columns of n values, one after
// another).
struct BatchEntryDev {
  SweepDev fwd, bwd;
  const int32_t* rhs_row = nullptr;
  const int32_t* sol_of = nullptr;
  int32_t n = 0;
  int64_t offset = 0;
  CsrDev a;                  // the factor's matrix (FP32-refined batches)
  int64_t first_column = 0;  // its first column's index among all the batch's columns
};

// LDS bytes the column solves need for n rows and the given total of both sweeps' levels (and
// `threads` per workgroup for refinement).
size_t solve_lds_bytes(int32_t n, bool fp32, int32_t levels);
size_t refine_lds_bytes(int32_t n, int threads, int32_t levels);

// One workgroup per column, the solution in LDS in factor order: rhs_row[u] is the factor row
// holding right-hand side entry u, sol_of[u] the factor row holding solution entry u; the
// sweeps are in factor order. x is column-major (ldx). fp32 sweeps in FP32 (to FP64 output).
hipError_t launch_solve_lds(hipStream_t stream, int threads, bool fp32, const SweepDev& fwd, const SweepDev& bwd,
                            int32_t n, const int32_t* rhs_row, const int32_t* sol_of, const ColumnsDev& rhs,
                            double* x, int64_t ldx, int32_t ncols);

// FP32 sweeps in LDS refined in FP64, one workgroup per column, blocks column-major (ld).
hipError_t launch_refine_lds(hipStream_t stream, int threads, const SweepDev& fwd, const SweepDev& bwd,
                             const CsrDev& a, const int32_t* rhs_row, const int32_t* sol_of, const double* b,
                             double* x0, double* x1, double* r, int64_t ld, int32_t ncols, double tol,
                             int32_t max_steps, int32_t* steps, double* ratio, int32_t floor_rule);

// FP64 forward then backward sweep over column tiles of width `tile` (1..64, a power of two),
// `threads` per workgroup (64..1024): t (packed) is the forward result, x the solution. t and x
// hold a pad row after their n rows.
hipError_t launch_solve64(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                          const RhsDev& rhs, double* t, double* x, int64_t ld, int32_t ncols);

// A sweep's sources for the windowed solve (window_code in analysis.hpp: a value in the LDS ring
// of the last `cap` positions, a row final by the level before the one before, or a row written
// by the level before): per position its head (dst, source, p0, p1) and its first four terms'
// sources (unused ones: the pad row), and every term's source. Values and diagonals come from the
// sweep's own records.
struct WindowDev {
  const int4* head = nullptr;
  const int4* dep4 = nullptr;
  const int32_t* dep = nullptr;
  int32_t cap = 0;
};
// The FP64 solve with each tile's last `cap` positions' values in an LDS ring (both sweeps' cap
// the same): a term in the ring is read from LDS, a term final three levels back is loaded two
// levels ahead, and only the others go to the block when their row is swept. LDS bytes for a tile width,
// the largest ring a compute unit's LDS allows, and the launch (at most 512 threads).
size_t window_lds_bytes(const SweepDev& fwd, const SweepDev& bwd, int tile, int32_t cap);
int32_t window_cap_for(const SweepDev& fwd, const SweepDev& bwd, int tile, size_t lds_cap);
hipError_t launch_solve64_window(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const WindowDev& wf,
                                 const SweepDev& bwd, const WindowDev& wb, const RhsDev& rhs, double* t, double* x,
                                 int64_t ld, int32_t ncols);

// The reach of a pair block's tiles: tile b's forward rows are the positions pos[lp[b (levels +
// 1) + l] ..) of each level l, in level order (levels = the forward sweep's).
struct ReachDev {
  const int32_t* pos = nullptr;
  const int32_t* lp = nullptr;
  const uint32_t* live = nullptr;  // per tile, a bit per forward row in its reach (words each)
  int32_t words = 0;
};

// LDS a reach solve needs (both sweeps' level pointers and a bit per forward row), and the solve
// itself: a pair block (rhs.plus/minus) whose forward sweep visits only each tile's reach.
// skip_loads: each row issues only its own inline terms' loads (short rows, a Cholesky factor's),
// else all four without branches (long rows, an LU factor's).
size_t reach_lds_bytes(const SweepDev& fwd, const SweepDev& bwd);
hipError_t launch_solve64_reach(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                                const RhsDev& rhs, const ReachDev& reach, double* t, double* x, int64_t ld,
                                int32_t ncols, bool skip_loads);

// The refinement launchers' floor_rule: 0 stops each column by refine_normwise's rule alone
// (accuracy.rs:99-120); 1 stops it also once its ratio is at most 2 u || |A| |x| ||inf / ||b||inf,
// the rounding floor of an FP64 residual, taken in the same residual pass (the correction after
// it would not halve the ratio)

// FP32 sweeps refined in FP64 to `tol` (refine_normwise's rule), one workgroup per column tile;
// the best iterate ends in x0. steps/ratio get each column's corrections and relative residual.
hipError_t launch_refine(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                         const CsrDev& a, const double* b, float* t, float* d, double* x0, double* x1, double* r,
                         int64_t ld, int32_t ncols, double tol, int32_t max_steps, int32_t* steps, double* ratio,
                         int32_t floor_rule);

// Layout changes between count vectors of n values, one after another (the host's &[Vec<f64>])
// and a row-interleaved block (n rows x ld, column c of row i at i * ld + c).
// FP64 solve of a block over partitioned sweeps (fwd: parts, then the top part; bwd: the top
// part first, then the parts), in three launches: every (tile, part) of the forward parts, then
// per tile the top part forward and backward, then every (tile, part) of the backward parts.
// Each row keeps its own terms and their order, so the bits are those of launch_solve64. With
// reach.lp set the block is pairs and each tile sweeps only its reach forward, as
// launch_solve64_reach does (reach.lp over fwd's levels, reach.live the tiles' bitmaps). The top
// kernel takes a configuration of its own (top_tile columns a workgroup, top_threads): its few
// long levels want more, narrower workgroups than the parts. A pair block keeps top_tile = tile.
struct PartLevels {
  int32_t fwd_parts = 0, fwd_top = 0, bwd_top = 0, bwd_parts = 0;  // the most levels of one part, and the top's
};

// The top parts of the two sweeps held in LDS (the host analysis's top_slots): slots for the top
// rows, then a halo slot for each row of the parts a forward top segment reads. The top's
// metadata is one image in device memory, laid out as top_lds_layout places it after the slots'
// values, so a workgroup stages it with one coalesced copy. Per sweep: a 64-byte record per
// segment (TopRecord: its slot, its term count, where its terms past the fourth start in the
// overflow arrays, its first four terms' slots and values), the segments' diagonals, the level
// starts, and the overflow terms' values and slots. (Only Cholesky factors are partitioned, and
// their values are fixed at upload, so the image keeps copies.)
struct TopRecord {
  int4 head;   // slot, terms, first overflow term, 0
  int4 dep;    // the first four terms' slots (unused ones: the segment's own slot)
  double2 v01, v23;  // their values (unused ones: 0)
};
struct TopPartDev {
  int32_t levels = 0, segs = 0, overflow = 0;
};
struct TopSlotsDev {
  int32_t rows = 0, halo = 0;
  const int32_t* rhs_row = nullptr;   // rows: each top row's forward right-hand side row
  const int32_t* halo_row = nullptr;  // halo: the forward block's row each halo slot holds
  const int32_t* out_row = nullptr;   // rows: each top row's solution row
  const int4* meta = nullptr;         // the metadata image, meta_units 16-byte units
  int32_t meta_units = 0;
  TopPartDev fwd, bwd;
};
// Byte offsets in LDS for a tile of `tile` columns: the slots' values (slot-major, the tile's
// columns side by side), then the metadata image (its arrays' offsets from its start), each
// array 16-byte aligned
struct TopLdsLayout {
  size_t values = 0, meta = 0, meta_bytes = 0, total = 0;
  struct Part {
    size_t rec = 0, diag = 0, lp = 0, ov_val = 0, ov_dep = 0;
  } fwd, bwd;
};
__host__ __device__ inline TopLdsLayout top_lds_layout(const TopSlotsDev& top, int tile) {
  TopLdsLayout o;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t here = at;
    at += (bytes + 15) / 16 * 16;
    return here;
  };
  auto part = [&](const TopPartDev& p, TopLdsLayout::Part& q) {
    q.rec = take((size_t)p.segs * sizeof(TopRecord));
    q.diag = take((size_t)p.segs * sizeof(double));
    q.lp = take((size_t)(p.levels + 1) * sizeof(int32_t));
    q.ov_val = take((size_t)p.overflow * sizeof(double));
    q.ov_dep = take((size_t)p.overflow * sizeof(int32_t));
  };
  part(top.fwd, o.fwd);
  part(top.bwd, o.bwd);
  o.meta_bytes = at;
  o.values = 0;
  o.meta = ((size_t)(top.rows + top.halo) * (size_t)tile * sizeof(double) + 15) / 16 * 16;
  o.total = o.meta + o.meta_bytes;
  return o;
}

// The independent parts of one sweep (forward or backward) in LDS slots, for the LDS parts
// kernel: part q's segments are records rec[d.seg0 ..] with their diagonals (a continuing
// segment's is 1), slot- and term-numbered within the part; its slots' source rows in_row and
// result rows out_row from d.slot0, the halo's rows (backward: top rows of the solution block)
// from d.halo0; its rounds (one wavefront's pass over up to 64 of a level's segments:
// (first segment in the part, segments, 1 where the round ends its level, 0)) from d.round0, a
// multiple of PARTS_LOOKAHEAD of them (the last ones empty); its terms past a segment's fourth
// (overflow) from d.ov0, the records' head.z counting from there.
constexpr int PARTS_LOOKAHEAD = 4;  // rounds whose records are loaded ahead of the one swept
struct PartDesc {
  int32_t seg0, segs, slot0, rows, halo0, halo, round0, rounds, ov0, ov, pad[2];
};
struct PartSlotsDev {
  int32_t parts = 0;
  const PartDesc* desc = nullptr;
  const TopRecord* rec = nullptr;
  const double* diag = nullptr;
  const int32_t* in_row = nullptr;
  const int32_t* out_row = nullptr;
  const int32_t* halo_row = nullptr;
  const int4* rounds = nullptr;
  const double* ov_val = nullptr;
  const int32_t* ov_dep = nullptr;
  int32_t most_slots = 0, most_ov = 0, most_rounds = 0;  // the largest over the parts
};
// Byte offsets in LDS for a tile of `tile` columns: the slots' values (slot-major, the tile's
// columns side by side), the overflow terms' values and slots, and the rounds (with
// PARTS_LOOKAHEAD empty ones after the last)
struct PartsLdsLayout {
  size_t values = 0, ov_val = 0, ov_dep = 0, rounds = 0, total = 0;
};
__host__ __device__ inline PartsLdsLayout parts_lds_layout(const PartSlotsDev& p, int tile) {
  PartsLdsLayout o;
  auto up = [](size_t bytes) { return (bytes + 15) / 16 * 16; };
  o.ov_val = up((size_t)p.most_slots * (size_t)tile * sizeof(double));
  o.ov_dep = o.ov_val + up((size_t)p.most_ov * sizeof(double));
  o.rounds = o.ov_dep + up((size_t)p.most_ov * sizeof(int32_t));
  o.total = o.rounds + (size_t)(p.most_rounds + PARTS_LOOKAHEAD) * sizeof(int4);
  return o;
}
// The LDS parts kernels' configuration: tile 0 takes the parts kernels over the block in global
// memory
struct PartsLds {
  const PartSlotsDev* fwd = nullptr;
  const PartSlotsDev* bwd = nullptr;
  int tile = 0, threads = 0;
};

// top_lds: the top kernel over TopSlotsDev in LDS (null: the top kernel over the block in global
// memory). parts_lds: the parts swept in LDS slots (tile 0: over the block in global memory).
// Blocks of vectors, and pair blocks without a reach, only, for either.
hipError_t launch_solve64_parts(hipStream_t stream, int tile, int threads, int top_tile, int top_threads,
                                const SweepDev& fwd, const SweepDev& bwd, const PartLevels& levels, const RhsDev& rhs,
                                const ReachDev& reach, const TopSlotsDev* top_lds, const PartsLds& parts_lds,
                                double* t, double* x, int64_t ld, int32_t ncols);

// Complex FP64 solve of a block: forward sweep b -> t, backward t -> x. A complex block holds
// unknown i's real parts in row 2i and its imaginary parts in row 2i + 1 (ld doubles a row).
// conjugate: solve with conj(b) and return conj of the solution.
hipError_t launch_solve_complex(hipStream_t stream, int tile, int threads, const ComplexSweepDev& fwd,
                                const ComplexSweepDev& bwd, const double* b, double* t, double* x, int64_t ld,
                                int32_t ncols, bool conjugate);

// Complex FP32 sweeps (the sweeps' val32 / diag32) refined in FP64 against the complex matrix a,
// one tile of `tile` columns per workgroup of `threads` (a power of two): b the right-hand sides,
// t and d FP32 complex blocks, x0 (the result) and x1 the iterates, r the residuals, all complex
// blocks of ld doubles a row; each column's corrections and final ratio into steps / ratio.
hipError_t launch_refine_complex(hipStream_t stream, int tile, int threads, const ComplexSweepDev& fwd,
                                 const ComplexSweepDev& bwd, const ZCsrDev& a, const double* b, float* t, float* d,
                                 double* x0, double* x1, double* r, int64_t ld, int32_t ncols, double tol,
                                 int32_t max_steps, int32_t* steps, double* ratio, int32_t floor_rule);

hipError_t launch_to_block(hipStream_t stream, const double* vectors, double* block, int64_t n, int64_t count,
                           int64_t ld);
hipError_t launch_from_block(hipStream_t stream, const double* block, double* vectors, int64_t n, int64_t count,
                             int64_t ld);

// Blocks below are addressed with strides: entry (i, c) at i * rs + c * cs (row-interleaved:
// rs = ld, cs = 1; column-major: rs = 1, cs = ld).

// pair_update (ybus_cuda/src/kernels.cu:24-58) on a solved block y (by unknown), row-interleaved
// or column-major; coef holds one double per column (the update's coefficient), and singular gets
// 1 or 0 for every column.
hipError_t launch_pair_update(hipStream_t stream, const double* y, double* out, int32_t n, int32_t ncols, int64_t rs,
                              int64_t cs, const double* base, const int32_t* plus, const int32_t* minus,
                              const double* alpha, const double* inverse_weight, int32_t* singular, double* coef);

// Outages of sets of branches (dc/outage.rs:126-241) over a solved pair block y (one column per
// pair; set s holds pairs set_ptr[s] .. set_ptr[s + 1]): per set, W = D^-1 - C^T Y (row-major at
// w + w_at[s]) and C^T x with x = base + sum alpha_k Y_k (at ctx + set_ptr[s]); then, with z = W^-1
// C^T x per set, out(:, s) = x + sum z_j Y_j (x alone where singular[s]).
hipError_t launch_outage_system(hipStream_t stream, const double* y, int64_t rs, int64_t cs, const double* base,
                                const int32_t* set_ptr, int32_t sets, const int32_t* plus, const int32_t* minus,
                                const double* alpha, const double* inverse_weight, const int32_t* w_at, double* w,
                                double* ctx);
hipError_t launch_outage_apply(hipStream_t stream, const double* y, int64_t rs, int64_t cs, int32_t n,
                               const double* base, const int32_t* set_ptr, const double* alpha, const double* z,
                               const int32_t* singular, double* out, int64_t ors, int64_t ocs, int32_t sets);

// column_max (kernels.cu:78-128): per column, the largest |v| * weight[row] (smallest row on a
// tie, -1 if every product is zero) and whether any value isn't finite. A row-interleaved block
// is split into `chunks` runs of rows whose results (pv, pr, pb: chunks x ncols each) are then
// combined; max_chunks gives the count that fills the device. A column-major block (rs = 1)
// takes one workgroup per column and no partials.
int32_t max_chunks(int32_t rows, int32_t ncols, int cus);
hipError_t launch_column_max(hipStream_t stream, const double* block, int32_t rows, int32_t ncols, int64_t rs,
                             int64_t cs, const double* weight, int32_t chunks, double* pv, int32_t* pr, int32_t* pb,
                             double* best, int32_t* best_row, int32_t* not_finite);

// The same maxima over M x + offset with row zero[c] of column c zeroed (the operator's product,
// pairs.rs:312-366), computed on the fly rather than stored.
hipError_t launch_product_max(hipStream_t stream, const CsrDev& m, const double* x, int64_t rs, int64_t cs,
                              const double* offset, const int32_t* zero, const double* weight, int32_t ncols,
                              int32_t chunks, double* pv, int32_t* pr, int32_t* pb, double* best, int32_t* best_row,
                              int32_t* not_finite);

// out = M x + offset for each column, then out(zero[c], c) = 0 where zero[c] >= 0; out's entry
// (r, c) at r * ors + c * ocs.
hipError_t launch_operator(hipStream_t stream, const CsrDev& m, const double* x, int64_t rs, int64_t cs,
                           const double* offset, const int32_t* zero, double* out, int64_t ors, int64_t ocs,
                           int32_t ncols);

// out[order[q]] = -acc[q] for each stored entry q (-0.0 everywhere when `any` is false: the
// adjoint of no vectors)
hipError_t launch_scatter_negated(hipStream_t stream, const double* acc, const int32_t* order, int64_t nnz, bool any,
                                  double* out);

// Batches: a persistent launch of `grid` workgroups that take (factor, column) items from the
// queue counter (which must hold 0 at the launch) until `total` are taken. prepare_batch allows
// the kernel lds_bytes of LDS (call it before a launch or a graph capture). per_cu: the class's
// workgroups a compute unit (one or two: workgroups of at most 512 threads, else 1,024).
hipError_t prepare_batch(size_t lds_bytes, int per_cu);
hipError_t launch_batch(hipStream_t stream, int grid, int threads, size_t lds_bytes, int per_cu,
                        const BatchEntryDev* entries, const int2* items, int32_t total, int32_t* queue,
                        const double* rhs, double* x);
// The FP32-refined batch: the same queue, each column refined in FP64 against its factor's matrix
// (x1 and r: scratch laid out like x; steps and ratio: one per column of the batch)
hipError_t prepare_batch_refine(size_t lds_bytes);
hipError_t launch_batch_refine(hipStream_t stream, int grid, int threads, size_t lds_bytes,
                               const BatchEntryDev* entries, const int2* items, int32_t total, int32_t* queue,
                               const double* rhs, double* x, double* x1, double* r, double tol, int32_t max_steps,
                               int32_t* steps, double* ratio, int32_t floor_rule);
// p[0..count) = value, as a kernel (so it can sit in a captured graph)
hipError_t launch_fill(hipStream_t stream, int32_t* p, int32_t value, int32_t count);

// The adjoint's matrix gradient on A's pattern: for each stored entry q = (row[q], col[q]),
// acc[q] = (first ? 0 : acc[q]) + sum over columns c ascending of bbar(row, c) * x(col, c); both
// blocks addressed with strides (rs, cs) as above.
hipError_t launch_sddmm(hipStream_t stream, const int32_t* row, const int32_t* col, int64_t nnz, const double* bbar,
                        const double* x, int64_t rs, int64_t cs, int32_t ncols, double* acc, bool first);

// ---- numeric refactorization of an LU factor on the device (lu.rs:662-757) ----
//
// Column k of the ordered matrix P A Q is gathered into x (its slots: U(:, k)'s rows in order,
// the diagonal last, then L(:, k)'s rows below the diagonal), then for each row j above the
// diagonal in U(:, k), in increasing order, u_jk = x[slot of j] and x[slot] -= l_ij * u_jk over
// L(:, j) below its diagonal, then the pivot check and l = x / pivot. Columns are taken in an
// order where each comes after the columns it reads (by level), each by one wavefront that
// waits for the L values it reads to be written.
struct RefactorDev {
  int32_t n = 0;
  // per position c of the column order: (column k, its first U entry, U entries, first L entry)
  // and (L entries below the diagonal, first and end matrix entry of its column, first target)
  const int4* head = nullptr;
  const int4* span = nullptr;
  // per U entry above a diagonal (at its position in U): (row j, first L entry below j's
  // diagonal, their count, 0)
  const int4* step = nullptr;
  const int32_t* a_slot = nullptr;     // per matrix entry: its slot in its column's x
  const double* a_val = nullptr;       // the matrix's values, as stored by the caller (CSC)
  const int32_t* target = nullptr;     // per multiply-subtract: the slot it updates
  // the new factors, in the uploaded CSC order; L's entries must hold REFACTOR_PENDING at the
  // launch (a column's values are written in place of it, which is what a reading column waits for)
  double* l_val = nullptr;
  double* u_val = nullptr;
  int32_t* control = nullptr;  // [0] next column to take (0 at launch), [1] the smallest column whose
                               // pivot fails (n at launch), [2] waits given up (0 at launch)
  double tau = 0.1;
  int32_t x_cap = 0;          // slots of x in LDS (at least the longest column's)
  uint32_t spin_limit = 0;    // polls of a value before a wait is given up
};
// The bits of a value not written yet: a signalling NaN, which no arithmetic result is (an
// operation on a NaN gives a quiet one)
constexpr uint64_t REFACTOR_PENDING = 0x7FF4C0DE0000DEADull;
// p[0..count) = value, as a kernel
hipError_t launch_fill_u64(hipStream_t stream, uint64_t* p, uint64_t value, int64_t count);
// dst[i] = src[idx[i]] for i < count
hipError_t launch_gather(hipStream_t stream, const double* src, const int32_t* idx, int64_t count, double* dst);
// The entries one batch of steps stages: 64, 128, 256 or 512 (a batch holds at least one step, so
// this bounds the longest column of L below its diagonal; the smallest that fits is fastest)
constexpr int32_t REFACTOR_STAGE = 512;
int32_t refactor_stage_for(int32_t longest_l);
// LDS of one wavefront
size_t refactor_lds_bytes(int32_t x_cap, int32_t stage, bool complex);
// `grid` single-wavefront workgroups take the columns, staging `stage` entries per batch;
// control[0], [2] must hold 0 and control[1] n at the launch. A complex factor's values (the
// matrix's, L's and U's) are (re, im) pairs of doubles, and each part of L's holds
// REFACTOR_PENDING at the launch.
hipError_t launch_refactor(hipStream_t stream, const RefactorDev& d, int grid, int32_t stage, bool complex);

// A sweep's value arrays rewritten from new factor values: every entry takes src[vpos[p]] (a
// position in the new L values followed by the new U values), every diagonal src[dpos[i]], in
// FP64 and FP32 and in the rows' inline records (their first four entries; zero past a row's end).
struct RefreshDev {
  int32_t n = 0;
  int64_t entries = 0;
  const int32_t* row_ptr = nullptr;
  const int32_t* vpos = nullptr;
  const int32_t* dpos = nullptr;  // null for a unit diagonal
  double* val64 = nullptr;
  float* val32 = nullptr;
  double* diag64 = nullptr;
  float* diag32 = nullptr;
  double2* val01 = nullptr;
  double2* val23 = nullptr;
  float4* v32 = nullptr;
  double* rdiag64 = nullptr;
  float* rdiag32 = nullptr;
};
hipError_t launch_refresh(hipStream_t stream, const RefreshDev& r, const double* src);
// A complex sweep's values (one (re, im) pair per entry, and per row unless its diagonal is unit)
// from new factor values: val[p] = src[vpos[p]], diag[i] = src[dpos[i]], and their FP32 copies
// val32 / diag32 rounded from them (either may be null)
hipError_t launch_refresh_complex(hipStream_t stream, int32_t n, int64_t entries, const int32_t* vpos,
                                  const int32_t* dpos, const double2* src, double2* val, double2* diag, float2* val32,
                                  float2* diag32);

// The workspace of the entry point that allocates nothing (ybus_hip_factor_solve_device_noalloc):
// a file-scope device array of SCRATCH_BYTES in kernels.hip, which every device the library runs
// on holds until the process ends. scratch_address gives its address on `device`, which must be
// the current device; the first call per device looks it up (hipGetSymbolAddress), later calls
// return the address kept for that device. Host code passes that address, never the array's
// name, to kernels and copies.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
hipError_t scratch_address(int device, void** out);

}  // namespace ybus_hip
