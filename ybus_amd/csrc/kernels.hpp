// Launchers of the device kernels (kernels.hip). Every launcher enqueues on `stream` and
// returns the launch's hipError_t; shapes are checked by the caller before any launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ybus_hip {

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
};

// A CSR matrix on the device.
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

// FP64 forward then backward sweep over column tiles of width `tile` (1..64, a power of two),
// `threads` per workgroup (64..1024): t (packed) is the forward result, x the solution.
hipError_t launch_solve64(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                          const RhsDev& rhs, double* t, double* x, int64_t ld, int32_t ncols);

// FP32 sweeps refined in FP64 to `tol` (refine_normwise's rule), one workgroup per column tile;
// the best iterate ends in x0. steps/ratio get each column's corrections and relative residual.
hipError_t launch_refine(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                         const CsrDev& a, const double* b, float* t, float* d, double* x0, double* x1, double* r,
                         int64_t ld, int32_t ncols, double tol, int32_t max_steps, int32_t* steps, double* ratio);

// Layout changes between count vectors of n values, one after another (the host's &[Vec<f64>])
// and a row-interleaved block (n rows x ld, column c of row i at i * ld + c).
hipError_t launch_to_block(hipStream_t stream, const double* vectors, double* block, int64_t n, int64_t count,
                           int64_t ld);
hipError_t launch_from_block(hipStream_t stream, const double* block, double* vectors, int64_t n, int64_t count,
                             int64_t ld);

// pair_update (ybus_cuda/src/kernels.cu:24-58) on a solved block y (n x ld, by unknown).
hipError_t launch_pair_update(hipStream_t stream, const double* y, double* out, int32_t n, int32_t ncols, int64_t ld,
                              const double* base, const int32_t* plus, const int32_t* minus, const double* alpha,
                              const double* inverse_weight, int32_t* singular);

// column_max (kernels.cu:78-128): per column, the largest |v| * weight[row] (smallest row on a
// tie, -1 if every product is zero) and whether any value isn't finite.
hipError_t launch_column_max(hipStream_t stream, const double* block, int32_t rows, int32_t ncols, int64_t ld,
                             const double* weight, double* best, int32_t* best_row, int32_t* not_finite);

// out = M x + offset for each column, then out[zero[c]][c] = 0 where zero[c] >= 0.
hipError_t launch_operator(hipStream_t stream, const CsrDev& m, const double* x, int64_t ldx, const double* offset,
                           const int32_t* zero, double* out, int64_t ldo, int32_t ncols);

}  // namespace ybus_hip
