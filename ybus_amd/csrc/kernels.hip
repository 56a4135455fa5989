// Device kernels of the factor-solve path, written for CDNA4 (gfx950, wave64).
//
// Layout: a block of right-hand sides is row-interleaved, n rows x ld, column c of row i at
// i * ld + c, so the lanes of a wavefront that take consecutive columns of one row load and
// store consecutive doubles: a sweep's gathers x[dep][c..c+TW) are coalesced.
//
// Work split: one workgroup owns a tile of TW columns for the whole solve and walks the level
// schedule with a workgroup barrier between levels; lane (slot, c) of the workgroup takes rows
// slot, slot + slots, ... of each level for column c. No workgroup waits on another, so there is
// no cross-XCD synchronisation, no spinning and no deadlock to rule out, and a launch ends when
// its last tile's last level does.
//
// Arithmetic: the library is compiled with -ffp-contract=off, and each row subtracts its terms
// one at a time in the order the CPU solve does (the sweep's packed entry order), so the FP64
// solves give ybus_sparse's bits exactly.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace ybus_hip {

namespace {

template <int TW>
struct Tile {
  int lane, slot, slots;
  int32_t col;
  bool on;
  __device__ Tile(int32_t ncols) {
    lane = (int)threadIdx.x % TW;
    slot = (int)threadIdx.x / TW;
    slots = (int)blockDim.x / TW;
    col = (int32_t)blockIdx.x * TW + lane;
    on = col < ncols;
  }
};

// A max that ignores NaN, as the reference's infinity-norm fold does (accuracy.rs:66-69)
__device__ __forceinline__ double fold_max(double m, double v) { return v > m ? v : m; }

// Reduces each lane's value over the workgroup's slots into out[lane] (every thread must call).
template <int TW>
__device__ void reduce_max(double* scratch, double mine, double* out, const Tile<TW>& t) {
  scratch[threadIdx.x] = mine;
  __syncthreads();
  for (int stride = t.slots / 2; stride > 0; stride /= 2) {
    if (t.slot < stride) scratch[threadIdx.x] = fold_max(scratch[threadIdx.x], scratch[threadIdx.x + stride * TW]);
    __syncthreads();
  }
  if (t.slot == 0) out[t.lane] = scratch[threadIdx.x];
  __syncthreads();
}

// ---- FP64 solve: forward into t, barrier, backward into x ----

template <int TW>
__global__ void __launch_bounds__(1024) solve64_kernel(SweepDev fwd, SweepDev bwd, RhsDev rhs, double* __restrict__ t,
                                                       double* __restrict__ x, int64_t ld, int32_t ncols) {
  const Tile<TW> g(ncols);
  for (int32_t l = 0; l < fwd.levels; ++l) {
    const int32_t end = fwd.level_ptr[l + 1];
    if (g.on) {
      for (int32_t k = fwd.level_ptr[l] + g.slot; k < end; k += g.slots) {
        const int32_t s = fwd.src[k];
        double acc;
        if (rhs.block) {
          acc = rhs.block[s * ld + g.col];
        } else {
          // e_plus - e_minus as the host builds it: 1.0, -1.0 or 0.0
          acc = (s == rhs.plus[g.col] ? 1.0 : 0.0) - (s == rhs.minus[g.col] ? 1.0 : 0.0);
        }
        const int32_t pe = fwd.row_ptr[k + 1];
        for (int32_t p = fwd.row_ptr[k]; p < pe; ++p) acc = acc - fwd.val64[p] * t[fwd.dep[p] * ld + g.col];
        t[fwd.dst[k] * ld + g.col] = fwd.diag64 ? acc / fwd.diag64[k] : acc;
      }
    }
    __syncthreads();
  }
  for (int32_t l = 0; l < bwd.levels; ++l) {
    const int32_t end = bwd.level_ptr[l + 1];
    if (g.on) {
      for (int32_t k = bwd.level_ptr[l] + g.slot; k < end; k += g.slots) {
        double acc = t[bwd.src[k] * ld + g.col];
        const int32_t pe = bwd.row_ptr[k + 1];
        for (int32_t p = bwd.row_ptr[k]; p < pe; ++p) acc = acc - bwd.val64[p] * x[bwd.dep[p] * ld + g.col];
        x[bwd.dst[k] * ld + g.col] = bwd.diag64 ? acc / bwd.diag64[k] : acc;
      }
    }
    __syncthreads();
  }
}

// ---- FP32 sweeps refined in FP64 ----

// Forward FP32 sweep of one tile: t[dst] = (float(src[src_row]) - sum val * t[dep]) / diag
template <int TW>
__device__ void forward32(const SweepDev& s, const double* __restrict__ src, float* __restrict__ t, int64_t ld,
                          const Tile<TW>& g, bool act) {
  for (int32_t l = 0; l < s.levels; ++l) {
    const int32_t end = s.level_ptr[l + 1];
    if (act) {
      for (int32_t k = s.level_ptr[l] + g.slot; k < end; k += g.slots) {
        float acc = (float)src[s.src[k] * ld + g.col];
        const int32_t pe = s.row_ptr[k + 1];
        for (int32_t p = s.row_ptr[k]; p < pe; ++p) acc = acc - s.val32[p] * t[s.dep[p] * ld + g.col];
        t[s.dst[k] * ld + g.col] = s.diag32 ? acc / s.diag32[k] : acc;
      }
    }
    __syncthreads();
  }
}

// Backward FP32 sweep of one tile into d (its own dependencies), updating the FP64 iterate:
// xn[dst] = xc[dst] + double(d) (or double(d) alone when xc is null)
template <int TW>
__device__ void backward32(const SweepDev& s, const float* __restrict__ t, float* __restrict__ d,
                           const double* __restrict__ xc, double* __restrict__ xn, int64_t ld, const Tile<TW>& g,
                           bool act) {
  for (int32_t l = 0; l < s.levels; ++l) {
    const int32_t end = s.level_ptr[l + 1];
    if (act) {
      for (int32_t k = s.level_ptr[l] + g.slot; k < end; k += g.slots) {
        float acc = t[s.src[k] * ld + g.col];
        const int32_t pe = s.row_ptr[k + 1];
        for (int32_t p = s.row_ptr[k]; p < pe; ++p) acc = acc - s.val32[p] * d[s.dep[p] * ld + g.col];
        const float v = s.diag32 ? acc / s.diag32[k] : acc;
        const int64_t at = s.dst[k] * ld + g.col;
        d[at] = v;
        xn[at] = xc ? xc[at] + (double)v : (double)v;
      }
    }
    __syncthreads();
  }
}

// r = b - A x for the tile's columns (each row summed from 0 in index order, as mul_vec's
// dot_kernel does, product.rs:17-27); returns this thread's partial max |r|
template <int TW>
__device__ double residual(const CsrDev& a, const double* __restrict__ b, const double* __restrict__ x,
                           double* __restrict__ r, int64_t ld, const Tile<TW>& g, bool act) {
  double m = 0.0;
  if (!act) return m;
  for (int32_t i = g.slot; i < a.rows; i += g.slots) {
    double sum = 0.0;
    const int32_t pe = a.ptr[i + 1];
    for (int32_t p = a.ptr[i]; p < pe; ++p) sum = sum + a.val[p] * x[a.col[p] * ld + g.col];
    const double v = b[i * ld + g.col] - sum;
    r[i * ld + g.col] = v;
    m = fold_max(m, fabs(v));
  }
  return m;
}

__device__ __forceinline__ double ratio_of(double r_norm, double b_norm) {
  return b_norm > 0.0 ? r_norm / b_norm : (r_norm > 0.0 ? __builtin_inf() : 0.0);
}

template <int TW>
__global__ void __launch_bounds__(1024) refine_kernel(SweepDev fwd, SweepDev bwd, CsrDev a, const double* __restrict__ b,
                                                      float* __restrict__ t, float* __restrict__ d,
                                                      double* __restrict__ x0, double* __restrict__ x1,
                                                      double* __restrict__ r, int64_t ld, int32_t ncols, double tol,
                                                      int32_t max_steps, int32_t* __restrict__ steps_out,
                                                      double* __restrict__ ratio_out) {
  __shared__ double scratch[1024];
  __shared__ double s_bnorm[TW], s_rnorm[TW], s_ratio[TW], s_best[TW];
  __shared__ int32_t s_best_buf[TW], s_steps[TW], s_active[TW];
  const Tile<TW> g(ncols);

  // ||b||inf per column
  double m = 0.0;
  if (g.on)
    for (int32_t i = g.slot; i < a.rows; i += g.slots) m = fold_max(m, fabs(b[i * ld + g.col]));
  reduce_max<TW>(scratch, m, s_bnorm, g);

  // x0 = solve(b), its residual and ratio
  forward32<TW>(fwd, b, t, ld, g, g.on);
  backward32<TW>(bwd, t, d, nullptr, x0, ld, g, g.on);
  m = residual<TW>(a, b, x0, r, ld, g, g.on);
  reduce_max<TW>(scratch, m, s_rnorm, g);
  if (g.slot == 0) {
    const double ratio = ratio_of(s_rnorm[g.lane], s_bnorm[g.lane]);
    s_ratio[g.lane] = ratio;
    s_best[g.lane] = ratio;
    s_best_buf[g.lane] = 0;
    s_steps[g.lane] = 0;
    s_active[g.lane] = g.on && max_steps > 0 && !(ratio <= tol);
  }
  __syncthreads();

  // refine_normwise (accuracy.rs:99-120): while a column is active, solve for its residual and
  // add the correction into the other buffer. A column stays active only while each step at
  // least halves its ratio, so its best iterate is always the current or the previous one.
  double* cur = x0;
  double* nxt = x1;
  for (int32_t step = 1; step <= max_steps; ++step) {
    bool any = false;
    for (int c = 0; c < TW; ++c) any = any || s_active[c];
    if (!any) break;
    const bool act = g.on && s_active[g.lane];
    forward32<TW>(fwd, r, t, ld, g, act);
    backward32<TW>(bwd, t, d, cur, nxt, ld, g, act);
    m = residual<TW>(a, b, nxt, r, ld, g, act);
    reduce_max<TW>(scratch, m, s_rnorm, g);
    if (g.slot == 0 && act) {
      const double previous = s_ratio[g.lane];
      const double ratio = ratio_of(s_rnorm[g.lane], s_bnorm[g.lane]);
      s_ratio[g.lane] = ratio;
      if (ratio < s_best[g.lane]) {
        s_best[g.lane] = ratio;
        s_best_buf[g.lane] = nxt == x1 ? 1 : 0;
        s_steps[g.lane] = step;
      }
      const bool stop = ratio != ratio || ratio > 0.5 * previous || ratio <= tol || step == max_steps;
      s_active[g.lane] = !stop;
    }
    __syncthreads();
    double* swap = cur;
    cur = nxt;
    nxt = swap;
  }

  // the best iterate into x0
  if (g.on && s_best_buf[g.lane] == 1)
    for (int32_t i = g.slot; i < a.rows; i += g.slots) x0[i * ld + g.col] = x1[i * ld + g.col];
  if (g.on && g.slot == 0) {
    steps_out[g.col] = s_steps[g.lane];
    ratio_out[g.col] = s_best[g.lane];
  }
}

// ---- LDS-resident column solves ----
//
// One workgroup per right-hand side: its whole solution vector lives in LDS in factor order,
// the forward and backward sweeps run in place there (row i reads its own value at slot i and
// its dependencies at theirs; the backward sweep of row i needs only rows already final), and
// global memory sees only the column's right-hand side read once and its solution written once,
// both contiguous. The sweep schedules used here are in factor order (src = dst = row).

template <typename T>
__device__ __forceinline__ const T* values_of(const SweepDev& s);
template <>
__device__ __forceinline__ const double* values_of<double>(const SweepDev& s) {
  return s.val64;
}
template <>
__device__ __forceinline__ const float* values_of<float>(const SweepDev& s) {
  return s.val32;
}
template <typename T>
__device__ __forceinline__ const T* diagonal_of(const SweepDev& s);
template <>
__device__ __forceinline__ const double* diagonal_of<double>(const SweepDev& s) {
  return s.diag64;
}
template <>
__device__ __forceinline__ const float* diagonal_of<float>(const SweepDev& s) {
  return s.diag32;
}

// One in-place sweep over xs, a barrier after each level. Each row's terms are fetched four at
// a time before any is subtracted (so their latencies overlap) and subtracted in order.
template <typename T>
__device__ void lds_sweep(const SweepDev& s, T* xs) {
  const T* __restrict__ val = values_of<T>(s);
  const T* __restrict__ diag = diagonal_of<T>(s);
  for (int32_t l = 0; l < s.levels; ++l) {
    const int32_t end = s.level_ptr[l + 1];
    for (int32_t k = s.level_ptr[l] + (int32_t)threadIdx.x; k < end; k += (int32_t)blockDim.x) {
      const int32_t i = s.dst[k];
      int32_t p = s.row_ptr[k];
      const int32_t pe = s.row_ptr[k + 1];
      T acc = xs[i];
      for (; p + 4 <= pe; p += 4) {
        const int32_t d0 = s.dep[p], d1 = s.dep[p + 1], d2 = s.dep[p + 2], d3 = s.dep[p + 3];
        const T v0 = val[p], v1 = val[p + 1], v2 = val[p + 2], v3 = val[p + 3];
        const T x0 = xs[d0], x1 = xs[d1], x2 = xs[d2], x3 = xs[d3];
        acc = acc - v0 * x0;
        acc = acc - v1 * x1;
        acc = acc - v2 * x2;
        acc = acc - v3 * x3;
      }
      for (; p < pe; ++p) acc = acc - val[p] * xs[s.dep[p]];
      xs[i] = diag ? acc / diag[k] : acc;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(1024) solve_lds_kernel(SweepDev fwd, SweepDev bwd, int32_t n,
                                                         const int32_t* __restrict__ rhs_row,
                                                         const int32_t* __restrict__ sol_of, ColumnsDev rhs,
                                                         double* __restrict__ x, int64_t ldx) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* xs = reinterpret_cast<T*>(smem);
  const int64_t col = blockIdx.x;
  if (rhs.block) {
    const double* b = rhs.block + col * rhs.ld;
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xs[rhs_row[u]] = (T)b[u];
  } else {
    // e_plus - e_minus as the host builds it: zeros, then 1.0 and -1.0
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) xs[i] = (T)0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int32_t p = rhs.plus[col], m = rhs.minus[col];
      if (p >= 0) xs[rhs_row[p]] = (T)1.0;
      if (m >= 0) xs[rhs_row[m]] = (T)-1.0;
    }
  }
  __syncthreads();
  lds_sweep<T>(fwd, xs);
  lds_sweep<T>(bwd, xs);
  double* out = x + col * ldx;
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) out[u] = (double)xs[sol_of[u]];
}

// Reduces one value per thread to thread 0's max (NaN ignored); every thread must call
__device__ double block_max(double* scratch, double mine) {
  scratch[threadIdx.x] = mine;
  __syncthreads();
  for (int stride = blockDim.x / 2; stride > 0; stride /= 2) {
    if ((int)threadIdx.x < stride) scratch[threadIdx.x] = fold_max(scratch[threadIdx.x], scratch[threadIdx.x + stride]);
    __syncthreads();
  }
  const double m = scratch[0];
  __syncthreads();
  return m;
}

// r = b - A x for one column (A by user rows, each row summed from 0 in index order as
// mul_vec's dot_kernel does); returns this thread's partial max |r|
__device__ double column_residual(const CsrDev& a, const double* __restrict__ b, const double* __restrict__ x,
                                  double* __restrict__ r) {
  double m = 0.0;
  for (int32_t u = threadIdx.x; u < a.rows; u += blockDim.x) {
    double sum = 0.0;
    const int32_t pe = a.ptr[u + 1];
    for (int32_t p = a.ptr[u]; p < pe; ++p) sum = sum + a.val[p] * x[a.col[p]];
    const double v = b[u] - sum;
    r[u] = v;
    m = fold_max(m, fabs(v));
  }
  return m;
}

// FP32 sweeps in LDS refined in FP64 (refine_normwise, accuracy.rs:83-122), one workgroup per
// column: x = solve(b); while the column is active, r = b - A x, d = solve(r), x' = x + d in
// the other buffer. A column continues only while each step at least halves its ratio, so its
// best iterate is the current or the previous one; it ends in x0.
__global__ void __launch_bounds__(1024) refine_lds_kernel(SweepDev fwd, SweepDev bwd, CsrDev a,
                                                          const int32_t* __restrict__ rhs_row,
                                                          const int32_t* __restrict__ sol_of,
                                                          const double* __restrict__ bblock, double* __restrict__ x0b,
                                                          double* __restrict__ x1b, double* __restrict__ rblock,
                                                          int64_t ld, double tol, int32_t max_steps,
                                                          int32_t* __restrict__ steps_out,
                                                          double* __restrict__ ratio_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* scratch = reinterpret_cast<double*>(smem);
  float* xs = reinterpret_cast<float*>(smem + blockDim.x * sizeof(double));
  const int32_t n = a.rows;
  const int64_t col = blockIdx.x;
  const double* b = bblock + col * ld;
  double* xa = x0b + col * ld;
  double* xb = x1b + col * ld;
  double* r = rblock + col * ld;

  double m = 0.0;
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) m = fold_max(m, fabs(b[u]));
  const double bnorm = block_max(scratch, m);

  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xs[rhs_row[u]] = (float)b[u];
  __syncthreads();
  lds_sweep<float>(fwd, xs);
  lds_sweep<float>(bwd, xs);
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xa[u] = (double)xs[sol_of[u]];
  __syncthreads();
  double ratio = ratio_of(block_max(scratch, column_residual(a, b, xa, r)), bnorm);
  double best = ratio;
  int32_t best_steps = 0;
  bool best_in_b = false;
  bool active = max_steps > 0 && !(ratio <= tol);
  double* cur = xa;
  double* nxt = xb;
  for (int32_t step = 1; active; ++step) {
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xs[rhs_row[u]] = (float)r[u];
    __syncthreads();
    lds_sweep<float>(fwd, xs);
    lds_sweep<float>(bwd, xs);
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) nxt[u] = cur[u] + (double)xs[sol_of[u]];
    __syncthreads();
    const double previous = ratio;
    ratio = ratio_of(block_max(scratch, column_residual(a, b, nxt, r)), bnorm);
    if (ratio < best) {
      best = ratio;
      best_steps = step;
      best_in_b = nxt == xb;
    }
    active = !(ratio != ratio || ratio > 0.5 * previous || ratio <= tol || step == max_steps);
    double* swap = cur;
    cur = nxt;
    nxt = swap;
  }
  if (best_in_b)
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xa[u] = xb[u];
  if (threadIdx.x == 0) {
    steps_out[col] = best_steps;
    ratio_out[col] = best;
  }
}

// ---- layout changes ----

constexpr int TR = 32;

__global__ void to_block_kernel(const double* __restrict__ v, double* __restrict__ block, int64_t n, int64_t count,
                                int64_t ld) {
  __shared__ double tile[TR][TR + 1];
  const int64_t i0 = (int64_t)blockIdx.x * TR, c0 = (int64_t)blockIdx.y * TR;
  for (int k = threadIdx.y; k < TR; k += blockDim.y) {
    const int64_t c = c0 + k, i = i0 + threadIdx.x;
    if (c < count && i < n) tile[k][threadIdx.x] = v[c * n + i];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < TR; k += blockDim.y) {
    const int64_t i = i0 + k, c = c0 + threadIdx.x;
    if (i < n && c < count) block[i * ld + c] = tile[threadIdx.x][k];
  }
}

__global__ void from_block_kernel(const double* __restrict__ block, double* __restrict__ v, int64_t n, int64_t count,
                                  int64_t ld) {
  __shared__ double tile[TR][TR + 1];
  const int64_t i0 = (int64_t)blockIdx.x * TR, c0 = (int64_t)blockIdx.y * TR;
  for (int k = threadIdx.y; k < TR; k += blockDim.y) {
    const int64_t i = i0 + k, c = c0 + threadIdx.x;
    if (i < n && c < count) tile[k][threadIdx.x] = block[i * ld + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < TR; k += blockDim.y) {
    const int64_t c = c0 + k, i = i0 + threadIdx.x;
    if (c < count && i < n) v[c * n + i] = tile[threadIdx.x][k];
  }
}

// ---- pairs, maxima, operator ----

// ybus_cuda/src/kernels.cu:24-58, over a block by unknown: x = base + alpha y,
// w = inverse_weight - c^T y, z = c^T x / w, out = x + z y; w == 0 flags the column singular
__global__ void pair_update_kernel(const double* __restrict__ y, double* __restrict__ out, int32_t n, int32_t ncols,
                                   int64_t rs, int64_t cs, const double* __restrict__ base, const int32_t* __restrict__ plus,
                                   const int32_t* __restrict__ minus, const double* __restrict__ alpha,
                                   const double* __restrict__ inverse_weight, int32_t* __restrict__ singular) {
  const int64_t index = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (index >= (int64_t)n * ncols) return;
  // consecutive threads take consecutive elements of whichever index has unit stride
  int32_t i, k;
  if (cs == 1) {
    i = (int32_t)(index / ncols);
    k = (int32_t)(index % ncols);
  } else {
    k = (int32_t)(index / n);
    i = (int32_t)(index % n);
  }
  const int32_t p = plus[k], m = minus[k];
  const double a = alpha[k];
  const int64_t c = k * cs;
  double cy = 0.0, cx = 0.0;
  if (p >= 0) {
    cy += y[p * rs + c];
    cx += base[p] + a * y[p * rs + c];
  }
  if (m >= 0) {
    cy -= y[m * rs + c];
    cx -= base[m] + a * y[m * rs + c];
  }
  const double w = -cy + inverse_weight[k];
  const double x = base[i] + a * y[i * rs + c];
  if (w == 0.0) {
    if (i == 0) singular[k] = 1;
    out[i * rs + c] = x;
    return;
  }
  out[i * rs + c] = x + (cx / w) * y[i * rs + c];
}

// One wavefront-wide column group per workgroup: lane = column, the workgroup's waves split
// the rows (each lane visits its rows in increasing order, so a strictly larger product wins),
// then the waves' results combine in LDS, the smaller row winning a tie.
__global__ void __launch_bounds__(256) column_max_kernel(const double* __restrict__ block, int32_t rows, int32_t ncols,
                                                         int64_t rs, int64_t cs, const double* __restrict__ weight,
                                                         double* __restrict__ best, int32_t* __restrict__ best_row,
                                                         int32_t* __restrict__ not_finite) {
  __shared__ double s_value[256];
  __shared__ int32_t s_row[256], s_bad[256];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, waves = blockDim.x / 64;
  const int32_t col = (int32_t)blockIdx.x * 64 + lane;
  double value = 0.0;
  int32_t row = -1, bad = 0;
  if (col < ncols) {
    for (int32_t i = wave; i < rows; i += waves) {
      const double v = block[i * rs + col * cs];
      if (!isfinite(v)) {
        bad = 1;
        continue;
      }
      const double product = fabs(v) * weight[i];
      if (product > value) {
        value = product;
        row = i;
      }
    }
  }
  s_value[threadIdx.x] = value;
  s_row[threadIdx.x] = row;
  s_bad[threadIdx.x] = bad;
  __syncthreads();
  if (wave == 0 && col < ncols) {
    for (int w = 1; w < waves; ++w) {
      const double v = s_value[w * 64 + lane];
      const int32_t r = s_row[w * 64 + lane];
      if (r >= 0 && (v > value || (v == value && (row < 0 || r < row)))) {
        value = v;
        row = r;
      }
      bad |= s_bad[w * 64 + lane];
    }
    best[col] = value;
    best_row[col] = row;
    not_finite[col] = bad;
  }
}

// Column-major blocks: one workgroup per column, threads over rows (each thread visits its rows
// in increasing order, so a strictly larger product wins), combined in LDS with the smaller row
// winning a tie: the same result as column_max_kernel.
__global__ void __launch_bounds__(256) column_max_cols_kernel(const double* __restrict__ block, int32_t rows,
                                                              int64_t cs, const double* __restrict__ weight,
                                                              double* __restrict__ best, int32_t* __restrict__ best_row,
                                                              int32_t* __restrict__ not_finite) {
  __shared__ double s_value[256];
  __shared__ int32_t s_row[256], s_bad[256];
  const int32_t col = (int32_t)blockIdx.x;
  const double* column = block + (int64_t)col * cs;
  double value = 0.0;
  int32_t row = -1, bad = 0;
  for (int32_t i = threadIdx.x; i < rows; i += blockDim.x) {
    const double v = column[i];
    if (!isfinite(v)) {
      bad = 1;
      continue;
    }
    const double product = fabs(v) * weight[i];
    if (product > value) {
      value = product;
      row = i;
    }
  }
  s_value[threadIdx.x] = value;
  s_row[threadIdx.x] = row;
  s_bad[threadIdx.x] = bad;
  __syncthreads();
  for (int stride = blockDim.x / 2; stride > 0; stride /= 2) {
    if ((int)threadIdx.x < stride) {
      const double v = s_value[threadIdx.x + stride];
      const int32_t r = s_row[threadIdx.x + stride];
      const double mv = s_value[threadIdx.x];
      const int32_t mr = s_row[threadIdx.x];
      if (r >= 0 && (v > mv || (v == mv && (mr < 0 || r < mr)))) {
        s_value[threadIdx.x] = v;
        s_row[threadIdx.x] = r;
      }
      s_bad[threadIdx.x] |= s_bad[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    best[col] = s_value[0];
    best_row[col] = s_row[0];
    not_finite[col] = s_bad[0];
  }
}

// out(r, c) = (sum_p M[r][p] x(p, c), summed from 0 in index order) + offset[r]: the host's
// M.mul_vec(theta) then + offset, term for term; then the zeroed entries. Entry (i, c) of x is
// at i * rs + c * cs and of out at r * ors + c * ocs; consecutive threads take consecutive
// entries of out.
__global__ void operator_kernel(CsrDev m, const double* __restrict__ x, int64_t rs, int64_t cs,
                                const double* __restrict__ offset, const int32_t* __restrict__ zero,
                                double* __restrict__ out, int64_t ors, int64_t ocs, int32_t ncols) {
  const int64_t index = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (index >= (int64_t)m.rows * ncols) return;
  int32_t r, c;
  if (ors == 1) {
    c = (int32_t)(index / m.rows);
    r = (int32_t)(index % m.rows);
  } else {
    r = (int32_t)(index / ncols);
    c = (int32_t)(index % ncols);
  }
  double sum = 0.0;
  for (int32_t p = m.ptr[r]; p < m.ptr[r + 1]; ++p) sum = sum + m.val[p] * x[m.col[p] * rs + c * cs];
  out[r * ors + c * ocs] = zero[c] == r ? 0.0 : sum + offset[r];
}

}  // namespace

// ---- launchers ----

#define DISPATCH_TILE(tile, F) \
  switch (tile) {              \
    case 1: F(1); break;       \
    case 2: F(2); break;       \
    case 4: F(4); break;       \
    case 8: F(8); break;       \
    case 16: F(16); break;     \
    case 32: F(32); break;     \
    case 64: F(64); break;     \
    default: return hipErrorInvalidValue; \
  }

hipError_t launch_solve64(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                          const RhsDev& rhs, double* t, double* x, int64_t ld, int32_t ncols) {
  if (threads < tile || threads > 1024 || threads % 64) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
#define L(TWV) hipLaunchKernelGGL(solve64_kernel<TWV>, dim3(grid), dim3(threads), 0, stream, fwd, bwd, rhs, t, x, ld, ncols)
  DISPATCH_TILE(tile, L)
#undef L
  return hipGetLastError();
}

hipError_t launch_refine(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                         const CsrDev& a, const double* b, float* t, float* d, double* x0, double* x1, double* r,
                         int64_t ld, int32_t ncols, double tol, int32_t max_steps, int32_t* steps, double* ratio) {
  if (threads < tile || threads > 1024 || threads % 64 || (threads & (threads - 1))) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
#define L(TWV)                                                                                                  \
  hipLaunchKernelGGL(refine_kernel<TWV>, dim3(grid), dim3(threads), 0, stream, fwd, bwd, a, b, t, d, x0, x1, r, ld, \
                     ncols, tol, max_steps, steps, ratio)
  DISPATCH_TILE(tile, L)
#undef L
  return hipGetLastError();
}

hipError_t launch_to_block(hipStream_t stream, const double* vectors, double* block, int64_t n, int64_t count,
                           int64_t ld) {
  dim3 grid((unsigned)((n + TR - 1) / TR), (unsigned)((count + TR - 1) / TR));
  hipLaunchKernelGGL(to_block_kernel, grid, dim3(TR, 8), 0, stream, vectors, block, n, count, ld);
  return hipGetLastError();
}

hipError_t launch_from_block(hipStream_t stream, const double* block, double* vectors, int64_t n, int64_t count,
                             int64_t ld) {
  dim3 grid((unsigned)((n + TR - 1) / TR), (unsigned)((count + TR - 1) / TR));
  hipLaunchKernelGGL(from_block_kernel, grid, dim3(TR, 8), 0, stream, block, vectors, n, count, ld);
  return hipGetLastError();
}

hipError_t launch_pair_update(hipStream_t stream, const double* y, double* out, int32_t n, int32_t ncols, int64_t rs,
                              int64_t cs, const double* base, const int32_t* plus, const int32_t* minus,
                              const double* alpha, const double* inverse_weight, int32_t* singular) {
  const int64_t total = (int64_t)n * ncols;
  const unsigned grid = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(pair_update_kernel, dim3(grid), dim3(256), 0, stream, y, out, n, ncols, rs, cs, base, plus,
                     minus, alpha, inverse_weight, singular);
  return hipGetLastError();
}

hipError_t launch_column_max(hipStream_t stream, const double* block, int32_t rows, int32_t ncols, int64_t rs,
                             int64_t cs, const double* weight, double* best, int32_t* best_row, int32_t* not_finite) {
  if (rs == 1) {
    hipLaunchKernelGGL(column_max_cols_kernel, dim3(ncols), dim3(256), 0, stream, block, rows, cs, weight, best,
                       best_row, not_finite);
    return hipGetLastError();
  }
  const unsigned grid = (unsigned)((ncols + 63) / 64);
  hipLaunchKernelGGL(column_max_kernel, dim3(grid), dim3(256), 0, stream, block, rows, ncols, rs, cs, weight, best,
                     best_row, not_finite);
  return hipGetLastError();
}

hipError_t launch_operator(hipStream_t stream, const CsrDev& m, const double* x, int64_t rs, int64_t cs,
                           const double* offset, const int32_t* zero, double* out, int64_t ors, int64_t ocs,
                           int32_t ncols) {
  const int64_t total = (int64_t)m.rows * ncols;
  const unsigned grid = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(operator_kernel, dim3(grid), dim3(256), 0, stream, m, x, rs, cs, offset, zero, out, ors, ocs,
                     ncols);
  return hipGetLastError();
}

size_t solve_lds_bytes(int32_t n, bool fp32) { return (size_t)n * (fp32 ? sizeof(float) : sizeof(double)); }

size_t refine_lds_bytes(int32_t n, int threads) { return (size_t)threads * sizeof(double) + (size_t)n * sizeof(float); }

hipError_t launch_solve_lds(hipStream_t stream, int threads, bool fp32, const SweepDev& fwd, const SweepDev& bwd,
                            int32_t n, const int32_t* rhs_row, const int32_t* sol_of, const ColumnsDev& rhs,
                            double* x, int64_t ldx, int32_t ncols) {
  if (threads < 64 || threads > 1024 || threads % 64) return hipErrorInvalidValue;
  const size_t bytes = solve_lds_bytes(n, fp32);
  if (fp32) {
    hipError_t e = hipFuncSetAttribute((const void*)solve_lds_kernel<float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(solve_lds_kernel<float>, dim3(ncols), dim3(threads), bytes, stream, fwd, bwd, n, rhs_row,
                       sol_of, rhs, x, ldx);
  } else {
    hipError_t e = hipFuncSetAttribute((const void*)solve_lds_kernel<double>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(solve_lds_kernel<double>, dim3(ncols), dim3(threads), bytes, stream, fwd, bwd, n, rhs_row,
                       sol_of, rhs, x, ldx);
  }
  return hipGetLastError();
}

hipError_t launch_refine_lds(hipStream_t stream, int threads, const SweepDev& fwd, const SweepDev& bwd,
                             const CsrDev& a, const int32_t* rhs_row, const int32_t* sol_of, const double* b,
                             double* x0, double* x1, double* r, int64_t ld, int32_t ncols, double tol,
                             int32_t max_steps, int32_t* steps, double* ratio) {
  if (threads < 64 || threads > 1024 || (threads & (threads - 1))) return hipErrorInvalidValue;
  const size_t bytes = refine_lds_bytes(a.rows, threads);
  hipError_t e = hipFuncSetAttribute((const void*)refine_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(refine_lds_kernel, dim3(ncols), dim3(threads), bytes, stream, fwd, bwd, a, rhs_row, sol_of, b, x0,
                     x1, r, ld, tol, max_steps, steps, ratio);
  return hipGetLastError();
}

}  // namespace ybus_hip
