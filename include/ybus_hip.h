/*
 * ybus_hip: the GPU factor-solve path of ybus, MI355X-native (HIP, gfx950), behind a C ABI.
 *
 * A Cholesky or LU factorization made on the host (by ybus_sparse) is uploaded once and solved
 * on the device for many right-hand sides: forward and back substitution, transposed, in FP64 or
 * as FP32 solves refined in FP64, for host vectors or for blocks kept on the device. Each entry
 * point replaces a method of the reference crate ybus_cuda; the comment above it names the
 * reference item (file:line in the reference repository) whose meaning, arguments and errors it
 * keeps. Conventions follow ybus_ffi's C API (ybus_ffi/src/status.rs:9-98,
 * ybus_ffi/src/lib.rs:54-80):
 *
 *  - every fallible call returns an int32 status, YBUS_HIP_OK (0) on success;
 *  - the message of the calling thread's last failure is read with ybus_hip_last_error_message;
 *  - handles are opaque, made by a _create/_upload call and released by their _free, and
 *    _free(NULL) does nothing;
 *  - what the reference treats as a caller bug (an assert/panic) is YBUS_HIP_INVALID_ARGUMENT.
 *
 * Index arrays are uint64 on the host (the reference's usize); the device holds int32 indices,
 * so a factor or matrix with 2^31 or more entries is YBUS_HIP_TOO_LARGE (sparse.rs:62-68).
 * Right-hand sides and solutions cross the boundary as `count` vectors of `dim` doubles stored
 * one after another, the memory of the reference's &[Vec<f64>].
 *
 * Threading: a context and everything made from it are used by one thread at a time; work on
 * a context runs in order on its one stream, and every call returns with that work finished.
 */
#ifndef YBUS_HIP_H
#define YBUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (ybus_ffi/src/status.rs:9-14; CudaError, ybus_cuda/src/lib.rs:88-124) ---- */
enum {
  YBUS_HIP_OK = 0,
  YBUS_HIP_NO_DEVICE = 1,        /* Gpu::new's Ok(None): no HIP runtime or no device */
  YBUS_HIP_INVALID_ARGUMENT = 2, /* a caller bug the reference asserts on */
  YBUS_HIP_SHAPE_MISMATCH = 3,   /* CudaError::ShapeMismatch */
  YBUS_HIP_TOO_LARGE = 4,        /* CudaError::TooLarge: beyond 32-bit device indices */
  YBUS_HIP_DEVICE_ERROR = 5,     /* CudaError::Driver: the HIP runtime failed */
  YBUS_HIP_OUT_OF_MEMORY = 6,    /* a device allocation failed */
  YBUS_HIP_INTERNAL = 7,         /* a bound inside a kernel was exceeded (a library bug) */
  YBUS_HIP_PATTERN_MISMATCH = 8  /* SparseError::PatternMismatch (lu.rs:662-667) */
};

/* Precision and direction of a device solve. */
enum {
  YBUS_HIP_FP64 = 0,         /* FP64 factor, FP64 arithmetic: the reference's solve, bit for bit */
  YBUS_HIP_FP32_REFINED = 1, /* FP32 factor and sweeps, FP64 residual and correction */
  YBUS_HIP_TRANSPOSE = 2     /* A^T x = b with the same factor, FP64 (lu.rs:502-533) */
};

typedef struct ybus_hip_ctx ybus_hip_ctx;           /* Gpu (lib.rs:166-173) */
typedef struct ybus_hip_factor ybus_hip_factor;     /* GpuFactor (factor.rs:40-62) */
typedef struct ybus_hip_block ybus_hip_block;       /* GpuBlock (pairs.rs:88-96) */
typedef struct ybus_hip_operator ybus_hip_operator; /* GpuOperator (pairs.rs:52-64) */
typedef struct ybus_hip_batch ybus_hip_batch;       /* many factors solved in one launch (new) */

/* Timings (factor.rs:21-31): seconds spent in solve_many's stages, and blocks solved. */
typedef struct {
  double upload;
  double solve;
  double download;
  uint64_t blocks;
} ybus_hip_timings;

/* ColumnMax (pairs.rs:39-50); row is -1 for none. */
typedef struct {
  int64_t row;
  double value;
  int32_t not_finite;
} ybus_hip_column_max;

/* Device kernel statistics of a factor since its last reset, from HIP events on the factor's
 * stream: launches and summed milliseconds of the solve kernels (one launch solves a block) and
 * of the adjoint's SDDMM kernel. */
typedef struct {
  uint64_t launches;
  double milliseconds;
  /* algorithmic bytes of those launches: per FP64 block of c columns, the factor's entries once
   * (8-byte value and 4-byte index each, S = entries of both sweeps) and row pointers, plus 16
   * bytes per column entry (right-hand side in, solution out): S * 12 + 8 (n + 1) + 16 n c */
  uint64_t bytes;
} ybus_hip_kernel_stats;

/* Outcome of a device refactorization (RefactorOutcome, lu.rs:88-99): the pivots were
 * reused and the factor now holds the new values, or a reused pivot failed the threshold and the
 * factor is unchanged (the reference would factor afresh here; on the device that is the
 * caller's host factorization and a new upload). */
enum {
  YBUS_HIP_REFACTOR_REUSED = 0,
  YBUS_HIP_REFACTOR_NEEDS_PIVOTING = 1
};

/* Statistics of a factor's refactorizations and value updates since its upload. */
typedef struct {
  uint64_t calls;           /* refactorizations run on the device */
  uint64_t reused;          /* of those, the ones whose pivots all passed */
  uint64_t updates;         /* value updates (ybus_hip_factor_update_lu_values) */
  double refactor_ms;       /* summed HIP-event time of the refactor kernel */
  double refresh_ms;        /* summed HIP-event time of rewriting the sweeps' values */
  uint64_t multiply_subtracts; /* of one refactorization */
} ybus_hip_refactor_stats;

/* The message of the calling thread's last failure (ybus_ffi/src/lib.rs:54-80): copies at most
 * len - 1 bytes and a NUL into buf, and returns the length the whole message needs, NUL
 * included. */
size_t ybus_hip_last_error_message(char *buf, size_t len);

/* ---- context: Gpu (ybus_cuda/src/lib.rs:175-219) ---- */

/* Gpu::new (lib.rs:179-207): the first device, or YBUS_HIP_NO_DEVICE without one. */
int32_t ybus_hip_ctx_create(ybus_hip_ctx **out);
void ybus_hip_ctx_free(ybus_hip_ctx *ctx);
/* Gpu::name (lib.rs:210-212): writes the device name like last_error_message; *needed gets
 * the length the whole name needs, NUL included. */
int32_t ybus_hip_ctx_name(ybus_hip_ctx *ctx, char *buf, size_t len, size_t *needed);
/* Gpu::memory (lib.rs:215-218): free and total device memory in bytes. */
int32_t ybus_hip_ctx_memory(ybus_hip_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- factor upload: GpuFactor::from_cholesky / from_lu (factor.rs:103-178) ---- */

/* GpuFactor::from_cholesky (factor.rs:103-119): P A P^T = L L^T with L in CSR form, each row
 * ending with its diagonal (Cholesky::l(), cholesky.rs:342-344), and P as perm: row i of
 * P A P^T is row perm[i] of A (Cholesky::permutation(), cholesky.rs:348-350). */
int32_t ybus_hip_factor_upload_cholesky(ybus_hip_ctx *ctx, uint64_t n, const uint64_t *l_ptr,
                                        const uint64_t *l_idx, const double *l_val,
                                        const uint64_t *perm, ybus_hip_factor **out);

/* GpuFactor::from_lu (factor.rs:123-133): P A Q = L U with L (unit diagonal stored first in
 * each column) and U (diagonal stored last) in CSC form (Lu::l(), Lu::u(), lu.rs:473-480),
 * row_perm as P (Lu::permutation(), lu.rs:483-485) and col_perm as Q
 * (Lu::column_permutation(), lu.rs:489-491). */
int32_t ybus_hip_factor_upload_lu(ybus_hip_ctx *ctx, uint64_t n, const uint64_t *l_ptr,
                                  const uint64_t *l_idx, const double *l_val,
                                  const uint64_t *u_ptr, const uint64_t *u_idx,
                                  const double *u_val, const uint64_t *row_perm,
                                  const uint64_t *col_perm, ybus_hip_factor **out);

/* Lu::refactor (lu.rs:662-757) on the device, for an uploaded real LU factor: the matrix A
 * (n x n, CSC: a_ptr n + 1, a_idx and a_val a_ptr[n] entries, rows increasing in each column)
 * is factored with the factor's own pivots and patterns, P A Q = L U, each column k by the
 * reference's loop (u_jk = x[j] for the rows j above the diagonal in U(:, k), in increasing
 * order, then x[i] -= l_ij * u_jk over L(:, j) below its diagonal; the pivot x[k] reused when it
 * is nonzero and at least pivot_threshold times the largest magnitude below it; l = x / pivot),
 * so the new values are the CPU refactor's bit for bit. *outcome is
 * YBUS_HIP_REFACTOR_REUSED (the factor's solves now use the new values) or
 * YBUS_HIP_REFACTOR_NEEDS_PIVOTING (the factor is unchanged). The first call fixes the matrix
 * pattern; a later call with another pattern, or a matrix with an entry outside the factor's
 * pattern, is YBUS_HIP_PATTERN_MISMATCH. pivot_threshold is the factorization's threshold
 * (LuOptions::pivot_threshold, lu.rs:248-255), in (0, 1]. A column of L with more than 512
 * entries below its diagonal, or a column of L and U together longer than 7,328, is beyond the
 * device refactor (YBUS_HIP_TOO_LARGE): refactor on the host and call
 * ybus_hip_factor_update_lu_values. A matrix attached with ybus_hip_factor_set_matrix that is
 * this one (CSC, the same pattern) takes the new values too, for FP32-refined solves; one
 * attached in another form keeps its values (attach the new matrix again). */
int32_t ybus_hip_factor_refactor_lu(ybus_hip_factor *f, uint64_t n, const uint64_t *a_ptr,
                                    const uint64_t *a_idx, const double *a_val,
                                    double pivot_threshold, int32_t *outcome);
/* The same for an uploaded complex LU factor (Lu<Complex64>: the reference's refactor is generic
 * over the element type): a_val holds (re, im) pairs, 2 * a_ptr[n] doubles. The arithmetic is
 * num-complex 0.4's (the product, then the difference; its quotient), and a pivot's magnitude is
 * its norm, hypot(re, im). */
int32_t ybus_hip_factor_refactor_lu_complex(ybus_hip_factor *f, uint64_t n, const uint64_t *a_ptr,
                                            const uint64_t *a_idx, const double *a_val,
                                            double pivot_threshold, int32_t *outcome);
/* New values for an uploaded real LU factor with the same pattern and permutations (a host
 * refactorization that reused its pivots): l_val and u_val in the uploaded CSC order. The
 * factor's sweeps are rewritten on the device, with no new analysis. The _complex form takes a
 * complex factor's values as (re, im) pairs. */
int32_t ybus_hip_factor_update_lu_values(ybus_hip_factor *f, const double *l_val, const double *u_val);
int32_t ybus_hip_factor_update_lu_values_complex(ybus_hip_factor *f, const double *l_val,
                                                 const double *u_val);
/* The refactorization statistics, and the L and U values of the factor (in the uploaded CSC
 * order: l_val L's entries, u_val U's, (re, im) pairs for a complex factor; either may be NULL). */
int32_t ybus_hip_factor_refactor_stats(const ybus_hip_factor *f, ybus_hip_refactor_stats *out);
int32_t ybus_hip_factor_lu_values(ybus_hip_factor *f, double *l_val, double *u_val);

/* Complex factors (new: the reference's GPU path uploads real factors only, factor.rs:103-133):
 * the arrays of ybus_hip_factor_upload_cholesky / _lu, with each value an interleaved (re, im)
 * pair of doubles (num_complex::Complex64's layout), so 2 * entries doubles per value array.
 * Cholesky<Complex64> is P A P^T = L L^H (cholesky.rs:206-311); the upload builds L^H as the
 * reference stores it, L's columns conjugated below the diagonal and each diagonal kept
 * (cholesky.rs:283-297). A complex factor takes the complex solves below and the accessors;
 * every other factor entry point returns YBUS_HIP_INVALID_ARGUMENT for it. */
int32_t ybus_hip_factor_upload_cholesky_complex(ybus_hip_ctx *ctx, uint64_t n, const uint64_t *l_ptr,
                                                const uint64_t *l_idx, const double *l_val,
                                                const uint64_t *perm, ybus_hip_factor **out);
int32_t ybus_hip_factor_upload_lu_complex(ybus_hip_ctx *ctx, uint64_t n, const uint64_t *l_ptr,
                                          const uint64_t *l_idx, const double *l_val,
                                          const uint64_t *u_ptr, const uint64_t *u_idx,
                                          const double *u_val, const uint64_t *row_perm,
                                          const uint64_t *col_perm, ybus_hip_factor **out);
/* 1 for a complex factor, else 0. */
int32_t ybus_hip_factor_is_complex(const ybus_hip_factor *f);

void ybus_hip_factor_free(ybus_hip_factor *f);

/* GpuFactor::dim (factor.rs:181-183). */
uint64_t ybus_hip_factor_dim(const ybus_hip_factor *f);
/* GpuFactor::block_width / set_block_width (factor.rs:186-199): set clamps to at least 1, and
 * at most 64 for a single-row factor. */
uint64_t ybus_hip_factor_block_width(const ybus_hip_factor *f);
void ybus_hip_factor_set_block_width(ybus_hip_factor *f, uint64_t width);
/* GpuFactor::timings / reset_timings (factor.rs:201-210). */
void ybus_hip_factor_timings(const ybus_hip_factor *f, ybus_hip_timings *out);
void ybus_hip_factor_reset_timings(ybus_hip_factor *f);

/* The matrix A itself, which FP32_REFINED needs for its FP64 residual b - A x: rows x rows in
 * CSR (row_major = 1, CsrMatrix) or CSC (row_major = 0, CscMatrix) form, as factored. No
 * counterpart in the reference, whose solves don't refine. */
int32_t ybus_hip_factor_set_matrix(ybus_hip_factor *f, int32_t row_major, const uint64_t *ptr,
                                   const uint64_t *idx, const double *val);

/* ---- host-vector solves ---- */

/* GpuFactor::solve_many (factor.rs:234-289): solves A x = b for `count` right-hand sides, in
 * blocks of block_width() columns, FP64. YBUS_HIP_SHAPE_MISMATCH is not possible here (the
 * caller's vectors are dim() long by construction of the ABI). */
int32_t ybus_hip_factor_solve_many(ybus_hip_factor *f, const double *rhs, uint64_t count,
                                   double *x);

/* A^T x = b for each right-hand side (Lu::solve_transpose, lu.rs:502-533; for a Cholesky factor
 * A^T = A), FP64, in blocks of block_width() columns. */
int32_t ybus_hip_factor_solve_transpose_many(ybus_hip_factor *f, const double *rhs,
                                             uint64_t count, double *x);

/* Complex FP64 solves of `count` right-hand sides of dim() interleaved (re, im) values each
 * (Factorization::solve_many, factorization.rs:29-92, over Complex64): forward_by_rows and
 * back_by_rows for a Cholesky factor (cholesky.rs:468-501), substitute for an LU factor
 * (lu.rs:770-785), in blocks of block_width() columns. Products and quotients are rounded as
 * num-complex 0.4 computes them. The kernel statistics count, per block of c columns,
 * S * 20 + 8 (n + 1) + 32 n c bytes. */
int32_t ybus_hip_factor_solve_many_complex(ybus_hip_factor *f, const double *rhs, uint64_t count,
                                           double *x);
/* A^T x = b over a complex factor: the plain transpose for an LU factor
 * (Lu::solve_transpose, lu.rs:502-533); for a Cholesky factor A^T = conj(A), solved as
 * conj(A^-1 conj(b)) (no counterpart in the reference). */
int32_t ybus_hip_factor_solve_transpose_many_complex(ybus_hip_factor *f, const double *rhs,
                                                     uint64_t count, double *x);

/* The stop rule of the factor's FP32-refined solves (new), each column's: YBUS_HIP_STOP_NORMWISE
 * (the default) is refine_normwise's (accuracy.rs:99-120): stop at `tolerance`, when a correction
 * fails to halve the ratio, on NaN, or after max_steps. YBUS_HIP_STOP_AT_FLOOR also stops as soon
 * as the ratio is at most 2 u || |A| |x| ||inf / ||b||inf (u = 2^-53), the rounding floor of an
 * FP64 residual, taken in the same residual pass: where `tolerance` is below the floor, the
 * correction after it would not halve the ratio, so the floor rule saves that correction and
 * keeps the same iterate. It applies to host, staged, device, complex and adjoint refinement. */
enum { YBUS_HIP_STOP_NORMWISE = 0, YBUS_HIP_STOP_AT_FLOOR = 1 };
int32_t ybus_hip_factor_set_refine_stop(ybus_hip_factor *f, int32_t rule);

/* The complex matrix A of a complex factor (new), which FP32-refined complex solves need for their
 * FP64 residual b - A x: rows x rows in CSR (row_major = 1) or CSC (row_major = 0) form, values as
 * interleaved (re, im) doubles (Complex64). The refined solves take the residual as
 * CsrMatrix::mul_vec does (product.rs:17-27), each row summed from zero in index order with
 * num-complex's product. The matrix keeps its values until it is attached again: after a device
 * refactor or value update, attach the new matrix before the next refined solve. */
int32_t ybus_hip_factor_set_matrix_complex(ybus_hip_factor *f, int32_t row_major, const uint64_t *ptr,
                                           const uint64_t *idx, const double *val);

/* Complex FP32 solves refined in FP64 (new; the reference's solve_refined_normwise over
 * Complex64, accuracy.rs:83-122): the factor's values rounded to FP32 and swept with FP32
 * arithmetic, the residual b - A x and the correction x + d in FP64, refine_normwise's stop rule
 * with |z| the hypotenuse of z's parts (num-complex's norm); steps and ratio (either may be
 * NULL) get each column's corrections and final relative residual. Right-hand sides and solutions
 * as in ybus_hip_factor_solve_many_complex. Needs ybus_hip_factor_set_matrix_complex. */
int32_t ybus_hip_factor_solve_many_refined_complex(ybus_hip_factor *f, const double *rhs, uint64_t count,
                                                   double tolerance, uint64_t max_steps, double *x,
                                                   uint64_t *steps, double *ratio);

/* FP32 solves refined in FP64 (new): each right-hand side follows refine_normwise
 * (accuracy.rs:83-122): solve, then while ||b - A x||inf / ||b||inf > tolerance and at most
 * max_steps times, solve for the residual and add the correction, stopping early when a step
 * fails to halve the ratio or gives NaN, and keeping the best iterate. steps[k] and
 * ratio[k] (either may be NULL) get each column's corrections and final relative residual.
 * Needs ybus_hip_factor_set_matrix. */
int32_t ybus_hip_factor_solve_many_refined(ybus_hip_factor *f, const double *rhs, uint64_t count,
                                           double tolerance, uint64_t max_steps, double *x,
                                           uint64_t *steps, double *ratio);

/* The adjoint of A x = b (new: the reference has no GPU gradient; its CPU counterpart is the
 * composition of Lu::solve_transpose, lu.rs:502-533, with an outer product). For `count` pairs
 * (x_bar_s, x_s), x_bar_s being the gradient of a scalar loss with respect to the solution x_s of
 * A x_s = b_s: b_bar_s = A^-T x_bar_s (the loss's gradient with respect to b_s), and
 * a_bar = -(sum over s of b_bar_s x_s^T) on A's stored pattern (its gradient with respect to A's
 * stored values), each entry summed over s in ascending order, so it is the same on every run.
 * mode YBUS_HIP_FP64: the transposed solve in FP64, the CPU's bits; YBUS_HIP_FP32_REFINED: FP32
 * transposed sweeps refined in FP64 against A^T (refine_normwise's rule, to `tolerance`, at most
 * max_steps corrections). Needs ybus_hip_factor_set_matrix; a_bar has one value per stored entry
 * of that matrix, in the order the caller stored them. */
int32_t ybus_hip_factor_adjoint(ybus_hip_factor *f, int32_t mode, double tolerance,
                                uint64_t max_steps, const double *x_bar, const double *x,
                                uint64_t count, double *b_bar, double *a_bar);

/* ---- vectors already in device memory (new) ----
 *
 * For callers that keep their vectors on the GPU (a PyTorch tensor, a previous kernel's output):
 * b, x, x_bar, b_bar hold `count` vectors of dim() doubles in device memory, vector c at
 * pointer + c * ld (ld >= dim()); a_bar is device memory for one value per stored entry of the
 * matrix given to ybus_hip_factor_set_matrix, in the caller's storage order. The caller's writes
 * to the inputs must be finished before the call (synchronize the stream that made them); the
 * call returns with its outputs written. Results are bit for bit those of the host-vector
 * calls. */
/* solve_many (mode YBUS_HIP_FP64), solve_transpose_many (YBUS_HIP_TRANSPOSE) or
 * solve_many_refined (YBUS_HIP_FP32_REFINED, to `tolerance` in at most max_steps corrections)
 * on device vectors. */
int32_t ybus_hip_factor_solve_device(ybus_hip_factor *f, int32_t mode, double tolerance,
                                     uint64_t max_steps, const double *b, uint64_t ld_b,
                                     uint64_t count, double *x, uint64_t ld_x);
/* A complex factor's solves on device vectors (new): b and x hold `count` vectors of dim() complex
 * values, interleaved (re, im) doubles, vector c at pointer + 2 c ld (ld counted in complex
 * values, ld >= dim()). mode YBUS_HIP_FP64 and YBUS_HIP_TRANSPOSE give the bits of
 * ybus_hip_factor_solve_many_complex / _solve_transpose_many_complex; YBUS_HIP_FP32_REFINED those of
 * ybus_hip_factor_solve_many_refined_complex (to `tolerance` in at most max_steps corrections). */
int32_t ybus_hip_factor_solve_device_complex(ybus_hip_factor *f, int32_t mode, double tolerance,
                                             uint64_t max_steps, const double *b, uint64_t ld_b,
                                             uint64_t count, double *x, uint64_t ld_x);
/* ybus_hip_factor_solve_device in YBUS_HIP_FP64 or YBUS_HIP_TRANSPOSE that allocates no device
 * memory, on this or any later call. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its workspace is a fixed device array of YBUS_HIP_SCRATCH_BYTES (64 MiB), which every device the
 * library runs on holds until the process ends; the vectors are solved in chunks of as many of
 * them as that array holds (at most block_width()), with the bits of ybus_hip_factor_solve_device.
 * A factor whose workspace for a single vector exceeds the array is YBUS_HIP_TOO_LARGE. Every call
 * on one device shares that device's array, so calls that overlap on one GPU (from several threads
 * or contexts) race; the array keeps the last call's contents, on which no later call's results
 * depend. The factor's own workspace and staged block are left as they are. */
#define YBUS_HIP_SCRATCH_BYTES (64u << 20)
int32_t ybus_hip_factor_solve_device_noalloc(ybus_hip_factor *f, int32_t mode, const double *b, uint64_t ld_b,
                                             uint64_t count, double *x, uint64_t ld_x);
/* ybus_hip_factor_adjoint on device vectors (x_bar, x and b_bar share the stride ld). */
int32_t ybus_hip_factor_adjoint_device(ybus_hip_factor *f, int32_t mode, double tolerance,
                                       uint64_t max_steps, const double *x_bar, const double *x,
                                       uint64_t ld, uint64_t count, double *b_bar,
                                       double *a_bar);

/* ---- device-resident blocks ---- */

/* Copies `count` right-hand sides to the device, into the factor's staged block (at most
 * block_width() of them), without solving: the device-resident solves below start from it. */
int32_t ybus_hip_factor_stage(ybus_hip_factor *f, const double *rhs, uint64_t count);

/* Stages k pairs (as pair_solves takes them) instead: solve_staged in YBUS_HIP_FP64 then builds
 * each right-hand side e_plus - e_minus on the device, as pair_solves does. */
int32_t ybus_hip_factor_stage_pairs(ybus_hip_factor *f, const int64_t *plus, const int64_t *minus,
                                    uint64_t k);

/* Solves the staged block on the device in `mode` (YBUS_HIP_FP64, _FP32_REFINED or _TRANSPOSE;
 * tolerance and max_steps are read by FP32_REFINED only), leaving the solutions there as a
 * block, which borrows the factor until freed. */
int32_t ybus_hip_factor_solve_staged(ybus_hip_factor *f, int32_t mode, double tolerance,
                                     uint64_t max_steps, ybus_hip_block **out);

/* GpuFactor::pair_solves (pairs.rs:153-202): solves A y = e_plus - e_minus for each of k pairs
 * (k <= block_width()), the right-hand sides built on the device; -1 means no entry.
 * INVALID_ARGUMENT for more pairs than the width, an unknown >= dim(), or plus == minus. */
int32_t ybus_hip_factor_pair_solves(ybus_hip_factor *f, const int64_t *plus, const int64_t *minus,
                                    uint64_t k, ybus_hip_block **out);

/* GpuFactor::pair_updates (pairs.rs:209-253): pair_solves, then each column's rank-one update of
 * base (dim() values): x = base + alpha y, w = inverse_weight - c^T y, z = c^T x / w,
 * out = x + z y, with w == 0 flagged singular and given x. */
int32_t ybus_hip_factor_pair_updates(ybus_hip_factor *f, const double *base, const int64_t *plus,
                                     const int64_t *minus, const double *alpha,
                                     const double *inverse_weight, uint64_t k,
                                     ybus_hip_block **out);

/* Outages of sets of branches (new: the reference takes a set out on the CPU, dc/outage.rs:126-241,
 * and its GPU path only single branches): set s holds the pairs set_ptr[s] .. set_ptr[s + 1] - 1
 * (at most 64) of plus/minus/alpha/inverse_weight, each a branch's pair with its alpha_k and
 * 1 / b_k. All the sets' pairs are solved as one block (Y = A^-1 C, at most block_width()
 * pairs), then per set: x = base + sum alpha_k Y_k (in order), W = D^-1 - C^T Y,
 * z = W^-1 C^T x, and the set's column x + sum z_j Y_j (in order). W and C^T x come from the
 * device; z is solved on the host by Gaussian elimination with partial pivoting (the reference
 * factors W with its sparse LU, so z agrees to rounding, not bit for bit). A set whose W is
 * singular gets x and is flagged in ybus_hip_block_singular. Returns a block of one column per
 * set, which borrows the factor like pair_updates's. */
int32_t ybus_hip_factor_outage_updates(ybus_hip_factor *f, const double *base, const uint64_t *set_ptr,
                                       uint64_t sets, const int64_t *plus, const int64_t *minus,
                                       const double *alpha, const double *inverse_weight,
                                       ybus_hip_block **out);

void ybus_hip_block_free(ybus_hip_block *b);
/* GpuBlock::columns (pairs.rs:258-260). */
uint64_t ybus_hip_block_columns(const ybus_hip_block *b);
/* GpuBlock::singular (pairs.rs:264-266): one byte per column. */
int32_t ybus_hip_block_singular(const ybus_hip_block *b, uint8_t *out);
/* Each column's refinement steps and relative residual (FP32_REFINED blocks; zeros otherwise). */
int32_t ybus_hip_block_refinement(const ybus_hip_block *b, uint64_t *steps, double *ratio);
/* GpuBlock::solutions (pairs.rs:269-290): columns() vectors of dim() values, by unknown. */
int32_t ybus_hip_block_solutions(ybus_hip_block *b, double *out);
/* GpuBlock::maxima (pairs.rs:293-308): each column's largest |value|, its unknown (the
 * smallest unknown on a tie) and whether any value isn't finite. */
int32_t ybus_hip_block_maxima(ybus_hip_block *b, ybus_hip_column_max *out);
/* GpuBlock::product (pairs.rs:373-392): the operator's M theta + offset for every column, with
 * column c's entry at zero[c] (-1 for none) set to zero: columns() vectors of rows() values. */
int32_t ybus_hip_block_product(ybus_hip_block *b, ybus_hip_operator *op, const int64_t *zero,
                               double *out);
/* GpuBlock::product_maxima (pairs.rs:396-413): product reduced to |value| * weight maxima. */
int32_t ybus_hip_block_product_maxima(ybus_hip_block *b, ybus_hip_operator *op,
                                      const int64_t *zero, ybus_hip_column_max *out);
/* Which kernels solve a block of `columns` right-hand sides (pairs != 0: a pair block) with this
 * factor:
 * YBUS_HIP_PATH_COLUMNS, one workgroup per right-hand side with its solution in LDS and the block
 * stored a column (a right-hand side) at a time, for factors whose solution fits in LDS and
 * blocks of at most twice as many columns as the device has compute units; or
 * YBUS_HIP_PATH_WIDE, tiles of columns in global memory with a row-interleaved block. Setting the
 * environment variable YBUS_HIP_PATH to "wide" or "columns" before an upload makes that factor
 * always take the wide path, or the column path wherever the factor fits. */
enum { YBUS_HIP_PATH_COLUMNS = 0, YBUS_HIP_PATH_WIDE = 1 };
int32_t ybus_hip_factor_path(const ybus_hip_factor *f, uint64_t columns, int32_t pairs);

/* The kernel configuration a block width takes: path (YBUS_HIP_PATH_*), workgroup size, column
 * tile width (1 on the column path), the time of the synthetic block it was chosen on, and
 * whether the wide path sweeps the factor's partitioned form (1: a Cholesky factor's independent
 * elimination subtrees side by side, then the top of the tree; 0: one schedule), and for a
 * partitioned block of vectors the top of the tree's own tile width and workgroup size (the
 * parts' own where the top takes the same), the tile width and workgroup size of the parts
 * swept with their values in LDS (0: the parts over the block in global memory), and whether a
 * one-schedule wide block keeps its last positions' values in an LDS ring (window 1). */
typedef struct {
  int32_t path;
  int32_t threads;
  int32_t tile;
  double milliseconds;
  int32_t parts;
  int32_t top_tile;
  int32_t top_threads;
  int32_t parts_lds_tile;
  int32_t parts_lds_threads;
  int32_t window;
} ybus_hip_tuning;

/* Autotuner (new): times every kernel configuration a block of `columns` right-hand sides can
 * take with this factor (the column path at 256 and 512 threads where the factor fits; the wide
 * path at 256, 512 and 1024 threads with column tiles of 1 to 64, over the factor's single and,
 * where it has one, partitioned schedule), each on a synthetic FP64 block
 * of that width, then, when a partitioned block of vectors wins, the same tile widths and sizes
 * for its top-of-tree kernel alone, then its parts swept in LDS (tiles of 1 to 16 columns at 256,
 * 512 and 1024 threads, where they fit) or, when one schedule wins, the same solve with its last positions in LDS at
 * each tile width and size, and keeps the fastest for this factor's later blocks of that width (up to the
 * next power of two) and kind: pair blocks (pairs != 0; pair_solves, pair_updates, staged pairs)
 * or blocks of vectors (solve_many, staged vectors, refinement). A pair block is timed on the
 * caller's own `columns` pairs when plus and minus are given (the pairs' reach sets the forward
 * sweep's work), else on synthetic ones. `out` (may be NULL) gets the choice. YBUS_HIP_PATH,
 * YBUS_HIP_THREADS and YBUS_HIP_TILE still override it. Whatever was staged is dropped. */
int32_t ybus_hip_factor_autotune(ybus_hip_factor *f, uint64_t columns, int32_t pairs,
                                 const int64_t *plus, const int64_t *minus, ybus_hip_tuning *out);

/* Kernel statistics of this factor's device solves (HIP events), and their reset. */
void ybus_hip_factor_kernel_stats(const ybus_hip_factor *f, ybus_hip_kernel_stats *out);
void ybus_hip_factor_reset_kernel_stats(ybus_hip_factor *f);

/* ---- batches of factors (new: the reference solves one factor per call, factor.rs:234-289) ----
 *
 * Many uploaded factors, each with its own right-hand sides, solved in one persistent launch per
 * class of factors (factors grouped by how many of their solution vectors fit a compute unit's
 * LDS). Workgroups take (factor, column) items from a device queue, longest first, so small and
 * large systems share the GPU without a launch per factor. The launches are captured once per
 * staging and mode in a HIP graph and replayed by each solve. A batch borrows its
 * factors: they must outlive it, and each must take the column path (dim() small enough for one
 * solution vector to fit in LDS; larger factors are rejected with YBUS_HIP_INVALID_ARGUMENT). */
int32_t ybus_hip_batch_create(ybus_hip_ctx *ctx, ybus_hip_batch **out);
void ybus_hip_batch_free(ybus_hip_batch *b);
int32_t ybus_hip_batch_add(ybus_hip_batch *b, ybus_hip_factor *f);
uint64_t ybus_hip_batch_factors(const ybus_hip_batch *b);
/* The stop rule of the batch's FP32-refined solves (ybus_hip_factor_set_refine_stop's rules). */
int32_t ybus_hip_batch_set_refine_stop(ybus_hip_batch *b, int32_t rule);
/* Copies the right-hand sides to the device: counts[i] vectors of factor i's dim(), for each
 * factor in the order added, all one after another. */
int32_t ybus_hip_batch_stage(ybus_hip_batch *b, const double *rhs, const uint64_t *counts);
/* Solves what is staged, in FP64 (YBUS_HIP_FP64: the CPU's bits) or as FP32 sweeps refined in
 * FP64 (YBUS_HIP_FP32_REFINED: each column to `tolerance` by refine_normwise's rule, at most
 * max_steps corrections; every factor needs ybus_hip_factor_set_matrix). The solutions stay on the
 * device until ybus_hip_batch_solutions. */
int32_t ybus_hip_batch_solve_staged(ybus_hip_batch *b, int32_t mode, double tolerance,
                                    uint64_t max_steps);
/* The staged right-hand sides' solutions, in the same arrangement as the right-hand sides. */
int32_t ybus_hip_batch_solutions(ybus_hip_batch *b, double *x);
/* Each staged column's corrections and final relative residual after an FP32-refined solve
 * (either may be NULL), in the arrangement of the staged right-hand sides. */
int32_t ybus_hip_batch_refinement(ybus_hip_batch *b, uint64_t *steps, double *ratio);
/* stage + solve_staged (FP64) + solutions. */
int32_t ybus_hip_batch_solve(ybus_hip_batch *b, const double *rhs, const uint64_t *counts,
                             double *x);
/* HIP-event timings of the batch's solves (one launch = one replay of its graph). */
void ybus_hip_batch_kernel_stats(const ybus_hip_batch *b, ybus_hip_kernel_stats *out);
void ybus_hip_batch_reset_kernel_stats(ybus_hip_batch *b);

/* ---- host-side symbolic analysis (no device needed; new) ----
 * The level schedule the device sweeps follow: level[i] of each factor row is one more than the
 * deepest row it depends on (0 for none). For a Cholesky factor (arrays as in upload_cholesky)
 * the forward sweep is L and the backward sweep L^T; supernode_start gets the first column of
 * each fundamental supernode of L's elimination tree (multifrontal.rs:94-138) and n, at most
 * n + 1 values, and *supernodes their count. For an LU factor (arrays as in upload_lu) the
 * sweeps are L, U, U^T and L^T (the transposed solve). Any output may be NULL. */
int32_t ybus_hip_analyze_cholesky(uint64_t n, const uint64_t *l_ptr, const uint64_t *l_idx,
                                  int32_t *forward_level, int32_t *backward_level,
                                  uint64_t *supernode_start, uint64_t *supernodes);
int32_t ybus_hip_analyze_lu(uint64_t n, const uint64_t *l_ptr, const uint64_t *l_idx,
                            const uint64_t *u_ptr, const uint64_t *u_idx, int32_t *forward_level,
                            int32_t *backward_level, int32_t *transpose_forward_level,
                            int32_t *transpose_backward_level);

/* ---- operator: GpuFactor::operator, GpuOperator (pairs.rs:66-149) ---- */

/* GpuFactor::operator (pairs.rs:112-149): a rows x dim() CSR matrix applied to solutions. */
int32_t ybus_hip_factor_operator(ybus_hip_factor *f, uint64_t rows, const uint64_t *ptr,
                                 const uint64_t *idx, const double *val, ybus_hip_operator **out);
void ybus_hip_operator_free(ybus_hip_operator *op);
uint64_t ybus_hip_operator_rows(const ybus_hip_operator *op);
/* GpuOperator::set_offset / set_weights (pairs.rs:73-85): rows() values each. */
int32_t ybus_hip_operator_set_offset(ybus_hip_operator *op, const double *offset);
int32_t ybus_hip_operator_set_weights(ybus_hip_operator *op, const double *weights);

#ifdef __cplusplus
}
#endif

#endif /* YBUS_HIP_H */
