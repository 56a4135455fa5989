// CPU oracle for the GPU factor-solve path: a restatement of ybus_sparse's compressed-matrix
// core, orderings, Cholesky and LU factorizations, CPU solves, residuals and refinement, and of
// ybus_powerflow's matrix builders for the synthetic workloads.
//
// TEST INFRASTRUCTURE ONLY. The product (ybus_amd, libybus_hip.so) never links or calls this
// code. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the
// checker and as the stand-in for ybus_sparse's host-side factor producer, whose factors the
// GPU path is defined to consume.
//
// Each function names the reference file:line it follows (paths relative to the reference
// repository root). Real (f64) arithmetic only; every floating-point operation is written in
// the reference's order and compiled with -ffp-contract=off, since Rust never fuses a*b+c.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

constexpr size_t NONE = SIZE_MAX;

// Error kinds, mirroring ybus_sparse's SparseError variants used on this path
// (ybus_sparse/src/sparse_error.rs).
enum class Err : int {
  NotSquare = 1,
  NotSymmetric = 2,
  NotPositiveDefinite = 3,
  SingularMatrix = 4,
  PatternMismatch = 5,
  ShapeMismatch = 6,
  OutOfBounds = 7,
  InvalidPermutation = 8,
  UnsupportedOrdering = 9,
  InvalidArgument = 10,
};

struct SparseError : std::runtime_error {
  Err kind;
  size_t pivot;
  SparseError(Err k, const std::string& what, size_t p = NONE)
      : std::runtime_error(what), kind(k), pivot(p) {}
};

// CompressedMatrix<f64, O> (ybus_sparse/src/compressed_matrix.rs:129-144): row-major = CSR,
// otherwise CSC. Minor indices strictly increase within each major index.
struct Compressed {
  size_t rows = 0, cols = 0;
  bool row_major = true;
  std::vector<size_t> ptr{0};
  std::vector<size_t> idx;
  std::vector<double> data;

  size_t major_dim() const { return row_major ? rows : cols; }
  size_t minor_dim() const { return row_major ? cols : rows; }
  size_t nnz() const { return idx.size(); }
  bool is_square() const { return rows == cols; }
};

struct Triplet {
  size_t row, col;
  double value;
};

enum class Ordering : int { Natural = 0, Rcm = 1, Amd = 2, Colamd = 3 };

// compressed_matrix.rs:254-321 (assemble / from_triplets)
Compressed from_triplets(size_t rows, size_t cols, bool row_major, const std::vector<Triplet>& t);
// compressed_matrix.rs:385-425 (reoriented: to_csr / to_csc)
Compressed reoriented(const Compressed& a, bool row_major);
// access.rs:175-219 (select), :229-250 (permute, permute_symmetric)
Compressed select(const Compressed& a, const std::vector<size_t>& rows, const std::vector<size_t>& cols);
Compressed permute_symmetric(const Compressed& a, const std::vector<size_t>& perm);
void check_permutation(const std::vector<size_t>& perm, size_t n);
// product.rs:17-46,93-105 (mul_vec through dot_kernel or scatter_kernel)
std::vector<double> mul_vec(const Compressed& a, const std::vector<double>& x);
// properties.rs:79-81 (is_hermitian, judged on values: stored zeros ignored)
bool is_symmetric(const Compressed& a);

// ordering.rs:68-130
struct Adjacency {
  std::vector<size_t> start, neighbours;
  static Adjacency of(const Compressed& a);
  size_t len() const { return start.size() - 1; }
  size_t degree(size_t i) const { return start[i + 1] - start[i]; }
};

// Mistakes the tests plant in AMD and COLAMD to show that the literal permutations they pin catch
// them (every other use passes none): degree lists or score buckets taken first in, first out;
// supervariables or supercolumns merged into the last of their kind rather than the first; no
// aggressive absorption.
enum Mutation : unsigned { FIFO_LISTS = 1, LAST_REPRESENTATIVE = 2, NO_AGGRESSIVE_ABSORPTION = 4 };

// ordering.rs:46-63 (CompressedMatrix::ordering)
std::vector<size_t> ordering(const Compressed& a, Ordering kind, unsigned mutations = 0);
std::vector<size_t> reverse_cuthill_mckee(const Adjacency& g);  // ordering.rs:135-234
std::vector<size_t> approximate_minimum_degree(const Adjacency& g,
                                               unsigned mutations = 0);  // ordering.rs:306-492
std::vector<size_t> column_approximate_minimum_degree(                  // colamd.rs:91-261
    size_t n_rows, size_t n_cols, const std::vector<std::pair<size_t, size_t>>& entries,
    unsigned mutations = 0);

// cholesky.rs:27-39 (SymbolicCholesky)
struct SymbolicCholesky {
  std::vector<size_t> perm;
  std::vector<size_t> parent;  // NONE for a root
  std::vector<size_t> postorder;
  std::vector<size_t> l_ptr, l_idx;
  std::vector<size_t> a_ptr, a_idx;
  size_t rows = 0;
  size_t dim() const { return perm.size(); }
  std::vector<size_t> column_counts() const;  // cholesky.rs:176-182
};

// cholesky.rs:321-328 (Cholesky<f64>)
struct Cholesky {
  Compressed l;          // CSR, each row ends with its diagonal
  Compressed l_adjoint;  // CSR of L^T, each row starts with its diagonal
  std::vector<size_t> perm, perm_inverse;
  size_t dim() const { return perm.size(); }
};

SymbolicCholesky analyze_cholesky(const Compressed& a, Ordering ordering);              // :50-58
SymbolicCholesky analyze_cholesky_permuted(const Compressed& a, std::vector<size_t> perm);  // :81-148
std::vector<size_t> postorder_of(const std::vector<size_t>& parent);                      // :527-557
Cholesky factor_cholesky(const SymbolicCholesky& s, const Compressed& a);                // :206-311
void forward_by_rows(const Compressed& l, double* y);                                     // :468-483
void back_by_rows(const Compressed& u, double* y);                                        // :487-501
void cholesky_substitute(const Cholesky& f, double* x, std::vector<double>& scratch);     // :570-579

// lu.rs:101-132 (Lu<f64>)
struct Lu {
  Compressed l;  // CSC, unit diagonal stored first in each column
  Compressed u;  // CSC, diagonal stored last in each column
  std::vector<size_t> row_perm, row_perm_inverse, col_perm;
  double pivot_threshold = 0.1;
  bool prefer_diagonal = true;
  std::vector<size_t> a_ptr, a_idx;  // the factored matrix's pattern, which refactor requires
  size_t dim() const { return row_perm.size(); }
};

struct SymbolicLu {
  std::vector<size_t> col_perm;
  bool prefer_diagonal = true;
  std::vector<size_t> a_ptr, a_idx;
};

SymbolicLu analyze_lu(const Compressed& a_csc, Ordering ordering);                       // lu.rs:142-153
SymbolicLu analyze_lu_with_permutation(const Compressed& a_csc, std::vector<size_t> q);  // lu.rs:163-178
Lu factor_lu(const SymbolicLu& s, const Compressed& a_csc, double pivot_threshold);      // lu.rs:226-245
Lu gilbert_peierls(const Compressed& a, const std::vector<size_t>& q, double tau,
                   bool prefer_diagonal);                                                 // lu.rs:268-414
void lu_substitute(const Lu& f, double* x, std::vector<double>& scratch);                 // lu.rs:770-785
// lu.rs:662-757 (Lu::refactor): new values on the same pattern, reusing the pivots; a pivot
// that fails the threshold factors afresh under the same column ordering (Repivoted)
enum class RefactorOutcome : int { Reused = 0, Repivoted = 1 };
RefactorOutcome lu_refactor(Lu& f, const Compressed& a_csc);
// The multiply-subtracts one refactor makes: the sum over columns k of the below-diagonal
// entries of L(:, j) for each j above the diagonal in U(:, k)
uint64_t lu_refactor_work(const Lu& f);
std::vector<double> lu_solve_transpose(const Lu& f, const double* b);                     // lu.rs:506-533

// factorization.rs:52-92 (solve_many / par_solve_many): rhs and solutions are `count`
// vectors of length n stored one after another.
void cholesky_solve_many(const Cholesky& f, const double* rhs, size_t count, double* out, int threads);
void lu_solve_many(const Lu& f, const double* rhs, size_t count, double* out, int threads);
void lu_solve_transpose_many(const Lu& f, const double* rhs, size_t count, double* out, int threads);

// accuracy.rs:57-79
double relative_residual(const Compressed& a, const double* x, const double* b, std::vector<double>* r);

struct NormwiseRefined {
  std::vector<double> x;
  size_t steps = 0;
  double relative_residual = 0.0;
};
// accuracy.rs:83-122, with the factorization's own solve
template <class Solve>
NormwiseRefined refine_normwise(const Compressed& a, const std::vector<double>& b, size_t max_steps,
                                double tolerance, Solve solve);

// multifrontal.rs:94-138 (fundamental supernodes) and :340-397, :620-660 (FrontTree)
struct Supernode {
  size_t start, end;
  std::vector<size_t> below;
  size_t parent;  // NONE for a root
};
std::vector<Supernode> supernodes(const SymbolicCholesky& s);
struct FrontTree {
  uint64_t fronts = 0, roots = 0, height = 0, work = 0, critical_work = 0, solve_work = 0,
           solve_critical_work = 0;
};
FrontTree front_tree(const std::vector<Supernode>& nodes);

// ---- refine_normwise body (a template, so it lives in the header) ----
template <class Solve>
NormwiseRefined refine_normwise(const Compressed& a, const std::vector<double>& b, size_t max_steps,
                                double tolerance, Solve solve) {
  const double half = 1.0 / (1.0 + 1.0);
  std::vector<double> x = solve(b);
  std::vector<double> r;
  double ratio = relative_residual(a, x.data(), b.data(), &r);
  NormwiseRefined best{x, 0, ratio};
  for (size_t step = 1; step <= max_steps; ++step) {
    if (ratio <= tolerance) break;
    std::vector<double> d = solve(r);
    for (size_t i = 0; i < x.size(); ++i) x[i] = x[i] + d[i];
    double previous = ratio;
    ratio = relative_residual(a, x.data(), b.data(), &r);
    if (ratio < best.relative_residual) best = NormwiseRefined{x, step, ratio};
    if (ratio != ratio || ratio > half * previous) break;
  }
  return best;
}

}  // namespace oracle
