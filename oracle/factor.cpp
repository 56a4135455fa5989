// Cholesky and LU factorizations and their CPU solves, in the oracle (test infrastructure
// only): a restatement of ybus_sparse/src/compressed_matrix/cholesky.rs, lu.rs,
// factorization.rs, accuracy.rs and multifrontal.rs (supernodes, front tree).
#include <algorithm>
#include <cmath>
#include <thread>

#include "oracle.hpp"

namespace oracle {

// cholesky.rs:176-182
std::vector<size_t> SymbolicCholesky::column_counts() const {
  std::vector<size_t> counts(dim(), 0);
  for (size_t j : l_idx) counts[j] += 1;
  return counts;
}

// cholesky.rs:50-58
SymbolicCholesky analyze_cholesky(const Compressed& a, Ordering kind) {
  if (kind == Ordering::Colamd)
    throw SparseError(Err::UnsupportedOrdering, "COLAMD orders columns only");
  std::vector<size_t> perm = ordering(a, kind);
  return analyze_cholesky_permuted(a, std::move(perm));
}

// cholesky.rs:81-148: Liu's elimination tree with path compression, then each row's pattern
// as its row subtree, sorted, ending with the diagonal.
SymbolicCholesky analyze_cholesky_permuted(const Compressed& a, std::vector<size_t> perm) {
  const size_t n = perm.size();
  std::vector<size_t> inverse(n, 0);
  for (size_t k = 0; k < n; ++k) inverse[perm[k]] = k;
  Adjacency graph = Adjacency::of(a);
  auto for_lower = [&](size_t i, auto f) {
    size_t old = perm[i];
    for (size_t k = graph.start[old]; k < graph.start[old + 1]; ++k) {
      size_t j = inverse[graph.neighbours[k]];
      if (j < i) f(j);
    }
  };
  std::vector<size_t> parent(n, NONE), ancestor(n, NONE);
  for (size_t i = 0; i < n; ++i) {
    for_lower(i, [&](size_t j) {
      while (j != i) {
        size_t next = ancestor[j];
        ancestor[j] = i;
        if (next == NONE) {
          parent[j] = i;
          break;
        }
        j = next;
      }
    });
  }
  SymbolicCholesky s;
  s.l_ptr.reserve(n + 1);
  s.l_ptr.push_back(0);
  std::vector<size_t> mark(n, NONE);
  for (size_t i = 0; i < n; ++i) {
    size_t start = s.l_idx.size();
    mark[i] = i;
    for_lower(i, [&](size_t j) {
      while (mark[j] != i) {
        mark[j] = i;
        s.l_idx.push_back(j);
        j = parent[j];
      }
    });
    std::sort(s.l_idx.begin() + start, s.l_idx.end());
    s.l_idx.push_back(i);
    s.l_ptr.push_back(s.l_idx.size());
  }
  s.postorder = postorder_of(parent);
  s.perm = std::move(perm);
  s.parent = std::move(parent);
  s.a_ptr = a.ptr;
  s.a_idx = a.idx;
  s.rows = a.rows;
  return s;
}

// cholesky.rs:527-557: children visited in increasing order, depth first; trees in increasing
// order of their roots
std::vector<size_t> postorder_of(const std::vector<size_t>& parent) {
  const size_t n = parent.size();
  std::vector<size_t> first_child(n, NONE), next_sibling(n, NONE);
  for (size_t node = n; node-- > 0;) {
    if (parent[node] != NONE) {
      next_sibling[node] = first_child[parent[node]];
      first_child[parent[node]] = node;
    }
  }
  std::vector<size_t> order;
  order.reserve(n);
  std::vector<size_t> stack;
  for (size_t root = 0; root < n; ++root) {
    if (parent[root] != NONE) continue;
    stack.push_back(root);
    while (!stack.empty()) {
      size_t node = stack.back();
      if (first_child[node] != NONE) {
        size_t child = first_child[node];
        first_child[node] = next_sibling[child];
        stack.push_back(child);
      } else {
        stack.pop_back();
        order.push_back(node);
      }
    }
  }
  return order;
}

// cholesky.rs:206-311: the up-looking factorization, storing the whole symbolic pattern
Cholesky factor_cholesky(const SymbolicCholesky& s, const Compressed& a) {
  const size_t n = s.dim();
  if (!a.row_major || a.rows != n || a.cols != n || a.ptr != s.a_ptr || a.idx != s.a_idx)
    throw SparseError(Err::PatternMismatch, "the matrix doesn't have the analyzed pattern");
  if (!is_symmetric(a)) throw SparseError(Err::NotSymmetric, "not symmetric");
  Compressed c = permute_symmetric(a, s.perm);

  std::vector<size_t> col_start(n + 1, 0);
  for (size_t j : s.l_idx) col_start[j + 1] += 1;
  for (size_t j = 0; j < n; ++j) col_start[j + 1] += col_start[j];
  for (size_t j = 0; j <= n; ++j) col_start[j] -= j;
  std::vector<size_t> col_fill = col_start;
  const size_t below = s.l_idx.size() - n;
  std::vector<size_t> col_rows(below, 0);
  std::vector<double> col_values(below, 0.0);

  std::vector<double> l_values(s.l_idx.size(), 0.0);
  std::vector<double> diagonal(n, 0.0), work(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    double remaining = 0.0;
    for (size_t p = c.ptr[i]; p < c.ptr[i + 1]; ++p) {
      size_t col = c.idx[p];
      if (col < i) {
        work[col] = c.data[p];
      } else {
        if (col == i) remaining = c.data[p];
        break;
      }
    }
    const size_t diagonal_slot = s.l_ptr[i + 1] - 1;
    for (size_t q = s.l_ptr[i]; q < diagonal_slot; ++q) {
      const size_t j = s.l_idx[q];
      const double l_ij = work[j] / diagonal[j];
      work[j] = 0.0;
      for (size_t r = col_start[j]; r < col_fill[j]; ++r) {
        size_t m = col_rows[r];
        work[m] = work[m] - col_values[r] * l_ij;
      }
      col_rows[col_fill[j]] = i;
      col_values[col_fill[j]] = l_ij;
      col_fill[j] += 1;
      l_values[q] = l_ij;
      remaining = remaining - (l_ij * l_ij + 0.0 * 0.0);
    }
    if (std::isnan(remaining) || remaining <= 0.0)
      throw SparseError(Err::NotPositiveDefinite, "not positive definite", s.perm[i]);
    diagonal[i] = std::sqrt(remaining);
    l_values[diagonal_slot] = diagonal[i];
  }

  Cholesky f;
  f.l.rows = f.l.cols = n;
  f.l.row_major = true;
  f.l.ptr = s.l_ptr;
  f.l.idx = s.l_idx;
  f.l.data = std::move(l_values);
  Compressed& lt = f.l_adjoint;
  lt.rows = lt.cols = n;
  lt.row_major = true;
  lt.ptr.clear();
  lt.ptr.push_back(0);
  lt.idx.reserve(s.l_idx.size());
  lt.data.reserve(s.l_idx.size());
  for (size_t j = 0; j < n; ++j) {
    lt.idx.push_back(j);
    lt.data.push_back(diagonal[j]);
    for (size_t r = col_start[j]; r < col_fill[j]; ++r) {
      lt.idx.push_back(col_rows[r]);
      lt.data.push_back(col_values[r]);
    }
    lt.ptr.push_back(lt.idx.size());
  }
  f.perm = s.perm;
  f.perm_inverse.assign(n, 0);
  for (size_t k = 0; k < n; ++k) f.perm_inverse[s.perm[k]] = k;
  return f;
}

// cholesky.rs:468-483
void forward_by_rows(const Compressed& l, double* y) {
  const size_t n = l.rows;
  for (size_t i = 0; i < n; ++i) {
    double remainder = y[i];
    double diagonal = 1.0;
    for (size_t p = l.ptr[i]; p < l.ptr[i + 1]; ++p) {
      size_t col = l.idx[p];
      if (col < i) {
        remainder = remainder - l.data[p] * y[col];
      } else {
        diagonal = l.data[p];
        break;
      }
    }
    y[i] = remainder / diagonal;
  }
}

// cholesky.rs:487-501
void back_by_rows(const Compressed& u, double* y) {
  const size_t n = u.rows;
  for (size_t i = n; i-- > 0;) {
    double remainder = y[i];
    double diagonal = 1.0;
    for (size_t p = u.ptr[i]; p < u.ptr[i + 1]; ++p) {
      size_t col = u.idx[p];
      if (col == i)
        diagonal = u.data[p];
      else if (col > i)
        remainder = remainder - u.data[p] * y[col];
    }
    y[i] = remainder / diagonal;
  }
}

// cholesky.rs:570-579
void cholesky_substitute(const Cholesky& f, double* x, std::vector<double>& scratch) {
  const size_t n = f.dim();
  scratch.resize(n);
  for (size_t k = 0; k < n; ++k) scratch[k] = x[f.perm[k]];
  forward_by_rows(f.l, scratch.data());
  back_by_rows(f.l_adjoint, scratch.data());
  for (size_t k = 0; k < n; ++k) x[f.perm[k]] = scratch[k];
}

// ---- LU ----

// lu.rs:142-153
SymbolicLu analyze_lu(const Compressed& a, Ordering kind) {
  if (a.row_major) throw SparseError(Err::InvalidArgument, "analyze_lu takes a CSC matrix");
  if (kind == Ordering::Colamd && !a.is_square()) throw SparseError(Err::NotSquare, "not square");
  SymbolicLu s;
  s.col_perm = ordering(a, kind);
  s.prefer_diagonal = kind != Ordering::Colamd;
  s.a_ptr = a.ptr;
  s.a_idx = a.idx;
  return s;
}

// lu.rs:163-178
SymbolicLu analyze_lu_with_permutation(const Compressed& a, std::vector<size_t> q) {
  if (!a.is_square()) throw SparseError(Err::NotSquare, "not square");
  check_permutation(q, a.cols);
  SymbolicLu s;
  s.col_perm = std::move(q);
  s.prefer_diagonal = true;
  s.a_ptr = a.ptr;
  s.a_idx = a.idx;
  return s;
}

// lu.rs:226-245
Lu factor_lu(const SymbolicLu& s, const Compressed& a, double pivot_threshold) {
  const size_t n = s.col_perm.size();
  if (a.row_major || a.rows != n || a.cols != n || a.ptr != s.a_ptr || a.idx != s.a_idx)
    throw SparseError(Err::PatternMismatch, "the matrix doesn't have the analyzed pattern");
  return gilbert_peierls(a, s.col_perm, pivot_threshold, s.prefer_diagonal);
}

// lu.rs:435-452
static void sort_columns(const std::vector<size_t>& ptr, std::vector<size_t>& idx, std::vector<double>& val) {
  std::vector<std::pair<size_t, double>> pairs;
  for (size_t c = 0; c + 1 < ptr.size(); ++c) {
    pairs.clear();
    for (size_t p = ptr[c]; p < ptr[c + 1]; ++p) pairs.emplace_back(idx[p], val[p]);
    std::sort(pairs.begin(), pairs.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t k = 0; k < pairs.size(); ++k) {
      idx[ptr[c] + k] = pairs[k].first;
      val[ptr[c] + k] = pairs[k].second;
    }
  }
}

// lu.rs:268-414: left-looking LU with threshold partial pivoting. The reach of each column
// through the graph of L by depth-first search gives a topological order; pivoted rows go to
// U, the largest of the rest (strict > in reverse postorder) is the candidate, and the
// diagonal of the ordered matrix is kept while it passes the threshold.
Lu gilbert_peierls(const Compressed& a, const std::vector<size_t>& q, double tau, bool prefer_diagonal) {
  if (!(tau > 0.0 && tau <= 1.0))
    throw SparseError(Err::InvalidArgument, "the pivot threshold must be in (0, 1]");
  const size_t n = q.size();
  std::vector<size_t> pinv(n, NONE);
  std::vector<size_t> l_ptr{0}, u_ptr{0}, l_idx, u_idx;
  std::vector<double> l_val, u_val;
  std::vector<double> x(n, 0.0);
  std::vector<size_t> visited(n, NONE);
  std::vector<size_t> postorder;
  postorder.reserve(n);
  std::vector<std::pair<size_t, size_t>> stack;

  for (size_t k = 0; k < n; ++k) {
    const size_t col = q[k];
    postorder.clear();
    auto first_child = [&](size_t row) { return pinv[row] == NONE ? (size_t)0 : l_ptr[pinv[row]] + 1; };
    for (size_t pa = a.ptr[col]; pa < a.ptr[col + 1]; ++pa) {
      const size_t start = a.idx[pa];
      if (visited[start] == k) continue;
      visited[start] = k;
      stack.emplace_back(start, first_child(start));
      while (!stack.empty()) {
        const size_t top = stack.size() - 1;
        size_t row = stack[top].first, next = stack[top].second;
        size_t end = pinv[row] == NONE ? 0 : l_ptr[pinv[row] + 1];
        size_t unvisited = NONE;
        while (next < end) {
          size_t child = l_idx[next];
          next += 1;
          if (visited[child] != k) {
            unvisited = child;
            break;
          }
        }
        stack[top].second = next;
        if (unvisited != NONE) {
          visited[unvisited] = k;
          stack.emplace_back(unvisited, first_child(unvisited));
        } else {
          stack.pop_back();
          postorder.push_back(row);
        }
      }
    }

    for (size_t pa = a.ptr[col]; pa < a.ptr[col + 1]; ++pa) x[a.idx[pa]] = a.data[pa];
    for (size_t r = postorder.size(); r-- > 0;) {
      const size_t row = postorder[r];
      if (pinv[row] != NONE) {
        const size_t s = pinv[row];
        const double x_row = x[row];
        for (size_t p = l_ptr[s] + 1; p < l_ptr[s + 1]; ++p) x[l_idx[p]] = x[l_idx[p]] - l_val[p] * x_row;
      }
    }

    double largest = 0.0;
    size_t pivot_row = NONE;
    for (size_t r = postorder.size(); r-- > 0;) {
      const size_t row = postorder[r];
      if (pinv[row] != NONE) {
        u_idx.push_back(pinv[row]);
        u_val.push_back(x[row]);
      } else {
        double magnitude = std::fabs(x[row]);
        if (magnitude > largest) {
          largest = magnitude;
          pivot_row = row;
        }
      }
    }
    if (pivot_row == NONE) throw SparseError(Err::SingularMatrix, "singular", col);
    if (prefer_diagonal && pinv[col] == NONE && visited[col] == k && std::fabs(x[col]) >= tau * largest)
      pivot_row = col;
    const double pivot = x[pivot_row];
    u_idx.push_back(k);
    u_val.push_back(pivot);
    u_ptr.push_back(u_idx.size());
    pinv[pivot_row] = k;

    l_idx.push_back(pivot_row);
    l_val.push_back(1.0);
    for (size_t r = postorder.size(); r-- > 0;) {
      const size_t row = postorder[r];
      if (pinv[row] == NONE) {
        l_idx.push_back(row);
        l_val.push_back(x[row] / pivot);
      }
      x[row] = 0.0;
    }
    l_ptr.push_back(l_idx.size());
  }

  Lu f;
  f.row_perm.assign(n, 0);
  for (size_t row = 0; row < n; ++row) f.row_perm[pinv[row]] = row;
  for (size_t& row : l_idx) row = pinv[row];
  sort_columns(l_ptr, l_idx, l_val);
  sort_columns(u_ptr, u_idx, u_val);
  f.l.rows = f.l.cols = n;
  f.l.row_major = false;
  f.l.ptr = std::move(l_ptr);
  f.l.idx = std::move(l_idx);
  f.l.data = std::move(l_val);
  f.u.rows = f.u.cols = n;
  f.u.row_major = false;
  f.u.ptr = std::move(u_ptr);
  f.u.idx = std::move(u_idx);
  f.u.data = std::move(u_val);
  f.row_perm_inverse.assign(n, 0);
  for (size_t k = 0; k < n; ++k) f.row_perm_inverse[f.row_perm[k]] = k;
  f.col_perm = q;
  f.pivot_threshold = tau;
  f.prefer_diagonal = prefer_diagonal;
  f.a_ptr = a.ptr;
  f.a_idx = a.idx;
  return f;
}

// lu.rs:662-757. Column k of the ordered matrix is scattered into x (indexed by pivot step),
// then for each row j above the diagonal in U(:, k), in increasing order (a topological order,
// L being lower triangular in step numbering), u_jk = x[j] and x[i] -= l_ij * u_jk over L(:, j)
// below its diagonal. The reference fuses nested column pairs (lu.rs:693-714), applying column
// j's term and then column j + 1's to each shared row: the same operations on each x[i] in the
// same order as the unfused loop written here, so the same values. The reused pivot x[k] must
// be nonzero and at least tau times the largest magnitude below it (a fold that starts from
// |x[k]| and takes b only when b > a); otherwise the matrix is factored afresh with the same
// column ordering. On an error f is unchanged.
RefactorOutcome lu_refactor(Lu& f, const Compressed& a) {
  const size_t n = f.dim();
  if (a.row_major || a.rows != n || a.cols != n || a.ptr != f.a_ptr || a.idx != f.a_idx)
    throw SparseError(Err::PatternMismatch, "the matrix's pattern differs from the factored one");
  const double tau = f.pivot_threshold;
  std::vector<double> l_val(f.l.nnz(), 0.0), u_val(f.u.nnz(), 0.0), x(n, 0.0);
  const std::vector<size_t>& step_of = f.row_perm_inverse;
  for (size_t k = 0; k < n; ++k) {
    const size_t col = f.col_perm[k];
    for (size_t p = a.ptr[col]; p < a.ptr[col + 1]; ++p) x[step_of[a.idx[p]]] = a.data[p];
    const size_t diagonal = f.u.ptr[k + 1] - 1;
    for (size_t p = f.u.ptr[k]; p < diagonal; ++p) {
      const size_t j = f.u.idx[p];
      const double u_jk = x[j];
      x[j] = 0.0;
      u_val[p] = u_jk;
      for (size_t q = f.l.ptr[j] + 1; q < f.l.ptr[j + 1]; ++q) x[f.l.idx[q]] = x[f.l.idx[q]] - l_val[q] * u_jk;
    }
    const double pivot = x[k];
    double largest = std::fabs(pivot);
    for (size_t q = f.l.ptr[k] + 1; q < f.l.ptr[k + 1]; ++q) {
      const double b = std::fabs(x[f.l.idx[q]]);
      if (b > largest) largest = b;
    }
    if (pivot == 0.0 || std::fabs(pivot) < tau * largest) {
      f = gilbert_peierls(a, f.col_perm, tau, f.prefer_diagonal);
      return RefactorOutcome::Repivoted;
    }
    u_val[diagonal] = pivot;
    x[k] = 0.0;
    l_val[f.l.ptr[k]] = 1.0;
    for (size_t q = f.l.ptr[k] + 1; q < f.l.ptr[k + 1]; ++q) {
      const size_t row = f.l.idx[q];
      l_val[q] = x[row] / pivot;
      x[row] = 0.0;
    }
  }
  f.l.data = std::move(l_val);
  f.u.data = std::move(u_val);
  return RefactorOutcome::Reused;
}

uint64_t lu_refactor_work(const Lu& f) {
  uint64_t work = 0;
  for (size_t k = 0; k < f.dim(); ++k)
    for (size_t p = f.u.ptr[k]; p + 1 < f.u.ptr[k + 1]; ++p) {
      const size_t j = f.u.idx[p];
      work += f.l.ptr[j + 1] - f.l.ptr[j] - 1;
    }
  return work;
}

// lu.rs:640-649 (back_substitute) inside lu.rs:770-785 (substitute)
void lu_substitute(const Lu& f, double* x, std::vector<double>& scratch) {
  const size_t n = f.dim();
  scratch.resize(n);
  double* y = scratch.data();
  for (size_t k = 0; k < n; ++k) y[k] = x[f.row_perm[k]];
  for (size_t j = 0; j < n; ++j) {
    const double y_j = y[j];
    for (size_t p = f.l.ptr[j] + 1; p < f.l.ptr[j + 1]; ++p) y[f.l.idx[p]] = y[f.l.idx[p]] - f.l.data[p] * y_j;
  }
  for (size_t j = n; j-- > 0;) {
    const size_t diagonal = f.u.ptr[j + 1] - 1;
    y[j] = y[j] / f.u.data[diagonal];
    const double y_j = y[j];
    for (size_t p = f.u.ptr[j]; p < diagonal; ++p) y[f.u.idx[p]] = y[f.u.idx[p]] - f.u.data[p] * y_j;
  }
  for (size_t k = 0; k < n; ++k) x[f.col_perm[k]] = y[k];
}

// lu.rs:506-533: U^T forward (columns of U read as rows, diagonal last), unit L^T back, then
// through P
std::vector<double> lu_solve_transpose(const Lu& f, const double* b) {
  const size_t n = f.dim();
  std::vector<double> y(n);
  for (size_t k = 0; k < n; ++k) y[k] = b[f.col_perm[k]];
  for (size_t j = 0; j < n; ++j) {
    const size_t diagonal = f.u.ptr[j + 1] - 1;
    double remainder = y[j];
    for (size_t p = f.u.ptr[j]; p < diagonal; ++p) remainder = remainder - f.u.data[p] * y[f.u.idx[p]];
    y[j] = remainder / f.u.data[diagonal];
  }
  for (size_t j = n; j-- > 0;) {
    double remainder = y[j];
    for (size_t p = f.l.ptr[j] + 1; p < f.l.ptr[j + 1]; ++p) remainder = remainder - f.l.data[p] * y[f.l.idx[p]];
    y[j] = remainder;
  }
  std::vector<double> x(n, 0.0);
  for (size_t k = 0; k < n; ++k) x[f.row_perm[k]] = y[k];
  return x;
}

// factorization.rs:52-92: each right-hand side solved with the same operations, serially or
// one per worker thread (par_solve_many), so both give the same bits.
template <class F>
static void solve_each(size_t count, size_t n, const double* rhs, double* out, int threads, F substitute) {
  auto work = [&](size_t from, size_t to) {
    std::vector<double> scratch;
    scratch.reserve(n);
    for (size_t c = from; c < to; ++c) {
      std::copy(rhs + c * n, rhs + (c + 1) * n, out + c * n);
      substitute(out + c * n, scratch);
    }
  };
  if (threads <= 1 || count < 2) {
    work(0, count);
    return;
  }
  std::vector<std::thread> pool;
  size_t t = std::min<size_t>((size_t)threads, count);
  for (size_t w = 0; w < t; ++w) {
    size_t from = count * w / t, to = count * (w + 1) / t;
    pool.emplace_back(work, from, to);
  }
  for (auto& th : pool) th.join();
}

void cholesky_solve_many(const Cholesky& f, const double* rhs, size_t count, double* out, int threads) {
  solve_each(count, f.dim(), rhs, out, threads,
             [&](double* x, std::vector<double>& s) { cholesky_substitute(f, x, s); });
}

void lu_solve_many(const Lu& f, const double* rhs, size_t count, double* out, int threads) {
  solve_each(count, f.dim(), rhs, out, threads, [&](double* x, std::vector<double>& s) { lu_substitute(f, x, s); });
}

void lu_solve_transpose_many(const Lu& f, const double* rhs, size_t count, double* out, int threads) {
  solve_each(count, f.dim(), rhs, out, threads, [&](double* x, std::vector<double>&) {
    std::vector<double> r = lu_solve_transpose(f, x);
    std::copy(r.begin(), r.end(), x);
  });
}

// accuracy.rs:57-79: ||b - A x||inf / ||b||inf, infinite when b is zero and the residual isn't
double relative_residual(const Compressed& a, const double* x, const double* b, std::vector<double>* r_out) {
  std::vector<double> xv(x, x + a.cols);
  std::vector<double> ax = mul_vec(a, xv);
  std::vector<double> r(a.rows);
  for (size_t i = 0; i < a.rows; ++i) r[i] = b[i] - ax[i];
  auto norm = [](const double* v, size_t len) {
    double m = 0.0;
    for (size_t i = 0; i < len; ++i)
      if (std::fabs(v[i]) > m) m = std::fabs(v[i]);
    return m;
  };
  double r_norm = norm(r.data(), r.size()), b_norm = norm(b, a.rows);
  double ratio = b_norm > 0.0 ? r_norm / b_norm : (r_norm > 0.0 ? INFINITY : 0.0);
  if (r_out) *r_out = std::move(r);
  return ratio;
}

// multifrontal.rs:94-138: fundamental supernodes; column j continues into j + 1 when j + 1 is
// its parent, j + 1 has one child, and column j holds one more entry than column j + 1.
std::vector<Supernode> supernodes(const SymbolicCholesky& s) {
  const size_t n = s.dim();
  std::vector<size_t> counts(n, 0);
  for (size_t j : s.l_idx) counts[j] += 1;
  std::vector<size_t> col_ptr(n + 1, 0);
  for (size_t j = 0; j < n; ++j) col_ptr[j + 1] = col_ptr[j] + counts[j];
  std::vector<size_t> fill = col_ptr;
  std::vector<size_t> col_rows(s.l_idx.size(), 0);
  for (size_t i = 0; i < n; ++i)
    for (size_t q = s.l_ptr[i]; q < s.l_ptr[i + 1]; ++q) col_rows[fill[s.l_idx[q]]++] = i;
  std::vector<size_t> children(n, 0);
  for (size_t p : s.parent)
    if (p != NONE) children[p] += 1;
  std::vector<Supernode> out;
  std::vector<size_t> node_of(n, 0);
  size_t start = 0;
  for (size_t j = 0; j < n; ++j) {
    bool continues = j + 1 < n && s.parent[j] == j + 1 && children[j + 1] == 1 && counts[j] == counts[j + 1] + 1;
    if (!continues) {
      size_t end = j + 1;
      Supernode sn{start, end, {}, NONE};
      for (size_t q = col_ptr[start]; q < col_ptr[start + 1]; ++q)
        if (col_rows[q] >= end) sn.below.push_back(col_rows[q]);
      for (size_t c = start; c < end; ++c) node_of[c] = out.size();
      out.push_back(std::move(sn));
      start = end;
    }
  }
  // multifrontal.rs:121-130: a supernode's parent is the supernode of its last column's parent
  for (auto& sn : out) sn.parent = s.parent[sn.end - 1] == NONE ? NONE : node_of[s.parent[sn.end - 1]];
  return out;
}

// multifrontal.rs:340-346 (front_work) and the FrontTree fold of SymbolicLdlt::front_tree
FrontTree front_tree(const std::vector<Supernode>& nodes) {
  const size_t n = nodes.size();
  struct Chain {
    uint64_t factor = 0, solve = 0, depth = 0;
  };
  std::vector<Chain> chain(n);
  FrontTree t;
  t.fronts = n;
  for (size_t k = 0; k < n; ++k) {
    const Supernode& node = nodes[k];
    uint64_t c = node.end - node.start;
    uint64_t m = c + node.below.size();
    uint64_t factor = 0;
    for (uint64_t j = 0; j < c; ++j) factor += (m - j - 1) * (m - j) / 2;
    uint64_t solve = c * (c > 0 ? c - 1 : 0) / 2 + c * node.below.size();
    Chain ours{chain[k].factor + factor, chain[k].solve + solve, chain[k].depth + 1};
    t.work += factor;
    t.solve_work += solve;
    if (node.parent != NONE) {
      Chain& theirs = chain[node.parent];
      theirs.factor = std::max(theirs.factor, ours.factor);
      theirs.solve = std::max(theirs.solve, ours.solve);
      theirs.depth = std::max(theirs.depth, ours.depth);
    } else {
      t.roots += 1;
      t.critical_work = std::max(t.critical_work, ours.factor);
      t.solve_critical_work = std::max(t.solve_critical_work, ours.solve);
      t.height = std::max(t.height, ours.depth);
    }
  }
  return t;
}

}  // namespace oracle
