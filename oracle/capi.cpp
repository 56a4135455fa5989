// C entry points of the oracle library (liboracle.so), for the Python test harness. Test
// infrastructure only. Handles are opaque; every call reports failure through its status and
// yo_last_error().
#include <cstring>
#include <memory>
#include <string>

#include "oracle.hpp"
#include "workloads.hpp"

using namespace oracle;

namespace {
thread_local std::string last_error;

struct Obj {
  virtual ~Obj() = default;
};
struct MatrixObj : Obj {
  Compressed m;
};
struct CholObj : Obj {
  SymbolicCholesky s;
  Cholesky f;
};
struct LuObj : Obj {
  Lu f;
};
struct SuperObj : Obj {
  std::vector<Supernode> nodes;
};
struct GridObj : Obj {
  Grid g;
};
struct DcObj : Obj {
  DcMatrices dc;
};
struct AcObj : Obj {
  AcJacobian ac;
};

Compressed view(uint64_t rows, uint64_t cols, int row_major, const uint64_t* ptr, const uint64_t* idx,
                const double* data) {
  Compressed m;
  m.rows = rows;
  m.cols = cols;
  m.row_major = row_major != 0;
  size_t major = m.row_major ? rows : cols;
  m.ptr.assign(ptr, ptr + major + 1);
  size_t nnz = m.ptr[major];
  m.idx.assign(idx, idx + nnz);
  m.data.assign(data, data + nnz);
  return m;
}

template <class F>
int guard(F f) {
  try {
    f();
    return 0;
  } catch (const SparseError& e) {
    last_error = e.what();
    return (int)e.kind;
  } catch (const std::exception& e) {
    last_error = e.what();
    return 99;
  }
}

void copy_matrix(const Compressed& m, uint64_t* ptr, uint64_t* idx, double* data) {
  for (size_t i = 0; i < m.ptr.size(); ++i) ptr[i] = m.ptr[i];
  for (size_t i = 0; i < m.idx.size(); ++i) idx[i] = m.idx[i];
  std::memcpy(data, m.data.data(), m.data.size() * sizeof(double));
}
}  // namespace

#define MAT_ARGS uint64_t rows, uint64_t cols, int row_major, const uint64_t *ptr, const uint64_t *idx, const double *data
#define MAT view(rows, cols, row_major, ptr, idx, data)

extern "C" {

size_t yo_last_error(char* buf, size_t len) {
  if (buf && len) {
    size_t n = std::min(len - 1, last_error.size());
    std::memcpy(buf, last_error.data(), n);
    buf[n] = 0;
  }
  return last_error.size() + 1;
}

void yo_free(void* h) { delete static_cast<Obj*>(h); }

// ---- matrices ----

int yo_from_triplets(uint64_t rows, uint64_t cols, int row_major, uint64_t count, const uint64_t* r,
                     const uint64_t* c, const double* v, void** out) {
  return guard([&] {
    std::vector<Triplet> t(count);
    for (size_t k = 0; k < count; ++k) t[k] = Triplet{r[k], c[k], v[k]};
    auto* o = new MatrixObj;
    o->m = from_triplets(rows, cols, row_major != 0, t);
    *out = o;
  });
}

void yo_matrix_sizes(void* h, uint64_t* out) {
  const Compressed& m = static_cast<MatrixObj*>(h)->m;
  out[0] = m.rows;
  out[1] = m.cols;
  out[2] = m.row_major;
  out[3] = m.nnz();
}

void yo_matrix_copy(void* h, uint64_t* ptr, uint64_t* idx, double* data) {
  copy_matrix(static_cast<MatrixObj*>(h)->m, ptr, idx, data);
}

int yo_reoriented(MAT_ARGS, int new_row_major, void** out) {
  return guard([&] {
    auto* o = new MatrixObj;
    o->m = reoriented(MAT, new_row_major != 0);
    *out = o;
  });
}

int yo_select(MAT_ARGS, uint64_t nr, const uint64_t* rsel, uint64_t nc, const uint64_t* csel, void** out) {
  return guard([&] {
    auto* o = new MatrixObj;
    o->m = select(MAT, std::vector<size_t>(rsel, rsel + nr), std::vector<size_t>(csel, csel + nc));
    *out = o;
  });
}

int yo_mul_vec(MAT_ARGS, const double* x, double* y) {
  return guard([&] {
    Compressed m = MAT;
    std::vector<double> r = mul_vec(m, std::vector<double>(x, x + m.cols));
    std::memcpy(y, r.data(), r.size() * sizeof(double));
  });
}

int yo_relative_residual(MAT_ARGS, const double* x, const double* b, double* ratio) {
  return guard([&] { *ratio = relative_residual(MAT, x, b, nullptr); });
}

int yo_ordering(MAT_ARGS, int kind, uint64_t* perm) {
  return guard([&] {
    std::vector<size_t> p = ordering(MAT, (Ordering)kind);
    for (size_t i = 0; i < p.size(); ++i) perm[i] = p[i];
  });
}

// The ordering with planted mistakes (Mutation flags), for the tests that show its literals catch them
int yo_ordering_mutated(MAT_ARGS, int kind, unsigned mutations, uint64_t* perm) {
  return guard([&] {
    std::vector<size_t> p = ordering(MAT, (Ordering)kind, mutations);
    for (size_t i = 0; i < p.size(); ++i) perm[i] = p[i];
  });
}

// ---- Cholesky ----

int yo_cholesky(MAT_ARGS, int kind, const uint64_t* perm, int symbolic_only, void** out, uint64_t* pivot) {
  *pivot = UINT64_MAX;
  std::unique_ptr<CholObj> o(new CholObj);
  int status = guard([&] {
    Compressed a = MAT;
    if (!a.row_major) throw SparseError(Err::InvalidArgument, "Cholesky takes a CSR matrix");
    if (perm) {
      if (!a.is_square()) throw SparseError(Err::NotSquare, "not square");
      std::vector<size_t> p(perm, perm + a.rows);
      check_permutation(p, a.rows);
      o->s = analyze_cholesky_permuted(a, std::move(p));
    } else {
      o->s = analyze_cholesky(a, (Ordering)kind);
    }
    if (!symbolic_only) {
      try {
        o->f = factor_cholesky(o->s, a);
      } catch (const SparseError& e) {
        *pivot = e.pivot;
        throw;
      }
    }
  });
  if (status == 0) *out = o.release();
  return status;
}

void yo_chol_sizes(void* h, uint64_t* out) {
  auto* o = static_cast<CholObj*>(h);
  out[0] = o->s.dim();
  out[1] = o->s.l_idx.size();
}

void yo_chol_copy_symbolic(void* h, uint64_t* perm, int64_t* parent, uint64_t* postorder, uint64_t* l_ptr,
                           uint64_t* l_idx, uint64_t* counts) {
  auto* o = static_cast<CholObj*>(h);
  const SymbolicCholesky& s = o->s;
  for (size_t i = 0; i < s.dim(); ++i) {
    perm[i] = s.perm[i];
    parent[i] = s.parent[i] == NONE ? -1 : (int64_t)s.parent[i];
    postorder[i] = s.postorder[i];
  }
  for (size_t i = 0; i < s.l_ptr.size(); ++i) l_ptr[i] = s.l_ptr[i];
  for (size_t i = 0; i < s.l_idx.size(); ++i) l_idx[i] = s.l_idx[i];
  std::vector<size_t> c = s.column_counts();
  for (size_t i = 0; i < c.size(); ++i) counts[i] = c[i];
}

void yo_chol_copy_factor(void* h, uint64_t* l_ptr, uint64_t* l_idx, double* l_val, uint64_t* lt_ptr,
                         uint64_t* lt_idx, double* lt_val) {
  auto* o = static_cast<CholObj*>(h);
  copy_matrix(o->f.l, l_ptr, l_idx, l_val);
  copy_matrix(o->f.l_adjoint, lt_ptr, lt_idx, lt_val);
}

int yo_chol_solve_many(void* h, uint64_t count, const double* rhs, double* out, int threads) {
  return guard([&] { cholesky_solve_many(static_cast<CholObj*>(h)->f, rhs, count, out, threads); });
}

int yo_chol_refine(void* h, MAT_ARGS, const double* b, uint64_t max_steps, double tol, double* x, uint64_t* steps,
                   double* ratio) {
  return guard([&] {
    const Cholesky& f = static_cast<CholObj*>(h)->f;
    Compressed a = MAT;
    std::vector<double> bv(b, b + a.rows);
    auto solve = [&](const std::vector<double>& rhs) {
      std::vector<double> out(rhs), scratch;
      cholesky_substitute(f, out.data(), scratch);
      return out;
    };
    NormwiseRefined r = refine_normwise(a, bv, max_steps, tol, solve);
    std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
    *steps = r.steps;
    *ratio = r.relative_residual;
  });
}

int yo_chol_supernodes(void* h, void** out) {
  return guard([&] {
    auto* o = new SuperObj;
    o->nodes = supernodes(static_cast<CholObj*>(h)->s);
    *out = o;
  });
}

void yo_super_sizes(void* h, uint64_t* out) {
  auto* o = static_cast<SuperObj*>(h);
  out[0] = o->nodes.size();
  size_t total = 0;
  for (auto& s : o->nodes) total += s.below.size();
  out[1] = total;
}

void yo_super_copy(void* h, uint64_t* start, uint64_t* end, int64_t* parent, uint64_t* below_ptr, uint64_t* below) {
  auto* o = static_cast<SuperObj*>(h);
  below_ptr[0] = 0;
  size_t at = 0;
  for (size_t k = 0; k < o->nodes.size(); ++k) {
    const Supernode& s = o->nodes[k];
    start[k] = s.start;
    end[k] = s.end;
    parent[k] = s.parent == NONE ? -1 : (int64_t)s.parent;
    for (size_t r : s.below) below[at++] = r;
    below_ptr[k + 1] = at;
  }
}

void yo_super_front_tree(void* h, uint64_t* out) {
  FrontTree t = front_tree(static_cast<SuperObj*>(h)->nodes);
  uint64_t v[7] = {t.fronts, t.roots, t.height, t.work, t.critical_work, t.solve_work, t.solve_critical_work};
  std::memcpy(out, v, sizeof v);
}

// ---- LU ----

int yo_lu(MAT_ARGS, int kind, double tau, const uint64_t* col_perm, void** out, uint64_t* pivot) {
  *pivot = UINT64_MAX;
  std::unique_ptr<LuObj> o(new LuObj);
  int status = guard([&] {
    Compressed a = MAT;
    if (a.row_major) throw SparseError(Err::InvalidArgument, "LU takes a CSC matrix");
    SymbolicLu s = col_perm ? analyze_lu_with_permutation(a, std::vector<size_t>(col_perm, col_perm + a.cols))
                            : analyze_lu(a, (Ordering)kind);
    try {
      o->f = factor_lu(s, a, tau);
    } catch (const SparseError& e) {
      *pivot = e.pivot;
      throw;
    }
  });
  if (status == 0) *out = o.release();
  return status;
}

void yo_lu_sizes(void* h, uint64_t* out) {
  auto* o = static_cast<LuObj*>(h);
  out[0] = o->f.dim();
  out[1] = o->f.l.nnz();
  out[2] = o->f.u.nnz();
}

void yo_lu_copy(void* h, uint64_t* l_ptr, uint64_t* l_idx, double* l_val, uint64_t* u_ptr, uint64_t* u_idx,
                double* u_val, uint64_t* row_perm, uint64_t* col_perm) {
  auto* o = static_cast<LuObj*>(h);
  copy_matrix(o->f.l, l_ptr, l_idx, l_val);
  copy_matrix(o->f.u, u_ptr, u_idx, u_val);
  for (size_t i = 0; i < o->f.dim(); ++i) {
    row_perm[i] = o->f.row_perm[i];
    col_perm[i] = o->f.col_perm[i];
  }
}

int yo_lu_solve_many(void* h, uint64_t count, const double* rhs, double* out, int threads) {
  return guard([&] { lu_solve_many(static_cast<LuObj*>(h)->f, rhs, count, out, threads); });
}

int yo_lu_solve_transpose_many(void* h, uint64_t count, const double* rhs, double* out, int threads) {
  return guard([&] { lu_solve_transpose_many(static_cast<LuObj*>(h)->f, rhs, count, out, threads); });
}

// lu.rs:662-757: outcome 0 = Reused, 1 = Repivoted (the handle then holds the fresh factors)
int yo_lu_refactor(void* h, MAT_ARGS, int* outcome) {
  return guard([&] {
    Compressed a = MAT;
    *outcome = (int)lu_refactor(static_cast<LuObj*>(h)->f, a);
  });
}

uint64_t yo_lu_refactor_work(void* h) { return lu_refactor_work(static_cast<LuObj*>(h)->f); }

int yo_lu_refine(void* h, MAT_ARGS, const double* b, uint64_t max_steps, double tol, double* x, uint64_t* steps,
                 double* ratio) {
  return guard([&] {
    const Lu& f = static_cast<LuObj*>(h)->f;
    Compressed a = MAT;
    std::vector<double> bv(b, b + a.rows);
    auto solve = [&](const std::vector<double>& rhs) {
      std::vector<double> out(rhs), scratch;
      lu_substitute(f, out.data(), scratch);
      return out;
    };
    NormwiseRefined r = refine_normwise(a, bv, max_steps, tol, solve);
    std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
    *steps = r.steps;
    *ratio = r.relative_residual;
  });
}

// ---- workloads ----

int yo_grid(uint64_t n, uint64_t seed, void** out) {
  return guard([&] {
    auto* o = new GridObj;
    o->g = make_grid(n, seed);
    *out = o;
  });
}

void yo_grid_sizes(void* h, uint64_t* out) {
  auto* o = static_cast<GridObj*>(h);
  out[0] = o->g.buses;
  out[1] = o->g.from.size();
}

void yo_grid_copy(void* h, double* bx, double* by, uint64_t* from, uint64_t* to, double* x) {
  const Grid& g = static_cast<GridObj*>(h)->g;
  std::memcpy(bx, g.bx.data(), g.buses * sizeof(double));
  std::memcpy(by, g.by.data(), g.buses * sizeof(double));
  for (size_t k = 0; k < g.from.size(); ++k) {
    from[k] = g.from[k];
    to[k] = g.to[k];
  }
  std::memcpy(x, g.x.data(), g.x.size() * sizeof(double));
}

void yo_grid_bridges(void* h, uint8_t* out) {
  std::vector<bool> b = bridges(static_cast<GridObj*>(h)->g);
  for (size_t k = 0; k < b.size(); ++k) out[k] = b[k];
}

int yo_dc(void* h, const uint8_t* in_service, void** bprime, void** bf, uint64_t* branch_rows) {
  return guard([&] {
    const Grid& g = static_cast<GridObj*>(h)->g;
    std::vector<bool> mask;
    if (in_service) mask.assign(in_service, in_service + g.from.size());
    DcMatrices dc = make_dc(g, in_service ? &mask : nullptr);
    auto* b = new MatrixObj;
    b->m = std::move(dc.bprime);
    auto* f = new MatrixObj;
    f->m = std::move(dc.bf);
    for (size_t k = 0; k < dc.branch_rows.size(); ++k) branch_rows[k] = dc.branch_rows[k];
    *bprime = b;
    *bf = f;
  });
}

int yo_jacobian(void* h, uint64_t seed, double shift, uint64_t shift_seed, void** jac, void** ac_out) {
  return guard([&] {
    auto* o = new AcObj;
    o->ac = make_jacobian(static_cast<GridObj*>(h)->g, seed, shift, shift_seed);
    auto* j = new MatrixObj;
    j->m = o->ac.jacobian;
    *jac = j;
    *ac_out = o;
  });
}

void yo_ac_sizes(void* h, uint64_t* out) {
  auto* o = static_cast<AcObj*>(h);
  out[0] = o->ac.pv.size();
  out[1] = o->ac.pq.size();
  out[2] = o->ac.vm.size();
  out[3] = o->ac.r.size();
}

void yo_ac_copy(void* h, uint64_t* pv, uint64_t* pq, double* vm, double* va, double* r, double* bc) {
  const AcJacobian& a = static_cast<AcObj*>(h)->ac;
  std::memcpy(r, a.r.data(), a.r.size() * sizeof(double));
  std::memcpy(bc, a.bc.data(), a.bc.size() * sizeof(double));
  for (size_t i = 0; i < a.pv.size(); ++i) pv[i] = a.pv[i];
  for (size_t i = 0; i < a.pq.size(); ++i) pq[i] = a.pq[i];
  std::memcpy(vm, a.vm.data(), a.vm.size() * sizeof(double));
  std::memcpy(va, a.va.data(), a.va.size() * sizeof(double));
}

void yo_uniforms(uint64_t seed, uint64_t stream, uint64_t count, double* out) {
  Rng r(seed, stream);
  for (size_t i = 0; i < count; ++i) out[i] = r.uniform();
}

void yo_normals(uint64_t seed, uint64_t stream, uint64_t count, double* out) {
  Rng r(seed, stream);
  for (size_t i = 0; i < count; ++i) out[i] = r.normal();
}

void yo_integers(uint64_t seed, uint64_t stream, uint64_t count, uint64_t* out) {
  Rng r(seed, stream);
  for (size_t i = 0; i < count; ++i) out[i] = r.next();
}

// Zero-sum injections: `count` scenarios of N(0,1) over `buses` buses, less their mean, with the
// slack (bus 0) dropped: `count` vectors of buses - 1 values, one after another.
void yo_injections(uint64_t seed, uint64_t stream, uint64_t buses, uint64_t count, double* out) {
  Rng r(seed, stream);
  std::vector<double> z(buses);
  for (size_t c = 0; c < count; ++c) {
    double sum = 0.0;
    for (size_t i = 0; i < buses; ++i) {
      z[i] = r.normal();
      sum = sum + z[i];
    }
    const double mean = sum / (double)buses;
    for (size_t i = 1; i < buses; ++i) out[c * (buses - 1) + (i - 1)] = z[i] - mean;
  }
}

}  // extern "C"
