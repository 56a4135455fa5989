#include "analysis.hpp"

#include <algorithm>
#include <numeric>

namespace ybus_hip {

std::vector<int32_t> levels_of(const RowForm& form) {
  const int64_t n = form.n;
  std::vector<int32_t> level(n, 0);
  auto visit = [&](int64_t i) {
    int32_t l = 0;
    for (int64_t p = form.ptr[i]; p < form.ptr[i + 1]; ++p) l = std::max(l, level[form.dep[p]] + 1);
    level[i] = l;
  };
  if (form.lower) {
    for (int64_t i = 0; i < n; ++i) visit(i);
  } else {
    for (int64_t i = n; i-- > 0;) visit(i);
  }
  return level;
}

Sweep schedule(const RowForm& form, const std::vector<int64_t>& src_map, const std::vector<int64_t>& dst_map) {
  const int64_t n = form.n;
  const std::vector<int32_t> level = levels_of(form);
  const int32_t levels = n ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  Sweep s;
  s.n = (int32_t)n;
  s.levels = levels;
  s.unit = form.unit;
  s.level_ptr.assign(levels + 1, 0);
  for (int64_t i = 0; i < n; ++i) s.level_ptr[level[i] + 1] += 1;
  for (int32_t l = 0; l < levels; ++l) {
    s.widest_level = std::max(s.widest_level, s.level_ptr[l + 1]);
    s.level_ptr[l + 1] += s.level_ptr[l];
  }
  std::vector<int32_t> cursor(s.level_ptr.begin(), s.level_ptr.end() - (levels ? 1 : 0));
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[cursor[level[i]]++] = i;  // ascending rows within a level
  s.src.resize(n);
  s.dst.resize(n);
  s.row_ptr.assign(n + 1, 0);
  s.dep.reserve(form.dep.size());
  s.val.reserve(form.val.size());
  const bool complex = !form.val_im.empty() || !form.diag_im.empty();
  if (complex) s.val_im.reserve(form.val.size());
  if (!form.unit) s.diag.resize(n);
  if (!form.unit && complex) s.diag_im.resize(n);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = order[k];
    s.src[k] = (int32_t)src_map[i];
    s.dst[k] = (int32_t)dst_map[i];
    for (int64_t p = form.ptr[i]; p < form.ptr[i + 1]; ++p) {
      s.dep.push_back((int32_t)dst_map[form.dep[p]]);
      s.val.push_back(form.val[p]);
      if (complex) s.val_im.push_back(form.val_im[p]);
    }
    s.row_ptr[k + 1] = (int32_t)s.dep.size();
    if (!form.unit) s.diag[k] = form.diag[i];
    if (!form.unit && complex) s.diag_im[k] = form.diag_im[i];
  }
  return s;
}

RowForm cholesky_forward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val) {
  RowForm f;
  f.n = n;
  f.lower = true;
  f.ptr.assign(n + 1, 0);
  f.diag.assign(n, 1.0);
  for (int64_t i = 0; i < n; ++i) {
    // forward_by_rows (cholesky.rs:468-483): entries left of the diagonal, ascending, until the
    // first entry at or right of it, which is the diagonal
    for (uint64_t p = l_ptr[i]; p < l_ptr[i + 1]; ++p) {
      if ((int64_t)l_idx[p] < i) {
        f.dep.push_back((int32_t)l_idx[p]);
        f.val.push_back(l_val[p]);
      } else {
        f.diag[i] = l_val[p];
        break;
      }
    }
    f.ptr[i + 1] = (int64_t)f.dep.size();
  }
  return f;
}

RowForm cholesky_backward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val) {
  // L^T in CSR is L in CSC: a counting sort of L's entries by column keeps rows ascending
  // (compressed_matrix.rs:385-425), the diagonal first (cholesky.rs:285-299)
  std::vector<int64_t> cptr(n + 1, 0);
  for (int64_t i = 0; i < n; ++i)
    for (uint64_t p = l_ptr[i]; p < l_ptr[i + 1]; ++p) cptr[l_idx[p] + 1] += 1;
  for (int64_t j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
  std::vector<int64_t> cursor(cptr.begin(), cptr.end() - 1);
  std::vector<int32_t> rows(cptr[n]);
  std::vector<double> vals(cptr[n]);
  for (int64_t i = 0; i < n; ++i)
    for (uint64_t p = l_ptr[i]; p < l_ptr[i + 1]; ++p) {
      int64_t q = cursor[l_idx[p]]++;
      rows[q] = (int32_t)i;
      vals[q] = l_val[p];
    }
  RowForm f;
  f.n = n;
  f.lower = false;
  f.ptr.assign(n + 1, 0);
  f.diag.assign(n, 1.0);
  for (int64_t j = 0; j < n; ++j) {
    // back_by_rows (cholesky.rs:487-501): the diagonal divides, entries right of it subtract,
    // in stored (ascending) order
    for (int64_t q = cptr[j]; q < cptr[j + 1]; ++q) {
      if (rows[q] == j) {
        f.diag[j] = vals[q];
      } else if (rows[q] > j) {
        f.dep.push_back(rows[q]);
        f.val.push_back(vals[q]);
      }
    }
    f.ptr[j + 1] = (int64_t)f.dep.size();
  }
  return f;
}

// The rows of a CSC triangle, each row's entries in ascending column order
static void csc_rows(int64_t n, const uint64_t* ptr, const uint64_t* idx, const double* val, bool below,
                     std::vector<int64_t>& rptr, std::vector<int32_t>& cols, std::vector<double>& vals) {
  rptr.assign(n + 1, 0);
  for (int64_t j = 0; j < n; ++j)
    for (uint64_t p = ptr[j]; p < ptr[j + 1]; ++p) {
      int64_t i = (int64_t)idx[p];
      if (below ? i > j : i < j) rptr[i + 1] += 1;
    }
  for (int64_t i = 0; i < n; ++i) rptr[i + 1] += rptr[i];
  std::vector<int64_t> cursor(rptr.begin(), rptr.end() - 1);
  cols.resize(rptr[n]);
  vals.resize(rptr[n]);
  for (int64_t j = 0; j < n; ++j)
    for (uint64_t p = ptr[j]; p < ptr[j + 1]; ++p) {
      int64_t i = (int64_t)idx[p];
      if (below ? i > j : i < j) {
        int64_t q = cursor[i]++;
        cols[q] = (int32_t)j;
        vals[q] = val[p];
      }
    }
}

RowForm lu_forward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val) {
  RowForm f;
  f.n = n;
  f.lower = true;
  f.unit = true;
  std::vector<int32_t> cols;
  csc_rows(n, l_ptr, l_idx, l_val, true, f.ptr, cols, f.val);
  f.dep = std::move(cols);
  return f;
}

RowForm lu_backward(int64_t n, const uint64_t* u_ptr, const uint64_t* u_idx, const double* u_val) {
  RowForm f;
  f.n = n;
  f.lower = false;
  std::vector<int64_t> rptr;
  std::vector<int32_t> cols;
  std::vector<double> vals;
  csc_rows(n, u_ptr, u_idx, u_val, false, rptr, cols, vals);
  f.ptr = rptr;
  f.dep.resize(cols.size());
  f.val.resize(vals.size());
  for (int64_t i = 0; i < n; ++i) {
    // descending columns: back_substitute (lu.rs:640-649) updates y[i] from column n-1 down
    int64_t a = rptr[i], b = rptr[i + 1];
    for (int64_t q = 0; q < b - a; ++q) {
      f.dep[a + q] = cols[b - 1 - q];
      f.val[a + q] = vals[b - 1 - q];
    }
  }
  f.diag.resize(n);
  for (int64_t j = 0; j < n; ++j) f.diag[j] = u_val[u_ptr[j + 1] - 1];
  return f;
}

RowForm lu_transpose_forward(int64_t n, const uint64_t* u_ptr, const uint64_t* u_idx, const double* u_val) {
  RowForm f;
  f.n = n;
  f.lower = true;
  f.ptr.assign(n + 1, 0);
  f.diag.resize(n);
  for (int64_t j = 0; j < n; ++j) {
    const uint64_t diagonal = u_ptr[j + 1] - 1;
    for (uint64_t p = u_ptr[j]; p < diagonal; ++p) {
      f.dep.push_back((int32_t)u_idx[p]);
      f.val.push_back(u_val[p]);
    }
    f.diag[j] = u_val[diagonal];
    f.ptr[j + 1] = (int64_t)f.dep.size();
  }
  return f;
}

RowForm lu_transpose_backward(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val) {
  RowForm f;
  f.n = n;
  f.lower = false;
  f.unit = true;
  f.ptr.assign(n + 1, 0);
  for (int64_t j = 0; j < n; ++j) {
    for (uint64_t p = l_ptr[j] + 1; p < l_ptr[j + 1]; ++p) {
      f.dep.push_back((int32_t)l_idx[p]);
      f.val.push_back(l_val[p]);
    }
    f.ptr[j + 1] = (int64_t)f.dep.size();
  }
  return f;
}

RowForm complex_form(RowForm re, const RowForm& im, bool conjugate) {
  re.val_im = im.val;
  if (conjugate)
    for (double& v : re.val_im) v = -v;
  if (!re.unit) re.diag_im = im.diag;
  return re;
}

Supernodes cholesky_supernodes(int64_t n, const uint64_t* l_ptr, const uint64_t* l_idx) {
  std::vector<int64_t> counts(n, 0), parent(n, -1), children(n, 0);
  for (int64_t i = 0; i < n; ++i)
    for (uint64_t p = l_ptr[i]; p < l_ptr[i + 1]; ++p) {
      int64_t j = (int64_t)l_idx[p];
      counts[j] += 1;
      // rows arrive in increasing order, so the first row below j is j's parent
      if (j < i && parent[j] < 0) parent[j] = i;
    }
  for (int64_t j = 0; j < n; ++j)
    if (parent[j] >= 0) children[parent[j]] += 1;
  Supernodes s;
  int64_t start = 0;
  for (int64_t j = 0; j < n; ++j) {
    bool continues = j + 1 < n && parent[j] == j + 1 && children[j + 1] == 1 && counts[j] == counts[j + 1] + 1;
    if (!continues) {
      s.start.push_back(start);
      int64_t size = j + 1 - start;
      s.largest = std::max(s.largest, size);
      if (size > 1) s.in_chains += size;
      start = j + 1;
    }
  }
  s.start.push_back(n);
  return s;
}

bool valid_compressed(int64_t major, int64_t minor, const uint64_t* ptr, const uint64_t* idx, std::string* why) {
  if (ptr[0] != 0) {
    *why = "the first pointer is not 0";
    return false;
  }
  for (int64_t i = 0; i < major; ++i) {
    if (ptr[i + 1] < ptr[i]) {
      *why = "pointers decrease";
      return false;
    }
    for (uint64_t p = ptr[i]; p < ptr[i + 1]; ++p)
      if ((int64_t)idx[p] >= minor) {
        *why = "an index is out of range";
        return false;
      }
  }
  return true;
}

int tile_width_for(int64_t ncols, int cus) {
  for (int tw = 64; tw > 1; tw /= 2)
    if ((ncols + tw - 1) / tw * 10 >= 9 * (int64_t)cus) return tw;
  return 1;
}

}  // namespace ybus_hip
