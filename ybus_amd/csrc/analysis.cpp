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

// Fills s's per-position arrays from `form`, position k holding row order[k]
static void emit_rows(const RowForm& form, const std::vector<int64_t>& src_map, const std::vector<int64_t>& dst_map,
                      const std::vector<int64_t>& order, Sweep& s) {
  const int64_t n = form.n;
  const bool complex = !form.val_im.empty() || !form.diag_im.empty();
  s.src.resize(n);
  s.dst.resize(n);
  s.row_ptr.assign(n + 1, 0);
  s.dep.reserve(form.dep.size());
  s.val.reserve(form.val.size());
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
  emit_rows(form, src_map, dst_map, order, s);
  return s;
}

Sweep split_rows(const Sweep& s, int32_t cap, double continuation_diag) {
  const int32_t levels = s.levels;
  if (!s.val_im.empty() || cap < 1) return s;
  // the level at which each destination row is final
  std::vector<int32_t> final_at(s.n + 1, 0);
  for (int32_t l = 0; l < levels; ++l)
    for (int32_t k = s.level_ptr[l]; k < s.level_ptr[l + 1]; ++k) final_at[s.dst[k]] = l;
  // each level's part: its first level and the end of its levels. A partitioned sweep's parts run
  // apart (each in its own workgroups, the top part before or after all the others), so a row's
  // segments stay within its part's levels, and a dependency in another part is final before
  // the part starts
  std::vector<int32_t> part_begin(levels, 0), part_end(levels, levels);
  if (s.parts)
    for (int32_t q = 0; q <= s.parts; ++q)
      for (int32_t l = s.part_level[q]; l < s.part_level[q + 1]; ++l) {
        part_begin[l] = s.part_level[q];
        part_end[l] = s.part_level[q + 1];
      }
  struct Segment {
    int32_t k, p0, p1;
    bool first, last;
  };
  std::vector<std::vector<Segment>> at(levels);
  for (int32_t l = 0; l < levels; ++l) {
    for (int32_t k = s.level_ptr[l]; k < s.level_ptr[l + 1]; ++k) {
      const int32_t p0 = s.row_ptr[k], p1 = s.row_ptr[k + 1];
      if (p1 - p0 <= cap) {
        at[l].push_back({k, p0, p1, true, true});
        continue;
      }
      const int32_t first = part_begin[l], end_level = part_end[l];
      // a dependency's level as this part sees it
      auto final_of = [&](int32_t d) {
        const int32_t f = final_at[d];
        return f >= first && f < end_level ? f : first - 1;
      };
      int32_t end = p1, level = l;
      bool last = true;
      while (end > p0) {
        int32_t begin = std::max(p0, end - cap);
        if (level == first) {
          begin = p0;
        } else {
          // every term left for the earlier segments runs at level - 1 at the latest, so its
          // dependency must be final by level - 2: this segment starts at the first term whose
          // dependency is final only at level - 1 or later, if it is not already before it
          for (int32_t q = p0; q < begin; ++q)
            if (final_of(s.dep[q]) >= level - 1) {
              begin = q;
              break;
            }
        }
        at[level].push_back({k, begin, end, begin == p0, last});
        last = false;
        end = begin;
        --level;
      }
    }
  }
  Sweep out;
  out.n = s.n;
  out.levels = levels;
  out.unit = s.unit;
  out.split = true;
  out.parts = s.parts;
  out.part_level = s.part_level;
  out.level_ptr.assign(levels + 1, 0);
  for (int32_t l = 0; l < levels; ++l) {
    out.level_ptr[l + 1] = out.level_ptr[l] + (int32_t)at[l].size();
    out.widest_level = std::max(out.widest_level, (int32_t)at[l].size());
  }
  const int32_t positions = out.level_ptr[levels];
  out.src.reserve(positions);
  out.dst.reserve(positions);
  out.row_ptr.reserve(positions + 1);
  out.row_ptr.push_back(0);
  out.dep.reserve(s.dep.size());
  out.val.reserve(s.val.size());
  if (!s.unit) out.diag.reserve(positions);
  for (int32_t l = 0; l < levels; ++l) {
    for (const Segment& g : at[l]) {
      out.src.push_back(g.first ? s.src[g.k] : ~s.dst[g.k]);
      out.dst.push_back(s.dst[g.k]);
      out.dep.insert(out.dep.end(), s.dep.begin() + g.p0, s.dep.begin() + g.p1);
      out.val.insert(out.val.end(), s.val.begin() + g.p0, s.val.begin() + g.p1);
      out.row_ptr.push_back((int32_t)out.dep.size());
      if (!s.unit) out.diag.push_back(g.last ? s.diag[g.k] : continuation_diag);
    }
  }
  return out;
}

TopSlots top_slots(const Sweep& f, const Sweep& b) {
  TopSlots out;
  const TopSlots bad;
  if (f.parts < 1 || b.parts != f.parts || f.n != b.n) return bad;
  auto range = [](const Sweep& s, TopSlots::Part& p) {
    p.level0 = s.part_level[s.parts];
    p.levels = s.part_level[s.parts + 1] - p.level0;
    p.seg0 = s.level_ptr[p.level0];
    p.segs = s.level_ptr[p.level0 + p.levels] - p.seg0;
    p.term0 = s.row_ptr[p.seg0];
    p.terms = s.row_ptr[p.seg0 + p.segs] - p.term0;
    p.slot.resize(p.segs);
    p.dep.resize(p.terms);
  };
  range(f, out.fwd);
  range(b, out.bwd);
  const int32_t n = f.n;
  // forward: the top rows in the order of their first segments, then the halo
  std::vector<int32_t> slot_t(n, -1), halo_t(n, -1);
  for (int32_t k = out.fwd.seg0; k < out.fwd.seg0 + out.fwd.segs; ++k) {
    const int32_t d = f.dst[k];
    if (slot_t[d] < 0) {
      if (f.src[k] < 0) return bad;  // a continuation before its row's first segment
      slot_t[d] = out.rows++;
      out.rhs_row.push_back(f.src[k]);
    }
    out.fwd.slot[k - out.fwd.seg0] = slot_t[d];
  }
  for (int32_t q = 0; q < out.fwd.terms; ++q) {
    const int32_t d = f.dep[out.fwd.term0 + q];
    if (slot_t[d] >= 0) {
      out.fwd.dep[q] = slot_t[d];
    } else {
      if (halo_t[d] < 0) {
        halo_t[d] = out.halo++;
        out.halo_row.push_back(d);
      }
      out.fwd.dep[q] = out.rows + halo_t[d];
    }
  }
  // backward: a first segment reads its row's forward value (so takes that row's slot) and
  // names the row's solution row; its terms read solution rows of the top
  std::vector<int32_t> slot_x(n, -1);
  out.out_row.assign(out.rows, -1);
  for (int32_t k = out.bwd.seg0; k < out.bwd.seg0 + out.bwd.segs; ++k) {
    const int32_t d = b.dst[k];
    if (b.src[k] >= 0) {
      const int32_t sl = slot_t[b.src[k]];
      if (sl < 0 || slot_x[d] >= 0 || out.out_row[sl] >= 0) return bad;
      slot_x[d] = sl;
      out.out_row[sl] = d;
    } else if (slot_x[d] < 0) {
      return bad;
    }
    out.bwd.slot[k - out.bwd.seg0] = slot_x[d];
  }
  for (int32_t sl = 0; sl < out.rows; ++sl)
    if (out.out_row[sl] < 0) return bad;
  for (int32_t q = 0; q < out.bwd.terms; ++q) {
    const int32_t sl = slot_x[b.dep[out.bwd.term0 + q]];
    if (sl < 0) return bad;
    out.bwd.dep[q] = sl;
  }
  return out;
}

PartSlots part_slots(const Sweep& f, const Sweep& b) {
  const PartSlots bad;
  if (f.parts < 1 || b.parts != f.parts || f.n != b.n) return bad;
  PartSlots out;
  out.parts = f.parts;
  std::vector<int32_t> slot_of(f.n, -1), halo_of(f.n, -1);
  // one sweep's parts; a term outside its part is a halo slot where `halo` allows it
  auto build = [&](const Sweep& s, PartSlots::Dir& d, bool halo) {
    d.seg_begin.assign(1, 0);
    d.slot_begin.assign(1, 0);
    d.halo_begin.assign(1, 0);
    for (int32_t q = 0; q < s.parts; ++q) {
      const int32_t k0 = s.level_ptr[s.part_level[q]], k1 = s.level_ptr[s.part_level[q + 1]];
      if (k1 <= k0) return false;
      const size_t first_row = d.out_row.size(), first_halo = d.halo_row.size();
      int32_t rows = 0, halos = 0;
      bool ok = true;
      for (int32_t k = k0; k < k1 && ok; ++k) {
        const int32_t r = s.dst[k];
        if (slot_of[r] < 0) {
          if (s.src[k] < 0) ok = false;  // a continuation before its row's first segment
          slot_of[r] = rows++;
          d.in_row.push_back(s.src[k]);
          d.out_row.push_back(r);
        }
        d.slot.push_back(slot_of[r]);
      }
      for (int32_t p = s.row_ptr[k0]; p < s.row_ptr[k1] && ok; ++p) {
        const int32_t r = s.dep[p];
        if (slot_of[r] >= 0) {
          d.dep.push_back(slot_of[r]);
        } else if (!halo) {
          ok = false;
        } else {
          if (halo_of[r] < 0) {
            halo_of[r] = halos++;
            d.halo_row.push_back(r);
          }
          d.dep.push_back(rows + halo_of[r]);
        }
      }
      for (size_t i = first_row; i < d.out_row.size(); ++i) slot_of[d.out_row[i]] = -1;
      for (size_t i = first_halo; i < d.halo_row.size(); ++i) halo_of[d.halo_row[i]] = -1;
      if (!ok) return false;
      d.seg_begin.push_back(k1 - s.level_ptr[s.part_level[0]]);
      d.slot_begin.push_back((int32_t)d.out_row.size());
      d.halo_begin.push_back((int32_t)d.halo_row.size());
    }
    return true;
  };
  // the forward parts read only their own rows; a backward part also reads top rows, final
  // before the backward parts start
  if (!build(f, out.fwd, false) || !build(b, out.bwd, true)) return bad;
  return out;
}

WindowCode window_code(const Sweep& s, int32_t cap) {
  WindowCode w;
  w.cap = cap;
  const int32_t positions = s.positions(), n = s.n;
  if (cap < 1) return w;
  std::vector<int32_t> level_of(positions), written(n + 1, -1);
  for (int32_t l = 0; l < s.levels; ++l)
    for (int32_t k = s.level_ptr[l]; k < s.level_ptr[l + 1]; ++k) level_of[k] = l;
  w.src.resize(positions);
  w.dep.resize(s.dep.size());
  auto code = [&](int32_t row, int32_t l, int32_t end) -> int32_t {
    const int32_t q = row >= 0 && row <= n ? written[row] : -1;
    if (q < 0) return row;
    if (q >= end - cap) return ~(q % cap);
    if (level_of[q] <= l - WINDOW_AHEAD - 1) return row;
    return row + n + 1;
  };
  for (int32_t k = 0; k < positions; ++k) {
    const int32_t l = level_of[k], end = s.level_ptr[l + 1];
    if (s.src[k] >= 0) {
      w.src[k] = s.src[k];
    } else {
      const int32_t c = code(s.dst[k], l, end);
      w.src[k] = c < 0 ? c : WINDOW_PARTIAL;
    }
    for (int32_t p = s.row_ptr[k]; p < s.row_ptr[k + 1]; ++p) w.dep[p] = code(s.dep[p], l, end);
    written[s.dst[k]] = k;
  }
  return w;
}

std::vector<int32_t> tree_parts(const RowForm& lower, int32_t target, int32_t& parts) {
  parts = 0;
  const int64_t n = lower.n;
  if (!lower.lower || n < 2 || target < 2) return {};
  // dependents of each row, and each row's first (smallest) dependent
  std::vector<int64_t> dptr(n + 1, 0);
  for (int64_t p = 0; p < lower.ptr[n]; ++p) dptr[lower.dep[p] + 1] += 1;
  for (int64_t r = 0; r < n; ++r) dptr[r + 1] += dptr[r];
  std::vector<int32_t> dependents(dptr[n]);
  {
    std::vector<int64_t> cursor(dptr.begin(), dptr.end() - 1);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t p = lower.ptr[i]; p < lower.ptr[i + 1]; ++p) dependents[cursor[lower.dep[p]]++] = (int32_t)i;
  }
  std::vector<int32_t> parent(n, -1), mark(n, -1);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = dptr[r]; e < dptr[r + 1]; ++e)
      if (parent[r] < 0 || dependents[e] < parent[r]) parent[r] = dependents[e];
  for (int64_t r = 0; r < n; ++r) {
    const int32_t up = parent[r];
    if (up < 0) continue;
    for (int64_t e = dptr[up]; e < dptr[up + 1]; ++e) mark[dependents[e]] = (int32_t)r;
    for (int64_t e = dptr[r]; e < dptr[r + 1]; ++e)
      if (dependents[e] != up && mark[dependents[e]] != (int32_t)r) return {};
  }
  // subtree sizes (a parent comes after its children) and children lists
  std::vector<int64_t> size(n, 1);
  for (int64_t r = 0; r < n; ++r)
    if (parent[r] >= 0) size[parent[r]] += size[r];
  std::vector<int64_t> cptr(n + 1, 0);
  for (int64_t r = 0; r < n; ++r)
    if (parent[r] >= 0) cptr[parent[r] + 1] += 1;
  for (int64_t r = 0; r < n; ++r) cptr[r + 1] += cptr[r];
  std::vector<int32_t> children(cptr[n]);
  {
    std::vector<int64_t> cursor(cptr.begin(), cptr.end() - 1);
    for (int64_t r = 0; r < n; ++r)
      if (parent[r] >= 0) children[cursor[parent[r]]++] = (int32_t)r;
  }
  // cut the largest subtree until every one is small enough
  const int64_t limit = std::max<int64_t>(1, n / (2 * (int64_t)target));
  auto larger = [&](int32_t a, int32_t b) { return size[a] < size[b] || (size[a] == size[b] && a > b); };
  std::vector<int32_t> heap;
  for (int64_t r = 0; r < n; ++r)
    if (parent[r] < 0) heap.push_back((int32_t)r);
  std::make_heap(heap.begin(), heap.end(), larger);
  std::vector<char> top(n, 0);
  while (!heap.empty() && size[heap.front()] > limit && (int64_t)heap.size() < 16 * (int64_t)target) {
    std::pop_heap(heap.begin(), heap.end(), larger);
    const int32_t r = heap.back();
    heap.pop_back();
    top[r] = 1;
    for (int64_t c = cptr[r]; c < cptr[r + 1]; ++c) {
      heap.push_back(children[c]);
      std::push_heap(heap.begin(), heap.end(), larger);
    }
  }
  if (heap.size() < 2) return {};
  // largest subtree first into the lightest part
  std::sort(heap.begin(), heap.end(), [&](int32_t a, int32_t b) { return larger(b, a); });
  const int32_t k = (int32_t)std::min<int64_t>(target, (int64_t)heap.size());
  std::vector<int64_t> load(k, 0);
  std::vector<int32_t> part_of(n, -1);
  for (int32_t r : heap) {
    const int32_t q = (int32_t)(std::min_element(load.begin(), load.end()) - load.begin());
    load[q] += size[r];
    part_of[r] = q;
  }
  // every other row: the top part, or its subtree root's part (parents come after children)
  for (int64_t r = n; r-- > 0;) {
    if (top[r])
      part_of[r] = k;
    else if (part_of[r] < 0)
      part_of[r] = part_of[parent[r]];
  }
  parts = k;
  return part_of;
}

Sweep schedule_parts(const RowForm& form, const std::vector<int64_t>& src_map, const std::vector<int64_t>& dst_map,
                     const std::vector<int32_t>& part_of, int32_t parts, bool top_first) {
  const int64_t n = form.n;
  Sweep bad;
  if ((int64_t)part_of.size() != n || parts < 1) return bad;
  // each row's level within its part, in the form's processing order
  std::vector<int32_t> level(n, 0);
  auto visit = [&](int64_t i) {
    const int32_t q = part_of[i];
    int32_t l = 0;
    for (int64_t p = form.ptr[i]; p < form.ptr[i + 1]; ++p) {
      const int32_t d = form.dep[p], qd = part_of[d];
      const bool allowed = top_first ? (qd == q || qd == parts) && !(q == parts && qd != parts)
                                     : (qd == q || q == parts);
      if (!allowed) return false;
      if (qd == q) l = std::max(l, level[d] + 1);
    }
    level[i] = l;
    return true;
  };
  if (form.lower) {
    for (int64_t i = 0; i < n; ++i)
      if (!visit(i)) return bad;
  } else {
    for (int64_t i = n; i-- > 0;)
      if (!visit(i)) return bad;
  }
  // levels per part, and each part's first global level
  std::vector<int32_t> depth(parts + 1, 0);
  for (int64_t i = 0; i < n; ++i) depth[part_of[i]] = std::max(depth[part_of[i]], level[i] + 1);
  Sweep s;
  s.n = (int32_t)n;
  s.unit = form.unit;
  s.parts = parts;
  s.part_level.assign(parts + 2, 0);
  for (int32_t q = 0; q <= parts; ++q) s.part_level[q + 1] = s.part_level[q] + depth[q];
  s.levels = s.part_level[parts + 1];
  s.level_ptr.assign(s.levels + 1, 0);
  std::vector<int32_t> global(n);
  for (int64_t i = 0; i < n; ++i) {
    global[i] = s.part_level[part_of[i]] + level[i];
    s.level_ptr[global[i] + 1] += 1;
  }
  for (int32_t l = 0; l < s.levels; ++l) {
    s.widest_level = std::max(s.widest_level, s.level_ptr[l + 1]);
    s.level_ptr[l + 1] += s.level_ptr[l];
  }
  std::vector<int32_t> cursor(s.level_ptr.begin(), s.level_ptr.end() - 1);
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[cursor[global[i]]++] = i;  // ascending rows within a level
  emit_rows(form, src_map, dst_map, order, s);
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
