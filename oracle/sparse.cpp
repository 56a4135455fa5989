// Compressed-matrix core and fill-reducing orderings of the oracle (test infrastructure only):
// a restatement of ybus_sparse/src/compressed_matrix.rs, access.rs, product.rs, ordering.rs
// and colamd.rs. Reference line numbers are given per function.
#include <algorithm>
#include <cmath>
#include <deque>
#include <map>

#include "oracle.hpp"

namespace oracle {

// compressed_matrix.rs:254-321. Triplets are grouped by major index with a counting sort that
// keeps input order, each group is sorted stably by minor index, and duplicates are summed in
// input order. Every position that appears is stored, even if its values sum to zero.
Compressed from_triplets(size_t rows, size_t cols, bool row_major, const std::vector<Triplet>& t) {
  for (const Triplet& e : t) {
    if (e.row >= rows || e.col >= cols)
      throw SparseError(Err::OutOfBounds, "triplet out of bounds");
  }
  const size_t major_dim = row_major ? rows : cols;
  auto major_minor = [&](size_t r, size_t c) {
    return row_major ? std::make_pair(r, c) : std::make_pair(c, r);
  };
  std::vector<size_t> start(major_dim + 1, 0);
  for (const Triplet& e : t) start[major_minor(e.row, e.col).first + 1] += 1;
  for (size_t i = 0; i < major_dim; ++i) start[i + 1] += start[i];
  std::vector<size_t> cursor = start;
  struct G {
    size_t minor, k;
    double value;
  };
  std::vector<G> grouped(t.size());
  for (size_t k = 0; k < t.size(); ++k) {
    auto mm = major_minor(t[k].row, t[k].col);
    grouped[cursor[mm.first]] = G{mm.second, k, t[k].value};
    cursor[mm.first] += 1;
  }
  Compressed out;
  out.rows = rows;
  out.cols = cols;
  out.row_major = row_major;
  out.ptr.assign(major_dim + 1, 0);
  out.idx.reserve(t.size());
  out.data.reserve(t.size());
  for (size_t major = 0; major < major_dim; ++major) {
    std::stable_sort(grouped.begin() + start[major], grouped.begin() + start[major + 1],
                     [](const G& a, const G& b) { return a.minor < b.minor; });
    for (size_t q = start[major]; q < start[major + 1]; ++q) {
      const G& g = grouped[q];
      bool first = !(out.idx.size() > out.ptr[major] && out.idx.back() == g.minor);
      if (first) {
        out.idx.push_back(g.minor);
        out.data.push_back(g.value);
      } else {
        out.data.back() = out.data.back() + g.value;
      }
    }
    out.ptr[major + 1] = out.idx.size();
  }
  return out;
}

// compressed_matrix.rs:385-425: a counting sort by minor index; visiting the major indices in
// order leaves each new segment sorted.
Compressed reoriented(const Compressed& a, bool row_major) {
  if (a.row_major == row_major) return a;
  const size_t minor_dim = a.minor_dim();
  Compressed out;
  out.rows = a.rows;
  out.cols = a.cols;
  out.row_major = row_major;
  out.ptr.assign(minor_dim + 1, 0);
  for (size_t m : a.idx) out.ptr[m + 1] += 1;
  for (size_t i = 0; i < minor_dim; ++i) out.ptr[i + 1] += out.ptr[i];
  std::vector<size_t> cursor = out.ptr;
  out.idx.assign(a.idx.size(), 0);
  out.data.assign(a.data.size(), 0.0);
  for (size_t major = 0; major < a.major_dim(); ++major) {
    for (size_t p = a.ptr[major]; p < a.ptr[major + 1]; ++p) {
      size_t minor = a.idx[p];
      out.idx[cursor[minor]] = major;
      out.data[cursor[minor]] = a.data[p];
      cursor[minor] += 1;
    }
  }
  return out;
}

// access.rs:175-219
Compressed select(const Compressed& a, const std::vector<size_t>& rows, const std::vector<size_t>& cols) {
  for (size_t r : rows)
    if (r >= a.rows) throw SparseError(Err::OutOfBounds, "row index out of range");
  for (size_t c : cols)
    if (c >= a.cols) throw SparseError(Err::OutOfBounds, "column index out of range");
  const std::vector<size_t>& majors = a.row_major ? rows : cols;
  const std::vector<size_t>& minors = a.row_major ? cols : rows;
  std::vector<std::vector<size_t>> positions(a.minor_dim());
  for (size_t position = 0; position < minors.size(); ++position)
    positions[minors[position]].push_back(position);
  Compressed out;
  out.rows = rows.size();
  out.cols = cols.size();
  out.row_major = a.row_major;
  out.ptr.clear();
  out.ptr.push_back(0);
  std::vector<std::pair<size_t, double>> segment;
  for (size_t major : majors) {
    segment.clear();
    for (size_t p = a.ptr[major]; p < a.ptr[major + 1]; ++p)
      for (size_t position : positions[a.idx[p]]) segment.emplace_back(position, a.data[p]);
    // sort_unstable_by_key on distinct positions (a position appears once per major index
    // unless the minor list repeats an index, where equal keys carry equal values)
    std::sort(segment.begin(), segment.end(),
              [](const auto& x, const auto& y) { return x.first < y.first; });
    for (auto& e : segment) {
      out.idx.push_back(e.first);
      out.data.push_back(e.second);
    }
    out.ptr.push_back(out.idx.size());
  }
  return out;
}

// access.rs (check_permutation): the right length, every index in range, none repeated
void check_permutation(const std::vector<size_t>& perm, size_t n) {
  if (perm.size() != n) throw SparseError(Err::InvalidPermutation, "wrong permutation length");
  std::vector<bool> seen(n, false);
  for (size_t i : perm) {
    if (i >= n || seen[i]) throw SparseError(Err::InvalidPermutation, "not a permutation");
    seen[i] = true;
  }
}

// access.rs:245-250
Compressed permute_symmetric(const Compressed& a, const std::vector<size_t>& perm) {
  if (!a.is_square()) throw SparseError(Err::NotSquare, "not square");
  check_permutation(perm, a.rows);
  return select(a, perm, perm);
}

// product.rs:17-27 (dot_kernel: each result summed from 0.0 in index order) and :32-46
// (scatter_kernel), chosen by orientation as mul_vec does (:93-105)
std::vector<double> mul_vec(const Compressed& a, const std::vector<double>& x) {
  if (x.size() != a.cols) throw SparseError(Err::ShapeMismatch, "x has the wrong length");
  std::vector<double> y;
  if (a.row_major) {
    y.assign(a.rows, 0.0);
    for (size_t i = 0; i < a.rows; ++i) {
      double sum = 0.0;
      for (size_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) sum = sum + a.data[p] * x[a.idx[p]];
      y[i] = sum;
    }
  } else {
    y.assign(a.rows, 0.0);
    for (size_t major = 0; major < a.cols; ++major)
      for (size_t p = a.ptr[major]; p < a.ptr[major + 1]; ++p)
        y[a.idx[p]] = y[a.idx[p]] + a.data[p] * x[major];
  }
  return y;
}

// properties.rs:79-81: A equals its transpose as a matrix (stored zeros ignored)
bool is_symmetric(const Compressed& a) {
  if (!a.is_square()) return false;
  auto value_at = [&](size_t major, size_t minor) {
    auto b = a.idx.begin() + a.ptr[major], e = a.idx.begin() + a.ptr[major + 1];
    auto it = std::lower_bound(b, e, minor);
    return (it != e && *it == minor) ? a.data[it - a.idx.begin()] : 0.0;
  };
  for (size_t major = 0; major < a.major_dim(); ++major)
    for (size_t p = a.ptr[major]; p < a.ptr[major + 1]; ++p)
      if (a.data[p] != value_at(a.idx[p], major)) return false;
  return true;
}

// ordering.rs:75-117: the graph of A + A^T without self-loops, each list sorted and unique
Adjacency Adjacency::of(const Compressed& a) {
  const size_t n = a.rows;
  std::vector<size_t> start(n + 1, 0);
  auto each_entry = [&](auto f) {
    for (size_t major = 0; major < a.major_dim(); ++major)
      for (size_t p = a.ptr[major]; p < a.ptr[major + 1]; ++p) {
        size_t row = a.row_major ? major : a.idx[p];
        size_t col = a.row_major ? a.idx[p] : major;
        if (row != col) f(row, col);
      }
  };
  each_entry([&](size_t row, size_t col) {
    start[row + 1] += 1;
    start[col + 1] += 1;
  });
  for (size_t i = 0; i < n; ++i) start[i + 1] += start[i];
  std::vector<size_t> cursor = start;
  std::vector<size_t> neighbours(start[n], 0);
  each_entry([&](size_t row, size_t col) {
    neighbours[cursor[row]++] = col;
    neighbours[cursor[col]++] = row;
  });
  size_t write = 0;
  std::vector<size_t> compact(n + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    std::sort(neighbours.begin() + start[i], neighbours.begin() + start[i + 1]);
    size_t previous = NONE;
    for (size_t k = start[i]; k < start[i + 1]; ++k) {
      size_t node = neighbours[k];
      if (previous != node) {
        neighbours[write++] = node;
        previous = node;
      }
    }
    compact[i + 1] = write;
  }
  neighbours.resize(write);
  Adjacency g;
  g.start = std::move(compact);
  g.neighbours = std::move(neighbours);
  return g;
}

// ordering.rs:46-63
std::vector<size_t> ordering(const Compressed& a, Ordering kind, unsigned mutations) {
  if (kind == Ordering::Colamd) {
    std::vector<std::pair<size_t, size_t>> entries;
    entries.reserve(a.nnz());
    for (size_t major = 0; major < a.major_dim(); ++major)
      for (size_t p = a.ptr[major]; p < a.ptr[major + 1]; ++p)
        entries.emplace_back(a.row_major ? major : a.idx[p], a.row_major ? a.idx[p] : major);
    return column_approximate_minimum_degree(a.rows, a.cols, entries, mutations);
  }
  if (!a.is_square()) throw SparseError(Err::NotSquare, "not square");
  switch (kind) {
    case Ordering::Natural: {
      std::vector<size_t> p(a.rows);
      for (size_t i = 0; i < a.rows; ++i) p[i] = i;
      return p;
    }
    case Ordering::Rcm:
      return reverse_cuthill_mckee(Adjacency::of(a));
    case Ordering::Amd:
      return approximate_minimum_degree(Adjacency::of(a), mutations);
    default:
      throw SparseError(Err::InvalidArgument, "unknown ordering");
  }
}

// ordering.rs:205-234: the breadth-first level structure from `root`: (levels, last level)
static std::pair<size_t, std::vector<size_t>> level_structure(const Adjacency& g, size_t root,
                                                              std::vector<size_t>& level) {
  std::vector<size_t> visited{root};
  level[root] = 0;
  std::deque<size_t> queue{root};
  while (!queue.empty()) {
    size_t node = queue.front();
    queue.pop_front();
    for (size_t k = g.start[node]; k < g.start[node + 1]; ++k) {
      size_t next = g.neighbours[k];
      if (level[next] == NONE) {
        level[next] = level[node] + 1;
        visited.push_back(next);
        queue.push_back(next);
      }
    }
  }
  size_t depth = level[visited.back()];
  std::vector<size_t> last;
  for (size_t node : visited)
    if (level[node] == depth) last.push_back(node);
  for (size_t node : visited) level[node] = NONE;
  return {depth + 1, last};
}

// ordering.rs:178-201 (George and Liu's pseudo-peripheral node)
static size_t pseudo_peripheral(const Adjacency& g, size_t seed, std::vector<size_t>& level) {
  size_t root = seed;
  auto [depth, last] = level_structure(g, root, level);
  for (;;) {
    size_t candidate = last[0];
    for (size_t node : last) {
      auto key = std::make_pair(g.degree(node), node);
      if (key < std::make_pair(g.degree(candidate), candidate)) candidate = node;
    }
    auto [cdepth, clast] = level_structure(g, candidate, level);
    if (cdepth > depth) {
      root = candidate;
      depth = cdepth;
      last = clast;
    } else {
      return root;
    }
  }
}

// ordering.rs:135-173
std::vector<size_t> reverse_cuthill_mckee(const Adjacency& g) {
  const size_t n = g.len();
  std::vector<size_t> order;
  order.reserve(n);
  std::vector<bool> numbered(n, false);
  std::vector<size_t> level(n, NONE);
  std::vector<size_t> candidates;
  for (size_t seed = 0; seed < n; ++seed) {
    if (numbered[seed]) continue;
    size_t root = pseudo_peripheral(g, seed, level);
    size_t head = order.size();
    numbered[root] = true;
    order.push_back(root);
    while (head < order.size()) {
      size_t node = order[head++];
      candidates.clear();
      for (size_t k = g.start[node]; k < g.start[node + 1]; ++k)
        if (!numbered[g.neighbours[k]]) candidates.push_back(g.neighbours[k]);
      std::sort(candidates.begin(), candidates.end(), [&](size_t a, size_t b) {
        return std::make_pair(g.degree(a), a) < std::make_pair(g.degree(b), b);
      });
      for (size_t next : candidates) {
        numbered[next] = true;
        order.push_back(next);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// ---- approximate minimum degree (ordering.rs:236-492) ----

namespace {
enum class Node { Variable, Merged, Element, Absorbed, Dense };

// ordering.rs:252-285: degree lists, each a doubly linked list; insert at the head (LIFO). With
// `fifo` (a planted mistake) a node goes to the tail instead.
struct DegreeLists {
  std::vector<size_t> head, next, previous;
  bool fifo = false;
  explicit DegreeLists(size_t n) : head(n + 1, NONE), next(n, NONE), previous(n, NONE) {}
  void insert(size_t node, size_t degree) {
    if (fifo && head[degree] != NONE) {
      size_t last = head[degree];
      while (next[last] != NONE) last = next[last];
      next[last] = node;
      previous[node] = last;
      next[node] = NONE;
      return;
    }
    next[node] = head[degree];
    previous[node] = NONE;
    if (head[degree] != NONE) previous[head[degree]] = node;
    head[degree] = node;
  }
  void remove(size_t node, size_t degree) {
    if (previous[node] != NONE)
      next[previous[node]] = next[node];
    else
      head[degree] = next[node];
    if (next[node] != NONE) previous[next[node]] = previous[node];
  }
};
}  // namespace

// ordering.rs:306-492. The supervariable merge groups variables by a hash of their sorted
// element and variable lists (ordering.rs:438-444) and merges, within a group, each later
// variable whose lists equal an earlier unmerged one's (:449-470). Groups are disjoint and
// list equality is an equivalence, so the merges are the same whatever order the groups are
// visited in; here a std::map keyed by the same wrapping hash visits them in key order.
std::vector<size_t> approximate_minimum_degree(const Adjacency& g, unsigned mutations) {
  const size_t n = g.len();
  std::vector<size_t> order;
  order.reserve(n);
  if (n == 0) return order;
  const size_t dense_threshold = (size_t)std::max(10.0 * std::sqrt((double)n), 16.0);
  std::vector<Node> kind(n);
  for (size_t i = 0; i < n; ++i)
    kind[i] = (g.degree(i) > dense_threshold && n > 2) ? Node::Dense : Node::Variable;

  std::vector<size_t> weight(n, 1);
  std::vector<std::vector<size_t>> variables(n), elements(n), element_list(n), members(n);
  for (size_t i = 0; i < n; ++i)
    for (size_t k = g.start[i]; k < g.start[i + 1]; ++k)
      if (kind[g.neighbours[k]] != Node::Dense) variables[i].push_back(g.neighbours[k]);
  std::vector<size_t> element_size(n, 0);
  std::vector<size_t> degree(n);
  for (size_t i = 0; i < n; ++i) degree[i] = variables[i].size();
  DegreeLists lists(n);
  lists.fifo = (mutations & FIFO_LISTS) != 0;
  size_t remaining = 0;
  for (size_t i = n; i-- > 0;) {
    if (kind[i] == Node::Variable) {
      lists.insert(i, degree[i]);
      remaining += 1;
    }
  }

  std::vector<size_t> in_pivot(n, NONE), outside(n, 0), outside_stamp(n, NONE), external(n, 0);
  size_t min_degree = 0;
  size_t step = 0;
  while (remaining > 0) {
    while (lists.head[min_degree] == NONE) min_degree += 1;
    const size_t p = lists.head[min_degree];
    lists.remove(p, min_degree);
    step += 1;

    in_pivot[p] = step;
    std::vector<size_t> pivot_list;
    for (size_t j : variables[p]) {
      if (kind[j] == Node::Variable && in_pivot[j] != step) {
        in_pivot[j] = step;
        pivot_list.push_back(j);
      }
    }
    std::vector<size_t> p_elements = std::move(elements[p]);
    elements[p].clear();
    for (size_t e : p_elements) {
      if (kind[e] != Node::Element) continue;
      for (size_t j : element_list[e]) {
        if (kind[j] == Node::Variable && in_pivot[j] != step) {
          in_pivot[j] = step;
          pivot_list.push_back(j);
        }
      }
      kind[e] = Node::Absorbed;
      element_list[e].clear();
    }
    variables[p].clear();
    kind[p] = Node::Element;
    remaining -= weight[p];
    order.push_back(p);
    order.insert(order.end(), members[p].begin(), members[p].end());
    members[p].clear();

    for (size_t i : pivot_list) {
      lists.remove(i, degree[i]);
      for (size_t e : elements[i]) {
        if (kind[e] == Node::Element) {
          if (outside_stamp[e] != step) {
            outside_stamp[e] = step;
            outside[e] = element_size[e];
          }
          outside[e] -= weight[i];
        }
      }
    }

    std::map<size_t, std::vector<size_t>> by_hash;
    for (size_t i : pivot_list) {
      std::vector<size_t> kept{p};
      size_t from_elements = 0;
      for (size_t e : elements[i]) {
        if (kind[e] != Node::Element) continue;
        if (outside[e] == 0 && !(mutations & NO_AGGRESSIVE_ABSORPTION)) {
          kind[e] = Node::Absorbed;
          element_list[e].clear();
        } else {
          from_elements += outside[e];
          kept.push_back(e);
        }
      }
      std::vector<size_t> neighbours;
      for (size_t j : variables[i])
        if (kind[j] == Node::Variable && in_pivot[j] != step) neighbours.push_back(j);
      size_t from_variables = 0;
      for (size_t j : neighbours) from_variables += weight[j];
      external[i] = from_variables + from_elements;
      std::sort(kept.begin(), kept.end());
      std::sort(neighbours.begin(), neighbours.end());
      uint64_t hash = 0;
      for (size_t x : kept) hash = hash * 31u + (uint64_t)x;
      for (size_t x : neighbours) hash = hash * 31u + (uint64_t)x;
      elements[i] = std::move(kept);
      variables[i] = std::move(neighbours);
      by_hash[(size_t)hash].push_back(i);
    }

    for (auto& entry : by_hash) {
      std::vector<size_t>& group = entry.second;
      if (group.size() <= 1) continue;
      if (mutations & LAST_REPRESENTATIVE) std::reverse(group.begin(), group.end());
      for (size_t a = 0; a < group.size(); ++a) {
        size_t i = group[a];
        if (kind[i] != Node::Variable) continue;
        for (size_t b = a + 1; b < group.size(); ++b) {
          size_t j = group[b];
          if (kind[j] == Node::Variable && elements[i] == elements[j] && variables[i] == variables[j]) {
            weight[i] += weight[j];
            weight[j] = 0;
            kind[j] = Node::Merged;
            std::vector<size_t> merged = std::move(members[j]);
            members[j].clear();
            members[i].push_back(j);
            members[i].insert(members[i].end(), merged.begin(), merged.end());
            elements[j].clear();
            variables[j].clear();
          }
        }
      }
    }

    std::vector<size_t> kept_list;
    for (size_t i : pivot_list)
      if (kind[i] == Node::Variable) kept_list.push_back(i);
    size_t size = 0;
    for (size_t i : kept_list) size += weight[i];
    element_size[p] = size;
    for (size_t i : kept_list) {
      size_t others = size - weight[i];
      size_t bound = std::min(std::min(remaining - weight[i], degree[i] + others), external[i] + others);
      degree[i] = bound;
      lists.insert(i, bound);
      min_degree = std::min(min_degree, bound);
    }
    element_list[p] = std::move(kept_list);
  }
  for (size_t i = 0; i < n; ++i)
    if (kind[i] == Node::Dense) order.push_back(i);
  return order;
}

// ---- COLAMD (colamd.rs) ----

namespace {
enum class Status { Live, Eliminated, Merged, SetAside };

// colamd.rs:35-87: score buckets, LIFO within a bucket, with a lowest-score pointer. With `fifo`
// (a planted mistake) a column goes to the tail of its bucket instead.
struct Buckets {
  std::vector<size_t> head, next, prev, score;
  size_t lowest;
  bool fifo = false;
  explicit Buckets(size_t n_cols)
      : head(n_cols + 1, NONE), next(n_cols, NONE), prev(n_cols, NONE), score(n_cols, 0), lowest(n_cols) {}
  void insert(size_t col, size_t s) {
    score[col] = s;
    lowest = std::min(lowest, s);
    if (fifo && head[s] != NONE) {
      size_t last = head[s];
      while (next[last] != NONE) last = next[last];
      next[last] = col;
      prev[col] = last;
      next[col] = NONE;
      return;
    }
    prev[col] = NONE;
    next[col] = head[s];
    if (head[s] != NONE) prev[head[s]] = col;
    head[s] = col;
    lowest = std::min(lowest, s);
  }
  void remove(size_t col) {
    size_t s = score[col];
    if (prev[col] != NONE)
      next[prev[col]] = next[col];
    else
      head[s] = next[col];
    if (next[col] != NONE) prev[next[col]] = prev[col];
  }
  size_t take_lowest() {
    while (lowest < head.size()) {
      if (head[lowest] != NONE) return head[lowest];
      lowest += 1;
    }
    return NONE;
  }
};
}  // namespace

// colamd.rs:91-261. Supercolumn detection keys candidates by (row count, row sum mod
// (n_cols + 1)) and merges a column into the first earlier representative with identical rows
// (:224-250); as with AMD the map's visiting order cannot change the outcome.
std::vector<size_t> column_approximate_minimum_degree(size_t n_rows, size_t n_cols,
                                                      const std::vector<std::pair<size_t, size_t>>& entries,
                                                      unsigned mutations) {
  std::vector<std::vector<size_t>> col_rows(n_cols);
  for (auto& e : entries) col_rows[e.second].push_back(e.first);
  for (auto& rows : col_rows) {
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
  }
  auto threshold = [](size_t n) { return std::max<size_t>(16, (size_t)(10.0 * std::sqrt((double)n))); };
  const size_t dense_col = threshold(std::min(n_rows, n_cols));
  const size_t dense_row = threshold(n_cols);
  std::vector<Status> status(n_cols, Status::Live);
  std::vector<size_t> dense_cols;
  for (size_t col = 0; col < n_cols; ++col) {
    if (col_rows[col].size() > dense_col) {
      status[col] = Status::SetAside;
      dense_cols.push_back(col);
    }
  }
  std::vector<std::vector<size_t>> row_cols(n_rows);
  for (size_t col = 0; col < n_cols; ++col)
    if (status[col] == Status::Live)
      for (size_t row : col_rows[col]) row_cols[row].push_back(col);
  std::vector<bool> row_alive(n_rows);
  for (size_t r = 0; r < n_rows; ++r) row_alive[r] = !row_cols[r].empty() && row_cols[r].size() <= dense_row;
  std::vector<size_t> null_cols;
  for (size_t col = 0; col < n_cols; ++col) {
    auto& rows = col_rows[col];
    rows.erase(std::remove_if(rows.begin(), rows.end(), [&](size_t r) { return !row_alive[r]; }), rows.end());
    if (status[col] == Status::Live && rows.empty()) {
      status[col] = Status::SetAside;
      null_cols.push_back(col);
    }
  }

  std::vector<size_t> weight(n_cols, 1);
  std::vector<std::vector<size_t>> members(n_cols);
  std::vector<size_t> row_size(n_rows);
  for (size_t r = 0; r < n_rows; ++r) row_size[r] = row_cols[r].size();
  size_t live_weight = 0;
  for (auto s : status)
    if (s == Status::Live) live_weight += 1;
  Buckets buckets(n_cols);
  buckets.fifo = (mutations & FIFO_LISTS) != 0;
  for (size_t col = 0; col < n_cols; ++col) {
    if (status[col] == Status::Live) {
      size_t score = 0;
      for (size_t row : col_rows[col]) score += row_size[row] - 1;
      buckets.insert(col, std::min(score, live_weight - 1));
    }
  }

  std::vector<size_t> order;
  order.reserve(n_cols);
  std::vector<size_t> col_mark(n_cols, 0), row_mark(n_rows, 0);
  std::vector<long long> outside(n_rows, 0);
  size_t stamp = 0;
  for (;;) {
    const size_t pivot = buckets.take_lowest();
    if (pivot == NONE) break;
    buckets.remove(pivot);
    status[pivot] = Status::Eliminated;
    live_weight -= weight[pivot];
    order.push_back(pivot);
    order.insert(order.end(), members[pivot].begin(), members[pivot].end());

    stamp += 1;
    std::vector<size_t> pivot_cols;
    for (size_t row : col_rows[pivot]) {
      if (!row_alive[row]) continue;
      for (size_t col : row_cols[row]) {
        if (status[col] == Status::Live && col_mark[col] != stamp) {
          col_mark[col] = stamp;
          pivot_cols.push_back(col);
        }
      }
      row_alive[row] = false;
    }
    const size_t p = row_cols.size();
    size_t psize = 0;
    for (size_t col : pivot_cols) psize += weight[col];
    row_size.push_back(psize);
    row_cols.push_back(pivot_cols);
    row_alive.push_back(true);
    row_mark.push_back(0);
    outside.push_back(0);
    for (size_t col : pivot_cols) {
      buckets.remove(col);
      auto& rows = col_rows[col];
      rows.erase(std::remove_if(rows.begin(), rows.end(), [&](size_t r) { return !row_alive[r]; }), rows.end());
      rows.push_back(p);
    }

    for (size_t col : pivot_cols) {
      for (size_t row : col_rows[col]) {
        if (row == p) continue;
        if (row_mark[row] != stamp) {
          row_mark[row] = stamp;
          outside[row] = (long long)row_size[row];
        }
        outside[row] -= (long long)weight[col];
      }
    }
    for (size_t col : pivot_cols)
      for (size_t row : col_rows[col])
        if (row != p && row_mark[row] == stamp && outside[row] == 0 && !(mutations & NO_AGGRESSIVE_ABSORPTION))
          row_alive[row] = false;
    std::vector<long long> scores;
    scores.reserve(pivot_cols.size());
    for (size_t col : pivot_cols) {
      auto& rows = col_rows[col];
      rows.erase(std::remove_if(rows.begin(), rows.end(), [&](size_t r) { return !row_alive[r]; }), rows.end());
      long long external = 0;
      for (size_t row : rows)
        if (row != p) external += outside[row];
      scores.push_back((long long)row_size[p] - (long long)weight[col] + external);
    }

    std::map<std::pair<size_t, size_t>, std::vector<size_t>> representatives;
    for (size_t index = 0; index < pivot_cols.size(); ++index) {
      size_t col = pivot_cols[index];
      std::sort(col_rows[col].begin(), col_rows[col].end());
      size_t sum = 0;
      for (size_t r : col_rows[col]) sum += r;
      auto key = std::make_pair(col_rows[col].size(), sum % (n_cols + 1));
      auto& candidates = representatives[key];
      size_t found = NONE, slot = 0;
      for (; slot < candidates.size(); ++slot)
        if (col_rows[pivot_cols[candidates[slot]]] == col_rows[col]) {
          found = candidates[slot];
          break;
        }
      if (found != NONE) {
        // merged into the earlier column (with LAST_REPRESENTATIVE, a planted mistake, the earlier
        // one into this one, which takes its place among the candidates)
        size_t keep = found, gone = index;
        if (mutations & LAST_REPRESENTATIVE) {
          keep = index;
          gone = found;
          candidates[slot] = index;
        }
        const size_t into = pivot_cols[keep], merged = pivot_cols[gone];
        weight[into] += weight[merged];
        std::vector<size_t> moved{merged};
        moved.insert(moved.end(), members[merged].begin(), members[merged].end());
        members[merged].clear();
        members[into].insert(members[into].end(), moved.begin(), moved.end());
        status[merged] = Status::Merged;
        scores[keep] -= (long long)weight[merged];
      } else {
        candidates.push_back(index);
      }
    }
    for (size_t k = 0; k < pivot_cols.size(); ++k) {
      size_t col = pivot_cols[k];
      if (status[col] == Status::Live) {
        size_t cap = live_weight - weight[col];
        size_t s = scores[k] > 0 ? (size_t)scores[k] : 0;
        buckets.insert(col, std::min(s, cap));
      }
    }
  }
  order.insert(order.end(), null_cols.begin(), null_cols.end());
  order.insert(order.end(), dense_cols.begin(), dense_cols.end());
  return order;
}

}  // namespace oracle
