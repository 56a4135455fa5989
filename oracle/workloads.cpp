// Synthetic workloads for the oracle (test infrastructure only): a platform-independent PRNG,
// the synthetic grid generator, and restatements of ybus_powerflow's B' (dc.rs:79-112,
// 368-407) and AC Newton Jacobian (ac.rs:55-174, 1309-1427) builders.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "workloads.hpp"

namespace oracle {

// ---- PRNG: splitmix64 seeding xoshiro256** ----

static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed, uint64_t stream) {
  uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull);
  for (auto& w : state) w = splitmix64(s);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
  const uint64_t result = rotl(state[1] * 5, 7) * 9;
  const uint64_t t = state[1] << 17;
  state[2] ^= state[0];
  state[3] ^= state[1];
  state[1] ^= state[2];
  state[0] ^= state[3];
  state[2] ^= t;
  state[3] = rotl(state[3], 45);
  return result;
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

double Rng::normal() {
  const double u1 = 1.0 - uniform();  // (0, 1]
  const double u2 = uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

size_t Rng::below(size_t n) { return (size_t)(next() % (uint64_t)n); }

// ---- grid generator ----

namespace {
struct Points {
  const std::vector<double>& x;
  const std::vector<double>& y;
  double d2(size_t i, size_t j) const {
    double dx = x[i] - x[j], dy = y[i] - y[j];
    return dx * dx + dy * dy;
  }
};

// The k nearest other buses of every bus (squared distance, then index), by a uniform cell grid
std::vector<std::vector<size_t>> nearest(const Points& pts, size_t n, size_t k) {
  const size_t cells = std::max<size_t>(1, (size_t)std::sqrt((double)n / 2.0));
  std::vector<std::vector<size_t>> grid(cells * cells);
  auto cell_of = [&](double v) { return std::min(cells - 1, (size_t)(v * (double)cells)); };
  for (size_t i = 0; i < n; ++i) grid[cell_of(pts.y[i]) * cells + cell_of(pts.x[i])].push_back(i);
  std::vector<std::vector<size_t>> out(n);
  std::vector<std::pair<double, size_t>> found;
  for (size_t i = 0; i < n; ++i) {
    const long cx = (long)cell_of(pts.x[i]), cy = (long)cell_of(pts.y[i]);
    for (long radius = 1;; ++radius) {
      found.clear();
      for (long gy = cy - radius; gy <= cy + radius; ++gy)
        for (long gx = cx - radius; gx <= cx + radius; ++gx) {
          if (gx < 0 || gy < 0 || gx >= (long)cells || gy >= (long)cells) continue;
          for (size_t j : grid[(size_t)gy * cells + (size_t)gx])
            if (j != i) found.emplace_back(pts.d2(i, j), j);
        }
      // Every bus within `radius - 1` cell widths has been seen, so the k-th nearest found is
      // exact once it lies within that distance (or the whole grid was searched)
      const double reach = (double)(radius - 1) / (double)cells;
      const bool whole = radius > (long)cells;
      if (found.size() >= k || whole) {
        std::sort(found.begin(), found.end());
        if (whole || (found.size() >= k && found[k - 1].first <= reach * reach)) {
          for (size_t q = 0; q < std::min(k, found.size()); ++q) out[i].push_back(found[q].second);
          break;
        }
      }
    }
  }
  return out;
}

struct Dsu {
  std::vector<size_t> p;
  explicit Dsu(size_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  size_t find(size_t a) {
    while (p[a] != a) a = p[a] = p[p[a]];
    return a;
  }
  bool unite(size_t a, size_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return false;
    if (a < b) std::swap(a, b);
    p[a] = b;
    return true;
  }
};

struct Edge {
  double d2;
  size_t i, j;
  bool operator<(const Edge& o) const {
    if (d2 != o.d2) return d2 < o.d2;
    if (i != o.i) return i < o.i;
    return j < o.j;
  }
};

std::vector<Edge> candidate_edges(const Points& pts, const std::vector<std::vector<size_t>>& nn) {
  std::vector<Edge> edges;
  for (size_t i = 0; i < nn.size(); ++i)
    for (size_t j : nn[i]) edges.push_back(Edge{pts.d2(i, j), std::min(i, j), std::max(i, j)});
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end(),
                          [](const Edge& a, const Edge& b) { return a.i == b.i && a.j == b.j; }),
              edges.end());
  return edges;
}
}  // namespace

// The topology: buses uniform in the unit square; a spanning tree as the minimum spanning
// forest of the 3-nearest-neighbour graph, its components joined by their nearest pairs; then
// the shortest unused 8-nearest-neighbour edges until there are round(1.35 N) branches. Each
// branch runs from its lower to its higher bus, and branches are listed in that order.
Grid make_grid(size_t n, uint64_t seed) {
  if (n < 2) throw SparseError(Err::InvalidArgument, "a grid needs at least two buses");
  Grid g;
  g.buses = n;
  Rng pos(seed, 1);
  g.bx.resize(n);
  g.by.resize(n);
  for (size_t i = 0; i < n; ++i) {
    g.bx[i] = pos.uniform();
    g.by[i] = pos.uniform();
  }
  Points pts{g.bx, g.by};
  const size_t k8 = std::min<size_t>(8, n - 1);
  auto nn8 = nearest(pts, n, k8);
  std::vector<std::vector<size_t>> nn3(n);
  for (size_t i = 0; i < n; ++i) nn3[i].assign(nn8[i].begin(), nn8[i].begin() + std::min<size_t>(3, nn8[i].size()));

  Dsu dsu(n);
  std::vector<std::pair<size_t, size_t>> chosen;
  for (const Edge& e : candidate_edges(pts, nn3))
    if (dsu.unite(e.i, e.j)) chosen.emplace_back(e.i, e.j);
  // Join components: the smallest (fewest buses, then lowest root) to its nearest outside bus
  for (;;) {
    std::vector<size_t> size(n, 0);
    for (size_t i = 0; i < n; ++i) size[dsu.find(i)] += 1;
    size_t best_root = SIZE_MAX;
    size_t roots = 0;
    for (size_t r = 0; r < n; ++r) {
      if (size[r] == 0) continue;
      roots += 1;
      if (best_root == SIZE_MAX || size[r] < size[best_root]) best_root = r;
    }
    if (roots <= 1) break;
    Edge best{INFINITY, 0, 0};
    for (size_t i = 0; i < n; ++i) {
      if (dsu.find(i) != best_root) continue;
      for (size_t j = 0; j < n; ++j) {
        if (dsu.find(j) == best_root) continue;
        Edge e{pts.d2(i, j), std::min(i, j), std::max(i, j)};
        if (e < best) best = e;
      }
    }
    dsu.unite(best.i, best.j);
    chosen.emplace_back(best.i, best.j);
  }
  const size_t target = (size_t)std::llround(1.35 * (double)n);
  std::vector<std::pair<size_t, size_t>> sorted_chosen = chosen;
  std::sort(sorted_chosen.begin(), sorted_chosen.end());
  for (const Edge& e : candidate_edges(pts, nn8)) {
    if (chosen.size() >= target) break;
    auto key = std::make_pair(e.i, e.j);
    if (std::binary_search(sorted_chosen.begin(), sorted_chosen.end(), key)) continue;
    chosen.push_back(key);
    sorted_chosen.insert(std::upper_bound(sorted_chosen.begin(), sorted_chosen.end(), key), key);
  }
  std::sort(chosen.begin(), chosen.end());
  g.from.resize(chosen.size());
  g.to.resize(chosen.size());
  for (size_t k = 0; k < chosen.size(); ++k) {
    g.from[k] = chosen[k].first;
    g.to[k] = chosen[k].second;
  }
  // Reactance log-uniform in [0.005, 0.2] pu
  Rng xr(seed, 2);
  g.x.resize(chosen.size());
  const double lo = std::log(0.005), hi = std::log(0.2);
  for (double& x : g.x) x = std::exp(lo + xr.uniform() * (hi - lo));
  return g;
}

// Bridges (Tarjan's low-link, iterative): branch k is a bridge iff removing it disconnects
std::vector<bool> bridges(const Grid& g) {
  const size_t n = g.buses, m = g.from.size();
  std::vector<std::vector<std::pair<size_t, size_t>>> adj(n);
  for (size_t k = 0; k < m; ++k) {
    adj[g.from[k]].emplace_back(g.to[k], k);
    adj[g.to[k]].emplace_back(g.from[k], k);
  }
  std::vector<size_t> disc(n, SIZE_MAX), low(n, 0);
  std::vector<bool> bridge(m, false);
  size_t time = 0;
  struct Frame {
    size_t node, via, next;
  };
  for (size_t root = 0; root < n; ++root) {
    if (disc[root] != SIZE_MAX) continue;
    std::vector<Frame> stack{{root, SIZE_MAX, 0}};
    disc[root] = low[root] = time++;
    while (!stack.empty()) {
      Frame& f = stack.back();
      if (f.next < adj[f.node].size()) {
        auto [to, k] = adj[f.node][f.next++];
        if (k == f.via) continue;
        if (disc[to] == SIZE_MAX) {
          disc[to] = low[to] = time++;
          stack.push_back(Frame{to, k, 0});
        } else {
          low[f.node] = std::min(low[f.node], disc[to]);
        }
      } else {
        Frame done = f;
        stack.pop_back();
        if (!stack.empty()) {
          size_t up = stack.back().node;
          low[up] = std::min(low[up], low[done.node]);
          if (low[done.node] > disc[up]) bridge[done.via] = true;
        }
      }
    }
  }
  return bridge;
}

// dc.rs:79-112 (Bdc::new): Bbus stamped per branch as (f,f,b), (f,t,-b), (t,f,-b), (t,t,b) with
// b = 1 / x / tap (tap 1), and Bf as (k,f,b), (k,t,-b); then B' = Bbus over the unknowns
// (buses 1..N-1: bus 0 is the slack and there are no generators, so pv_pq() is the natural
// order, dc.rs:333-336), selected as factor_over does (dc.rs:368-372).
DcMatrices make_dc(const Grid& g, const std::vector<bool>* in_service) {
  const size_t nb = g.buses, nl = g.from.size();
  std::vector<Triplet> bbus, bf;
  std::vector<size_t> rows;
  for (size_t k = 0; k < nl; ++k) {
    if (in_service && !(*in_service)[k]) continue;
    const size_t f = g.from[k], t = g.to[k];
    const double b = 1.0 / g.x[k] / 1.0;
    const size_t row = rows.size();
    rows.push_back(k);
    bbus.push_back({f, f, b});
    bbus.push_back({f, t, -b});
    bbus.push_back({t, f, -b});
    bbus.push_back({t, t, b});
    bf.push_back({row, f, b});
    bf.push_back({row, t, -b});
  }
  DcMatrices m;
  m.branch_rows = rows;
  Compressed bbus_m = from_triplets(nb, nb, true, bbus);
  Compressed bf_m = from_triplets(rows.size(), nb, true, bf);
  std::vector<size_t> unknown(nb - 1);
  for (size_t i = 0; i + 1 < nb; ++i) unknown[i] = i + 1;
  std::vector<size_t> all_rows(rows.size());
  std::iota(all_rows.begin(), all_rows.end(), 0);
  m.bprime = select(bbus_m, unknown, unknown);
  m.bf = select(bf_m, all_rows, unknown);
  return m;
}

// ---- complex arithmetic as num-complex 0.4 defines it ----
namespace {
struct Cx {
  double re, im;
};
inline Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
inline Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx conj(Cx a) { return {a.re, -a.im}; }
inline Cx neg(Cx a) { return {-a.re, -a.im}; }
inline Cx div(Cx a, Cx b) {
  double norm_sqr = b.re * b.re + b.im * b.im;
  double re = a.re * b.re + a.im * b.im;
  double im = a.im * b.re - a.re * b.im;
  return {re / norm_sqr, im / norm_sqr};
}
inline Cx divr(Cx a, double r) { return {a.re / r, a.im / r}; }
inline double norm(Cx a) { return std::hypot(a.re, a.im); }
inline Cx from_polar(double r, double t) { return {r * std::cos(t), r * std::sin(t)}; }

struct CxCsr {
  size_t n;
  std::vector<size_t> ptr, idx;
  std::vector<Cx> data;
};

// compressed_matrix.rs:254-321 for complex values: grouped by row, sorted stably, summed in order
CxCsr cx_from_triplets(size_t n, const std::vector<std::pair<size_t, size_t>>& pos, const std::vector<Cx>& val) {
  std::vector<size_t> start(n + 1, 0);
  for (auto& p : pos) start[p.first + 1] += 1;
  for (size_t i = 0; i < n; ++i) start[i + 1] += start[i];
  std::vector<size_t> cursor = start;
  struct G {
    size_t minor;
    Cx v;
  };
  std::vector<G> grouped(pos.size());
  for (size_t k = 0; k < pos.size(); ++k) grouped[cursor[pos[k].first]++] = G{pos[k].second, val[k]};
  CxCsr m;
  m.n = n;
  m.ptr.assign(n + 1, 0);
  for (size_t r = 0; r < n; ++r) {
    std::stable_sort(grouped.begin() + start[r], grouped.begin() + start[r + 1],
                     [](const G& a, const G& b) { return a.minor < b.minor; });
    for (size_t q = start[r]; q < start[r + 1]; ++q) {
      if (m.idx.size() > m.ptr[r] && m.idx.back() == grouped[q].minor) {
        m.data.back() = add(m.data.back(), grouped[q].v);
      } else {
        m.idx.push_back(grouped[q].minor);
        m.data.push_back(grouped[q].v);
      }
    }
    m.ptr[r + 1] = m.idx.size();
  }
  return m;
}
}  // namespace

// The AC operating point and Newton Jacobian. r = x * U[0.1, 0.5], charging b ~ U[0, 0.05],
// 20% of the non-slack buses PV (a seeded Fisher-Yates draw), |V| ~ U[0.95, 1.05] and angles
// N(0, 0.1) rad (the slack's angle 0). Ybus follows makeYbus's stamps (ac.rs:55-174: four
// branch stamps, then one shunt stamp per bus); the Jacobian follows JacobianAssembly::new and
// refill (ac.rs:1311-1427) with pv_pq = PV buses then PQ buses (topology.rs:277-283).
AcJacobian make_jacobian(const Grid& g, uint64_t seed, double shift, uint64_t shift_seed) {
  const size_t nb = g.buses, nl = g.from.size();
  Rng ac(seed, 3);
  std::vector<double> r(nl), bc(nl);
  for (size_t k = 0; k < nl; ++k) r[k] = g.x[k] * (0.1 + 0.4 * ac.uniform());
  for (size_t k = 0; k < nl; ++k) bc[k] = 0.05 * ac.uniform();
  std::vector<size_t> others(nb - 1);
  std::iota(others.begin(), others.end(), 1);
  for (size_t i = others.size(); i > 1; --i) std::swap(others[i - 1], others[ac.below(i)]);
  const size_t npv = (size_t)std::llround(0.2 * (double)(nb - 1));
  std::vector<bool> is_pv(nb, false);
  for (size_t q = 0; q < npv; ++q) is_pv[others[q]] = true;
  AcJacobian out;
  out.r = r;
  out.bc = bc;
  for (size_t i = 1; i < nb; ++i)
    if (is_pv[i]) out.pv.push_back(i);
  for (size_t i = 1; i < nb; ++i)
    if (!is_pv[i]) out.pq.push_back(i);
  std::vector<Cx> v(nb);
  out.vm.resize(nb);
  out.va.resize(nb);
  for (size_t i = 0; i < nb; ++i) {
    out.vm[i] = 0.95 + 0.1 * ac.uniform();
    out.va[i] = i == 0 ? 0.0 : 0.1 * ac.normal();
  }
  if (shift > 0.0) {
    Rng step(shift_seed, 4);
    for (size_t i = 0; i < nb; ++i) {
      out.vm[i] += shift * (2.0 * step.uniform() - 1.0);
      const double da = shift * (2.0 * step.uniform() - 1.0);
      if (i != 0) out.va[i] += da;
    }
  }
  for (size_t i = 0; i < nb; ++i) v[i] = from_polar(out.vm[i], out.va[i]);
  // ac.rs:137-151 with tap 1 and no shift
  std::vector<std::pair<size_t, size_t>> positions;
  std::vector<Cx> stamps;
  for (size_t k = 0; k < nl; ++k) {
    const size_t f = g.from[k], t = g.to[k];
    Cx ys = div(Cx{1.0, 0.0}, Cx{r[k], g.x[k]});
    Cx tap = from_polar(1.0, 0.0 * 3.141592653589793 / 180.0);
    Cx ytt = add(ys, Cx{0.0, bc[k] / 2.0});
    Cx yff = div(ytt, mul(tap, conj(tap)));
    Cx yft = div(neg(ys), conj(tap));
    Cx ytf = div(neg(ys), tap);
    yff = add(yff, Cx{0.0, 0.0});  // no magnetizing admittance
    positions.insert(positions.end(), {{f, f}, {f, t}, {t, f}, {t, t}});
    stamps.insert(stamps.end(), {yff, yft, ytf, ytt});
  }
  for (size_t i = 0; i < nb; ++i) {
    positions.push_back({i, i});
    stamps.push_back(divr(Cx{0.0, 0.0}, 100.0));
  }
  CxCsr y = cx_from_triplets(nb, positions, stamps);
  std::vector<Cx> current(nb);
  for (size_t i = 0; i < nb; ++i) {
    Cx sum{0.0, 0.0};
    for (size_t p = y.ptr[i]; p < y.ptr[i + 1]; ++p) sum = add(sum, mul(y.data[p], v[y.idx[p]]));
    current[i] = sum;
  }
  std::vector<size_t> in_pvpq(nb, NONE), in_pq(nb, NONE);
  std::vector<size_t> pv_pq = out.pv;
  pv_pq.insert(pv_pq.end(), out.pq.begin(), out.pq.end());
  const size_t npvpq = pv_pq.size();
  for (size_t k = 0; k < npvpq; ++k) in_pvpq[pv_pq[k]] = k;
  for (size_t k = 0; k < out.pq.size(); ++k) in_pq[out.pq[k]] = k;
  enum Part { ReVa, ReVm, ImVa, ImVm };
  std::vector<Triplet> entries;
  for (size_t i = 0; i < nb; ++i) {
    for (size_t p = y.ptr[i]; p < y.ptr[i + 1]; ++p) {
      const size_t j = y.idx[p];
      struct B {
        size_t row, col;
        Part part;
      } blocks[4] = {
          {in_pvpq[i], in_pvpq[j], ReVa},
          {in_pvpq[i], in_pq[j] == NONE ? NONE : npvpq + in_pq[j], ReVm},
          {in_pq[i] == NONE ? NONE : npvpq + in_pq[i], in_pvpq[j], ImVa},
          {in_pq[i] == NONE ? NONE : npvpq + in_pq[i], in_pq[j] == NONE ? NONE : npvpq + in_pq[j], ImVm},
      };
      for (const B& b : blocks) {
        if (b.row == NONE || b.col == NONE) continue;
        const Cx entry = y.data[p];
        Cx derivative;
        if (b.part == ReVa || b.part == ImVa) {
          Cx inner = i == j ? sub(current[i], mul(entry, v[i])) : sub(Cx{0.0, 0.0}, mul(entry, v[j]));
          derivative = mul(mul(v[i], Cx{0.0, 1.0}), conj(inner));
        } else {
          Cx unit_j = divr(v[j], norm(v[j]));
          Cx scaled = mul(v[i], conj(mul(entry, unit_j)));
          derivative = i == j ? add(scaled, mul(conj(current[i]), unit_j)) : scaled;
        }
        double value = (b.part == ReVa || b.part == ReVm) ? derivative.re : derivative.im;
        entries.push_back({b.row, b.col, value});
      }
    }
  }
  const size_t n = npvpq + out.pq.size();
  out.jacobian = from_triplets(n, n, false, entries);
  return out;
}

}  // namespace oracle
