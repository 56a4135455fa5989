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

#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>

#include "kernels.hpp"

namespace ybus_hip {

namespace {

template <int TW>
struct Tile {
  int lane, slot, slots;
  int32_t col;
  bool on;
  __device__ Tile(int32_t ncols, int32_t tile = (int32_t)blockIdx.x) {
    lane = (int)threadIdx.x % TW;
    slot = (int)threadIdx.x / TW;
    slots = (int)blockDim.x / TW;
    col = tile * TW + lane;
    on = col < ncols;
  }
};

// The tile that workgroup blockIdx.x of a one-dimensional grid takes. The GPU hands workgroups
// to its XCDs in turn (workgroup i to XCD i mod xcds), and each XCD has an L2 of its own; tiles
// narrower than a 128-byte line share that line in every row of the block, so runs of
// consecutive tiles are given to the workgroups of one XCD (ceil(g / xcds) tiles to each of the
// first g mod xcds XCDs, floor to the rest). A bijection on [0, g); xcds = 1 is the identity.
__device__ __forceinline__ int32_t tile_of_workgroup(int32_t xcds) {
  const int32_t b = (int32_t)blockIdx.x, g = (int32_t)gridDim.x;
  if (xcds <= 1) return b;
  const int32_t xcd = b % xcds, idx = b / xcds, q = g / xcds, r = g % xcds;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + idx;
}

// Bytes of an n-entry vector in LDS, rounded up to 16 so what follows it stays aligned
__host__ __device__ __forceinline__ size_t lds_vector_bytes(int32_t n, size_t elem) {
  return ((size_t)n * elem + 15) / 16 * 16;
}

// A max that ignores NaN, as the reference's infinity-norm fold does (accuracy.rs:66-69)
__device__ __forceinline__ double fold_max(double m, double v) { return v > m ? v : m; }

// ||r||inf / ||b||inf as relative_residual forms it (accuracy.rs:57-78): 0 for a zero residual,
// infinite for a nonzero one against a zero b
__device__ __forceinline__ double ratio_of(double r_norm, double b_norm) {
  return b_norm > 0.0 ? r_norm / b_norm : (r_norm > 0.0 ? __builtin_inf() : 0.0);
}

// The stop rule's floor: twice the unit roundoff's share of ||A x|| that no FP64 residual gets
// below, 2 u || |A| |x| ||inf / ||b||inf (0 without the floor rule, or for b = 0)
__device__ __forceinline__ double floor_of(bool floor_rule, double ax_norm, double b_norm) {
  return floor_rule && b_norm > 0.0 ? 2.0 * 0x1p-53 * (ax_norm / b_norm) : 0.0;
}

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

// ---- wide tiles ----

// Element k of an array in global memory, loaded as global memory. A pointer the compiler cannot
// trace to a kernel argument (one read from a struct in memory, as a batch entry's) is loaded
// with flat instructions, which count against the LDS counter too: every wait on an LDS read
// would then also wait for all the record fetches in flight. Every array these kernels read
// this way is device memory.
template <typename T>
struct NativeOf {
  using type = T;
};
template <>
struct NativeOf<int4> {
  typedef int type __attribute__((ext_vector_type(4)));
};
template <>
struct NativeOf<float4> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct NativeOf<double2> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <typename T>
__device__ __forceinline__ T global_at(const T* p, int64_t k) {
  using N = typename NativeOf<T>::type;
  const N v = ((const __attribute__((address_space(1))) N*)p)[k];
  return __builtin_bit_cast(T, v);
}

// One row's record in registers
template <typename T>
struct Record {
  int32_t dst, src, p0, p1;
  int32_t dep[RECORD_ENTRIES];
  T val[RECORD_ENTRIES];
  T diag;
};

// p[k], as a global load where G (see global_at), else as the compiler chooses
template <bool G, typename P>
__device__ __forceinline__ P load_at(const P* p, int64_t k) {
  if constexpr (G)
    return global_at(p, k);
  else
    return p[k];
}

template <typename T, bool G = false>
__device__ __forceinline__ void fetch_values(const SweepDev& s, int32_t k, Record<T>& r) {
  if constexpr (std::is_same<T, double>::value) {
    const double2 a = load_at<G>(s.rec.val01, k), b = load_at<G>(s.rec.val23, k);
    r.val[0] = a.x;
    r.val[1] = a.y;
    r.val[2] = b.x;
    r.val[3] = b.y;
    r.diag = load_at<G>(s.rec.diag64, k);
  } else {
    const float4 a = load_at<G>(s.rec.val32, k);
    r.val[0] = a.x;
    r.val[1] = a.y;
    r.val[2] = a.z;
    r.val[3] = a.w;
    r.diag = load_at<G>(s.rec.diag32, k);
  }
}

// Issues the (independent, coalesced) loads of position k's record (G: as global loads)
template <typename T, bool G = false>
__device__ __forceinline__ void fetch(const SweepDev& s, int32_t k, Record<T>& r) {
  const int4 h = load_at<G>(s.rec.head, k), d = load_at<G>(s.rec.dep, k);
  r.dst = h.x;
  r.src = h.y;
  r.p0 = h.z;
  r.p1 = h.w;
  r.dep[0] = d.x;
  r.dep[1] = d.y;
  r.dep[2] = d.z;
  r.dep[3] = d.w;
  fetch_values<T, G>(s, k, r);
}

// Waits until every field of r has arrived. A record slot is refilled while the loads of its
// previous fetch may still be in flight (a thread with no row in that level never reads it); the
// compiler orders such a refill behind every load then in flight, so waiting on the slot's own
// loads first keeps the other slots' prefetches in flight.
template <typename T>
__device__ __forceinline__ void await_record(const Record<T>& r) {
  __asm__ volatile("" ::"v"(r.dst), "v"(r.src), "v"(r.p0), "v"(r.p1), "v"(r.dep[0]), "v"(r.dep[1]), "v"(r.dep[2]),
                   "v"(r.dep[3]), "v"(r.val[0]), "v"(r.val[1]), "v"(r.val[2]), "v"(r.val[3]), "v"(r.diag));
}

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

// Copies a sweep's level pointers into LDS (every thread must call; ends with a barrier)
__device__ void stage_levels(const SweepDev& s, int32_t* lp) {
  for (int32_t l = threadIdx.x; l <= s.levels; l += blockDim.x) lp[l] = s.level_ptr[l];
  __syncthreads();
}

// Terms of a row past its inline four are taken this many at a time: the chunk's dependency
// indices and values are loaded together, then its dependencies' values together, so a long row
// (a separator's near the top of a Cholesky schedule, most rows of an LU) waits for two round
// trips per chunk rather than per term.
constexpr int OVERFLOW_CHUNK = 8;

// A thread's record of position k of a level ending at `end`, clamped to the level's last row
// (so every fetch issues the same loads, and the wait counts the compiler places stay exact)
// `at` maps a level-list index to a record position (the identity when a tile sweeps every row;
// a tile's own list when it sweeps only some).
template <typename T, class At>
__device__ __forceinline__ void fetch_clamped(const SweepDev& s, int32_t k, int32_t end, Record<T>& r, At& at) {
  fetch<T>(s, at(k < end ? k : end - 1), r);
}

// The record a thread starts level m with (a harmless refetch of the last level past the end;
// a level may be empty in a tile's own list, and then the fetch takes the list's first entry)
template <typename T, class At>
__device__ __forceinline__ void fetch_first(const SweepDev& s, const int32_t* lp, int32_t m, int slot, Record<T>& r,
                                            At& at) {
  if (m >= s.levels) m = s.levels - 1;
  const int32_t begin = lp[m], end = lp[m + 1];
  const int32_t k = begin + slot;
  fetch<T>(s, at(k < end ? k : (end > begin ? end - 1 : lp[0])), r);
}

// Every dependency is live (a tile that sweeps every row)
struct AllLive {
  __device__ __forceinline__ bool operator()(int32_t) const { return true; }
};
struct Identity {
  __device__ __forceinline__ int32_t operator()(int32_t k) const { return k; }
};

// A row's right-hand side and inline dependencies, loaded (or in flight)
template <typename T>
struct Gathered {
  T rhs;
  T x[RECORD_ENTRIES];
};

// `live(d)` tells whether row d of dep_buf holds a value; a dead row reads as 0 (its load goes to
// the pad row, which every tile keeps in cache). With Skip (std::true_type) a row issues only the
// loads of the terms it has: the pad row's loads hit in cache but still take the vector memory
// path's bandwidth, which a sweep of short rows (about two terms each in a grid's Cholesky factor)
// is bound by; a sweep of long rows (an LU's) is faster issuing all four without the branches.
template <typename T, int TW, class Rhs, class Live, class Skip = std::false_type>
__device__ __forceinline__ void gather(const Record<T>& r, const T* __restrict__ dep_buf, int64_t ld, int32_t pad,
                                       const Tile<TW>& g, Rhs& rhs, Live& live, Gathered<T>& o, Skip = Skip()) {
  o.rhs = rhs(r.src);
  const int32_t len = r.p1 - r.p0;
#pragma unroll
  for (int i = 0; i < RECORD_ENTRIES; ++i) {
    if (Skip::value && len <= i) {
      o.x[i] = (T)0;
      continue;
    }
    const bool on = live(r.dep[i]);
    const T v = dep_buf[(on ? r.dep[i] : pad) * ld + g.col];
    o.x[i] = on ? v : (T)0;
  }
}

// A row's value: its right-hand side minus its terms in packed order (the inline four from o,
// the rest in chunks), over its diagonal
template <typename T, int TW, class Live>
__device__ __forceinline__ T finish(const SweepDev& s, const T* __restrict__ val, const Record<T>& r,
                                    const Gathered<T>& o, const T* __restrict__ dep_buf, int64_t ld, int32_t pad,
                                    const Tile<TW>& g, Live& live) {
  const int32_t len = r.p1 - r.p0;
  T a = o.rhs;
#pragma unroll
  for (int i = 0; i < RECORD_ENTRIES; ++i)
    if (len > i) a = a - r.val[i] * o.x[i];
  for (int32_t p = r.p0 + RECORD_ENTRIES; p < r.p1; p += OVERFLOW_CHUNK) {
    int32_t dd[OVERFLOW_CHUNK];
    T vv[OVERFLOW_CHUNK], xx[OVERFLOW_CHUNK];
#pragma unroll
    for (int i = 0; i < OVERFLOW_CHUNK; ++i) {
      const int32_t pi = p + i < r.p1 ? p + i : r.p1 - 1;
      dd[i] = s.dep[pi];
      vv[i] = val[pi];
    }
#pragma unroll
    for (int i = 0; i < OVERFLOW_CHUNK; ++i) {
      const bool on = live(dd[i]);
      const T v = dep_buf[(on ? dd[i] : pad) * ld + g.col];
      xx[i] = on ? v : (T)0;
    }
#pragma unroll
    for (int i = 0; i < OVERFLOW_CHUNK; ++i)
      if (p + i < r.p1) a = a - vv[i] * xx[i];
  }
  return s.unit ? a : a / r.diag;
}

// One level of a tile: rows begin + slot, + slots, ... from the thread's prefetched first
// record. The rounds are software-pipelined: a row's gathers are issued, then the previous row's
// value is stored, then the next row's record is fetched, and only then does the thread wait,
// for the gathers. A wait for a load waits for every memory operation issued before it, stores
// included, so storing after the next gathers keeps the store's latency off the wait; the
// records alternate between two registers sets, so no register copy waits on loads in flight.
template <typename T, int TW, class Rhs, class Store, class At, class Live, class Skip>
__device__ __forceinline__ void wide_level(const SweepDev& s, const T* __restrict__ val, int32_t begin, int32_t end,
                                           const T* __restrict__ dep_buf, int64_t ld, const Tile<TW>& g,
                                           const Record<T>& first, Rhs& rhs, Store& store, At& at, Live& live,
                                           Skip skip) {
  const int32_t pad = s.n;
  int32_t k = begin + g.slot;
  if (k >= end) return;
  const int32_t step = g.slots;
  Gathered<T> o;
  gather<T, TW>(first, dep_buf, ld, pad, g, rhs, live, o, skip);
  if (k + step >= end) {
    store(first.dst, finish<T, TW>(s, val, first, o, dep_buf, ld, pad, g, live));
    return;
  }
  Record<T> b, c;
  fetch_clamped<T>(s, k + step, end, b, at);
  T v = finish<T, TW>(s, val, first, o, dep_buf, ld, pad, g, live);
  int32_t dst = first.dst;
  for (k += step;; k += 2 * step) {
    // b holds row k
    gather<T, TW>(b, dep_buf, ld, pad, g, rhs, live, o, skip);
    store(dst, v);
    fetch_clamped<T>(s, k + step, end, c, at);
    v = finish<T, TW>(s, val, b, o, dep_buf, ld, pad, g, live);
    dst = b.dst;
    if (k + step >= end) break;
    // c holds row k + step
    gather<T, TW>(c, dep_buf, ld, pad, g, rhs, live, o, skip);
    store(dst, v);
    await_record<T>(b);
    fetch_clamped<T>(s, k + 2 * step, end, b, at);
    v = finish<T, TW>(s, val, c, o, dep_buf, ld, pad, g, live);
    dst = c.dst;
    if (k + 2 * step >= end) break;
    await_record<T>(c);
  }
  store(dst, v);
}

// One sweep of a tile: every row (as its record gives it) is rhs(src) minus its terms against
// dep_buf (the destination block), divided by its diagonal, and handed to store(dst, value).
// Lane (slot, c) takes rows lp[l] + slot + j slots of level l. The first record of each level
// is fetched two levels ahead (two records alternate, so no register copies wait on loads in
// flight). A barrier ends each level: it orders the global stores of a workgroup's waves, which
// share one CU's L1. Every thread takes every barrier; `act` only masks the rows.
template <typename T, int TW, class Rhs, class Store, class At = Identity, class Live = AllLive,
          class Skip = std::false_type>
__device__ void wide_sweep(const SweepDev& s, const int32_t* lp, const T* __restrict__ dep_buf, int64_t ld,
                           const Tile<TW>& g, bool act, Rhs rhs, Store store, At at = At(), Live live = Live(),
                           Skip skip = Skip()) {
  const T* __restrict__ val = values_of<T>(s);
  const int32_t levels = s.levels;
  if (levels == 0 || lp[levels] == lp[0]) return;
  Record<T> a, b;
  fetch_first<T>(s, lp, 0, g.slot, a, at);
  fetch_first<T>(s, lp, 1, g.slot, b, at);
  for (int32_t l = 0; l < levels; l += 2) {
    if (act) wide_level<T, TW>(s, val, lp[l], lp[l + 1], dep_buf, ld, g, a, rhs, store, at, live, skip);
    __syncthreads();
    await_record<T>(a);
    fetch_first<T>(s, lp, l + 2, g.slot, a, at);
    if (l + 1 >= levels) break;
    if (act) wide_level<T, TW>(s, val, lp[l + 1], lp[l + 2], dep_buf, ld, g, b, rhs, store, at, live, skip);
    __syncthreads();
    await_record<T>(b);
    fetch_first<T>(s, lp, l + 3, g.slot, b, at);
  }
}

// ---- FP64 solve: forward into t, barrier, backward into x ----

// Split: the sweeps are split (SweepDev::split), so a segment with src < 0 starts from the
// partial value in its row's own slot of the block it writes (t forward, x backward).
template <int TW, bool Split>
__global__ void __launch_bounds__(1024) solve64_kernel(SweepDev fwd, SweepDev bwd, RhsDev rhs, double* __restrict__ t,
                                                       double* __restrict__ x, int64_t ld, int32_t ncols,
                                                       int32_t xcds) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lpf = reinterpret_cast<int32_t*>(smem);
  int32_t* lpb = lpf + fwd.levels + 1;
  stage_levels(fwd, lpf);
  stage_levels(bwd, lpb);
  const Tile<TW> g(ncols, tile_of_workgroup(xcds));
  const int32_t col = g.col;
  const int32_t plus = !rhs.block && g.on ? rhs.plus[col] : -1;
  const int32_t minus = !rhs.block && g.on ? rhs.minus[col] : -1;
  auto from_rhs = [&](int32_t src) -> double {
    if (Split && src < 0) return t[(~src) * ld + col];
    if (rhs.block) return rhs.block[src * ld + col];
    // e_plus - e_minus as the host builds it: 1.0, -1.0 or 0.0
    return (src == plus ? 1.0 : 0.0) - (src == minus ? 1.0 : 0.0);
  };
  wide_sweep<double, TW>(fwd, lpf, t, ld, g, g.on, from_rhs, [&](int32_t dst, double v) { t[dst * ld + col] = v; });
  wide_sweep<double, TW>(bwd, lpb, x, ld, g, g.on,
                         [&](int32_t src) { return Split && src < 0 ? x[(~src) * ld + col] : t[src * ld + col]; },
                         [&](int32_t dst, double v) { x[dst * ld + col] = v; });
}

// ---- FP64 solve with the last positions' values in LDS ----

// A source as window_code encodes it (the pad row, a row that is never written, is old)
constexpr int32_t WINDOW_PARTIAL_SOURCE = INT32_MIN;

// A thread's position of a level for the windowed sweep (k < 0: none there): its window head and
// inline sources, the sweep's values and diagonal, and the values loaded ahead for it (its
// right-hand side, then its inline terms that are old)
struct WinRow {
  int32_t k;
  int4 h, d;
  double2 v01, v23;
  double diag;
  double pre[RECORD_ENTRIES + 1];
};

// Position k's record if k < end (else position 0's, and k = -1)
__device__ __forceinline__ void win_fetch(const SweepDev& s, const WindowDev& w, int32_t k, int32_t end, WinRow& x) {
  const bool on = k < end;
  const int32_t at = on ? k : 0;
  x.k = on ? k : -1;
  x.h = w.head[at];
  x.d = w.dep4[at];
  x.v01 = s.rec.val01[at];
  x.v23 = s.rec.val23[at];
  x.diag = s.rec.diag64[at];
}

// Loads x's right-hand side and its old inline terms (their rows are final by the start of the
// level before x's), where the thread has the position and its lane a column
template <int TW, class Rhs>
__device__ __forceinline__ void win_load_ahead(WinRow& x, const double* __restrict__ dep_buf, int64_t ld, int32_t n,
                                               const Tile<TW>& g, Rhs& rhs) {
  const bool go = x.k >= 0 && g.on;
  x.pre[0] = go && x.h.y >= 0 ? rhs(x.h.y) : 0.0;
  const int32_t dd[RECORD_ENTRIES] = {x.d.x, x.d.y, x.d.z, x.d.w};
#pragma unroll
  for (int i = 0; i < RECORD_ENTRIES; ++i) {
    const bool old = go && dd[i] >= 0 && dd[i] <= n;
    const double v = dep_buf[(int64_t)(old ? dd[i] : n) * ld + g.col];
    x.pre[1 + i] = old ? v : 0.0;
  }
}

// x's value: its source minus its terms in order, over its diagonal, each value from the ring,
// from what was loaded ahead (ahead: x's old terms and right-hand side were), or from the block.
// The block's loads (a first row's rare late terms or partial value; every old one of a row not
// loaded ahead) are issued and waited for inside their own branch: a wait placed after it would
// also wait for the loads issued ahead for the next levels, which come after x's in the queue.
template <int TW, class Rhs>
__device__ __forceinline__ double win_value(const SweepDev& s, const WindowDev& w, const WinRow& x, bool ahead,
                                            const double* ring, const double* __restrict__ dep_buf, int64_t ld,
                                            int32_t n, const Tile<TW>& g, Rhs& rhs) {
  const int32_t len = x.h.w - x.h.z;
  const int32_t dd[RECORD_ENTRIES] = {x.d.x, x.d.y, x.d.z, x.d.w};
  const double vv[RECORD_ENTRIES] = {x.v01.x, x.v01.y, x.v23.x, x.v23.y};
  double a = x.h.y >= 0 ? (ahead ? x.pre[0] : 0.0) : (x.h.y != WINDOW_PARTIAL_SOURCE ? ring[(~x.h.y) * TW + g.lane] : 0.0);
  double xv[RECORD_ENTRIES];
  bool block = x.h.y == WINDOW_PARTIAL_SOURCE || (x.h.y >= 0 && !ahead);
#pragma unroll
  for (int i = 0; i < RECORD_ENTRIES; ++i) {
    xv[i] = dd[i] < 0 ? ring[(~dd[i]) * TW + g.lane] : (ahead && dd[i] <= n ? x.pre[1 + i] : 0.0);
    block = block || (len > i && dd[i] >= 0 && (dd[i] > n || !ahead));
  }
  if (block) {
    if (x.h.y >= 0 && !ahead) a = rhs(x.h.y);
    if (x.h.y == WINDOW_PARTIAL_SOURCE) a = dep_buf[(int64_t)x.h.x * ld + g.col];
#pragma unroll
    for (int i = 0; i < RECORD_ENTRIES; ++i)
      if (len > i && dd[i] >= 0 && (dd[i] > n || !ahead))
        xv[i] = dep_buf[(int64_t)(dd[i] > n ? dd[i] - n - 1 : dd[i]) * ld + g.col];
    __asm__ volatile("" ::"v"(a), "v"(xv[0]), "v"(xv[1]), "v"(xv[2]), "v"(xv[3]));
  }
#pragma unroll
  for (int i = 0; i < RECORD_ENTRIES; ++i)
    if (len > i) a = a - vv[i] * xv[i];
  for (int32_t p = x.h.z + RECORD_ENTRIES; p < x.h.w; ++p) {
    const int32_t e = w.dep[p];
    a = a - s.val64[p] * (e < 0 ? ring[(~e) * TW + g.lane] : dep_buf[(int64_t)(e > n ? e - n - 1 : e) * ld + g.col]);
  }
  return s.unit ? a : a / x.diag;
}

// One sweep of a tile with its last w.cap positions' values in the LDS ring (slot q mod cap, the
// tile's TW columns side by side). Level l: the thread fetches its first position of level l + 3,
// loads the old values of its first position of level l + 2 (final since the barrier before
// level l: the encoding's old values are three levels back), then sweeps its positions of level
// l (the first from the values loaded at level l - 2; the second fetched at the level's start,
// before the loads ahead, so waiting for it does not wait for them; any further ones after),
// storing each value into the block and its ring slot, and a barrier ends the level. Four row
// sets take the four roles in turn (the loop unrolled by four), so no register copy waits on
// loads in flight. A ring slot is written by position k only after every reader of the position
// it held (k - cap) is done: a reader of position q at level l has q >= the level's end - cap >
// k - cap for every k it sweeps beside. Every row is computed as wide_level computes it, so the
// bits are the same.
template <int TW, class Rhs, class Store>
__device__ void window_sweep(const SweepDev& s, const WindowDev& w, const int32_t* lp, double* ring,
                             const double* __restrict__ dep_buf, int64_t ld, const Tile<TW>& g, bool act, Rhs rhs,
                             Store store) {
  const int32_t levels = s.levels, n = s.n, cap = w.cap, slots = g.slots;
  if (levels == 0 || lp[levels] == 0) return;
  auto end_of = [&](int32_t l) { return l < levels ? lp[l + 1] : 0; };
  auto first_of = [&](int32_t l) { return l < levels ? lp[l] + g.slot : 0; };
  auto sweep_row = [&](const WinRow& x, bool ahead) {
    const double v = win_value<TW>(s, w, x, ahead, ring, dep_buf, ld, n, g, rhs);
    store(x.h.x, v);
    ring[(x.k % cap) * TW + g.lane] = v;
  };
  auto level = [&](int32_t l, WinRow& a, WinRow& c, WinRow& d) {
    const int32_t end = end_of(l);
    WinRow e;
    win_fetch(s, w, a.k >= 0 ? a.k + slots : end, end, e);
    win_fetch(s, w, first_of(l + 3), end_of(l + 3), d);
    win_load_ahead<TW>(c, dep_buf, ld, n, g, rhs);
    if (act && a.k >= 0) {
      sweep_row(a, true);
      if (e.k >= 0) {
        sweep_row(e, false);
        for (int32_t k = e.k + slots; k < end; k += slots) {
          WinRow f;
          win_fetch(s, w, k, end, f);
          sweep_row(f, false);
        }
      }
    }
    __syncthreads();
  };
  // a: level 0 (its values loaded now), b: level 1 (loaded now), c: level 2, d: fetched at level 0
  WinRow a, b, c, d;
  win_fetch(s, w, first_of(0), end_of(0), a);
  win_fetch(s, w, first_of(1), end_of(1), b);
  win_fetch(s, w, first_of(2), end_of(2), c);
  win_load_ahead<TW>(a, dep_buf, ld, n, g, rhs);
  win_load_ahead<TW>(b, dep_buf, ld, n, g, rhs);
  for (int32_t l = 0; l < levels; l += 4) {
    level(l, a, c, d);
    if (l + 1 >= levels) break;
    level(l + 1, b, d, a);
    if (l + 2 >= levels) break;
    level(l + 2, c, a, b);
    if (l + 3 >= levels) break;
    level(l + 3, d, b, c);
  }
}

template <int TW>
__global__ void __launch_bounds__(512) solve64_window_kernel(SweepDev fwd, WindowDev wf, SweepDev bwd, WindowDev wb,
                                                              RhsDev rhs, double* __restrict__ t,
                                                              double* __restrict__ x, int64_t ld, int32_t ncols,
                                                              int32_t xcds) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* ring = reinterpret_cast<double*>(smem);
  int32_t* lpf = reinterpret_cast<int32_t*>(smem + (size_t)wf.cap * TW * sizeof(double));
  int32_t* lpb = lpf + fwd.levels + 1;
  for (int32_t l = threadIdx.x; l <= fwd.levels; l += blockDim.x) lpf[l] = fwd.level_ptr[l];
  stage_levels(bwd, lpb);
  const Tile<TW> g(ncols, tile_of_workgroup(xcds));
  const int32_t col = g.col;
  const int32_t plus = !rhs.block && g.on ? rhs.plus[col] : -1;
  const int32_t minus = !rhs.block && g.on ? rhs.minus[col] : -1;
  auto from_rhs = [&](int32_t src) -> double {
    if (rhs.block) return rhs.block[(int64_t)src * ld + col];
    return (src == plus ? 1.0 : 0.0) - (src == minus ? 1.0 : 0.0);
  };
  window_sweep<TW>(fwd, wf, lpf, ring, t, ld, g, g.on, from_rhs,
                   [&](int32_t dst, double v) { t[(int64_t)dst * ld + col] = v; });
  window_sweep<TW>(bwd, wb, lpb, ring, x, ld, g, g.on, [&](int32_t src) { return t[(int64_t)src * ld + col]; },
                   [&](int32_t dst, double v) { x[(int64_t)dst * ld + col] = v; });
}

// A pair block whose forward sweep visits only each tile's reach: the rows of the forward
// sweep that depend, directly or through other rows, on one of the tile's plus or minus
// unknowns. Every other forward row of the tile is exactly +0.0 in the CPU's solve (its
// right-hand side is +0.0, its terms are products with +0.0, and its diagonal is positive or
// unit), so it is neither computed nor stored: a bitmap in LDS marks the reach, a dependency
// outside it reads as 0, and the backward sweep reads its right-hand side only inside it.
// SkipLoads: each row issues only its own terms' loads (a factor of short rows: a grid's
// Cholesky factor), or all four inline loads without branches (an LU's long rows)
template <int TW, bool SkipLoads>
__global__ void __launch_bounds__(1024) solve64_reach_kernel(SweepDev fwd, SweepDev bwd, RhsDev rhs, ReachDev reach,
                                                             double* __restrict__ t, double* __restrict__ x, int64_t ld,
                                                             int32_t ncols, int32_t xcds) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lpf = reinterpret_cast<int32_t*>(smem);
  int32_t* lpb = lpf + fwd.levels + 1;
  uint32_t* live = reinterpret_cast<uint32_t*>(lpb + bwd.levels + 1);
  const int32_t words = fwd.n / 32 + 1;  // a bit for every row and for the pad row (always clear)
  const int32_t tile = tile_of_workgroup(xcds);
  const int32_t* __restrict__ tile_lp = reach.lp + (int64_t)tile * (fwd.levels + 1);
  for (int32_t l = threadIdx.x; l <= fwd.levels; l += blockDim.x) lpf[l] = tile_lp[l];
  for (int32_t l = threadIdx.x; l <= bwd.levels; l += blockDim.x) lpb[l] = bwd.level_ptr[l];
  for (int32_t w = threadIdx.x; w < words; w += blockDim.x) live[w] = 0u;
  __syncthreads();
  for (int32_t k = tile_lp[0] + threadIdx.x; k < tile_lp[fwd.levels]; k += blockDim.x) {
    const int32_t row = fwd.dst[reach.pos[k]];
    atomicOr(&live[row >> 5], 1u << (row & 31));
  }
  __syncthreads();
  const Tile<TW> g(ncols, tile);
  const int32_t col = g.col;
  const int32_t plus = g.on ? rhs.plus[col] : -1;
  const int32_t minus = g.on ? rhs.minus[col] : -1;
  const int32_t* __restrict__ pos = reach.pos;
  auto at = [&](int32_t k) { return pos[k]; };
  auto in_reach = [&](int32_t row) { return (live[row >> 5] >> (row & 31)) & 1u; };
  auto from_pair = [&](int32_t src) -> double {
    return (src == plus ? 1.0 : 0.0) - (src == minus ? 1.0 : 0.0);
  };
  using Skip = std::integral_constant<bool, SkipLoads>;
  wide_sweep<double, TW>(fwd, lpf, t, ld, g, g.on, from_pair, [&](int32_t dst, double v) { t[dst * ld + col] = v; },
                         at, in_reach, Skip());
  const int32_t pad = fwd.n;
  auto from_t = [&](int32_t src) -> double {
    const bool on = in_reach(src);
    const double v = t[(on ? src : pad) * ld + col];
    return on ? v : 0.0;
  };
  wide_sweep<double, TW>(bwd, lpb, x, ld, g, g.on, from_t, [&](int32_t dst, double v) { x[dst * ld + col] = v; },
                         Identity(), AllLive(), Skip());
}

// ---- complex solves ----

// Complex arithmetic rounded as num-complex 0.4 writes it (the reference's element type):
// a * b = (a.re b.re - a.im b.im, a.re b.im + a.im b.re), and
// a / b = ((a.re b.re + a.im b.im) / s, (a.im b.re - a.re b.im) / s) with s = b.re b.re + b.im b.im.
// Each product and sum is rounded on its own (the file is built without contraction). The FP32
// sweeps of a refined solve use the same formulas in float.
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  const double s = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / s, (a.y * b.x - a.x * b.y) / s);
}
__device__ __forceinline__ float2 zsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 zmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 zdiv(float2 a, float2 b) {
  const float s = b.x * b.x + b.y * b.y;
  return make_float2((a.x * b.x + a.y * b.y) / s, (a.y * b.x - a.x * b.y) / s);
}
// |z| as num-complex's norm() takes it: the hypotenuse of its parts
__device__ __forceinline__ double zabs(double2 v) { return hypot(v.x, v.y); }

// A complex value of a block of E (double or float) values
template <typename E>
struct Cx;
template <>
struct Cx<double> {
  using T = double2;
  static __device__ __forceinline__ T make(double re, double im) { return make_double2(re, im); }
  static __device__ __forceinline__ const T* val(const ComplexSweepDev& s) { return s.val; }
  static __device__ __forceinline__ const T* diag(const ComplexSweepDev& s) { return s.diag; }
};
template <>
struct Cx<float> {
  using T = float2;
  static __device__ __forceinline__ T make(float re, float im) { return make_float2(re, im); }
  static __device__ __forceinline__ const T* val(const ComplexSweepDev& s) { return s.val32; }
  static __device__ __forceinline__ const T* diag(const ComplexSweepDev& s) { return s.diag32; }
};

// Unknown `row` of column `col` in a complex block (real parts in row 2 row, imaginary in 2 row + 1)
template <typename E>
__device__ __forceinline__ typename Cx<E>::T zload(const E* __restrict__ blk, int32_t row, int64_t ld, int32_t col) {
  const int64_t at = 2 * (int64_t)row * ld + col;
  return Cx<E>::make(blk[at], blk[at + ld]);
}
template <typename E>
__device__ __forceinline__ void zstore(E* __restrict__ blk, int32_t row, int64_t ld, int32_t col,
                                       typename Cx<E>::T v) {
  const int64_t at = 2 * (int64_t)row * ld + col;
  blk[at] = v.x;
  blk[at + ld] = v.y;
}

// Terms of a complex row taken at a time: their indices and values are loaded together, then
// their dependencies together, so a row waits for two round trips per chunk rather than per term
constexpr int COMPLEX_CHUNK = 4;

// One complex sweep of a tile, row by row as forward_by_rows / back_by_rows and the LU
// substitutions do (cholesky.rs:468-501, lu.rs:502-533,640-649,770-785): the row's right-hand
// side, minus each term's value times its dependency in stored order, divided by the diagonal
// unless it is unit. E is the block's (and the arithmetic's) precision: double for the FP64
// solve, float for a refined solve's sweeps over the factor's FP32 copy. Lane (slot, c) takes
// rows lp[l] + slot + j slots of level l when `act`, loading the next row's header while it
// works on one; a barrier ends each level, and every thread takes every barrier.
template <int TW, typename E, class Rhs, class Store>
__device__ void complex_sweep(const ComplexSweepDev& s, const int32_t* lp, const E* __restrict__ dep_buf, int64_t ld,
                              const Tile<TW>& g, bool act, Rhs rhs, Store store) {
  using C = typename Cx<E>::T;
  const int4* __restrict__ head = s.head;
  const int32_t* __restrict__ dep = s.dep;
  const C* __restrict__ val = Cx<E>::val(s);
  const C* __restrict__ dg = Cx<E>::diag(s);
  for (int32_t l = 0; l < s.levels; ++l) {
    const int32_t kb = lp[l] + g.slot, ke = lp[l + 1];
    if (act && kb < ke) {
      int4 h = head[kb];
      for (int32_t k = kb; k < ke; k += g.slots) {
        const int4 next = head[k + g.slots < ke ? k + g.slots : k];
        const C diag = s.diag ? dg[k] : Cx<E>::make(1, 0);
        C acc = rhs(h.y);
        for (int32_t p = h.z; p < h.w; p += COMPLEX_CHUNK) {
          int32_t d[COMPLEX_CHUNK];
          C v[COMPLEX_CHUNK], x[COMPLEX_CHUNK];
#pragma unroll
          for (int q = 0; q < COMPLEX_CHUNK; ++q) {
            const int32_t at = p + q < h.w ? p + q : p;  // past the row's end: reload its first, unused
            d[q] = dep[at];
            v[q] = val[at];
          }
          // only the row's own terms are gathered: a reload past its end would hit in cache but
          // still take the vector memory path's bandwidth
#pragma unroll
          for (int q = 0; q < COMPLEX_CHUNK; ++q)
            x[q] = p + q < h.w ? zload<E>(dep_buf, d[q], ld, g.col) : Cx<E>::make(0, 0);
#pragma unroll
          for (int q = 0; q < COMPLEX_CHUNK; ++q)
            if (p + q < h.w) acc = zsub(acc, zmul(v[q], x[q]));
        }
        if (s.diag) acc = zdiv(acc, diag);
        store(h.x, acc);
        h = next;
      }
    }
    __syncthreads();
  }
}

template <int TW>
__global__ void __launch_bounds__(1024) solve_complex_kernel(ComplexSweepDev fwd, ComplexSweepDev bwd,
                                                             const double* __restrict__ b, double* __restrict__ t,
                                                             double* __restrict__ x, int64_t ld, int32_t ncols,
                                                             int32_t conjugate, int32_t xcds) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lpf = reinterpret_cast<int32_t*>(smem);
  int32_t* lpb = lpf + fwd.levels + 1;
  for (int32_t l = threadIdx.x; l <= fwd.levels; l += blockDim.x) lpf[l] = fwd.level_ptr[l];
  for (int32_t l = threadIdx.x; l <= bwd.levels; l += blockDim.x) lpb[l] = bwd.level_ptr[l];
  __syncthreads();
  const Tile<TW> g(ncols, tile_of_workgroup(xcds));
  const int32_t col = g.col;
  complex_sweep<TW, double>(
      fwd, lpf, t, ld, g, g.on,
      [&](int32_t src) {
        double2 v = zload<double>(b, src, ld, col);
        if (conjugate) v.y = -v.y;
        return v;
      },
      [&](int32_t dst, double2 v) { zstore<double>(t, dst, ld, col, v); });
  complex_sweep<TW, double>(
      bwd, lpb, x, ld, g, g.on, [&](int32_t src) { return zload<double>(t, src, ld, col); },
      [&](int32_t dst, double2 v) { zstore<double>(x, dst, ld, col, v); });
  // the backward sweep read its own results unconjugated; each lane now conjugates its column
  if (conjugate && g.on)
    for (int32_t i = g.slot; i < bwd.n; i += g.slots) x[(2 * (int64_t)i + 1) * ld + col] = -x[(2 * (int64_t)i + 1) * ld + col];
}

// ---- complex FP32 solves refined in FP64 ----

// One FP32 solve of a tile's active columns: src (an FP64 complex block: the right-hand sides or
// the residuals) rounded to float as its rows are read, forward into t, backward into d (both
// FP32 complex blocks), the sweeps over the factor's FP32 values with FP32 accumulation
template <int TW>
__device__ void zsolve32(const ComplexSweepDev& fwd, const ComplexSweepDev& bwd, const int32_t* lpf,
                         const int32_t* lpb, const double* __restrict__ src, float* __restrict__ t,
                         float* __restrict__ d, int64_t ld, const Tile<TW>& g, bool act) {
  const int32_t col = g.col;
  complex_sweep<TW, float>(
      fwd, lpf, t, ld, g, act,
      [&](int32_t row) {
        const double2 v = zload<double>(src, row, ld, col);
        return make_float2((float)v.x, (float)v.y);
      },
      [&](int32_t dst, float2 v) { zstore<float>(t, dst, ld, col, v); });
  complex_sweep<TW, float>(
      bwd, lpb, d, ld, g, act, [&](int32_t row) { return zload<float>(t, row, ld, col); },
      [&](int32_t dst, float2 v) { zstore<float>(d, dst, ld, col, v); });
}

// r = b - A x over the tile's active columns in FP64 (CsrMatrix::mul_vec's order: each row summed
// from zero over its entries in index order, product.rs:17-27, then subtracted from b); returns
// the thread's largest |r_i|
// (ax non-null: also the thread's partial max of sum |a_ij| |x_j| into *ax)
template <int TW>
__device__ double zresidual(const ZCsrDev& a, const double* __restrict__ b, const double* __restrict__ x,
                            double* __restrict__ r, int64_t ld, const Tile<TW>& g, bool act, double* ax = nullptr) {
  double m = 0.0, axm = 0.0;
  if (ax) *ax = 0.0;
  if (!act) return m;
  for (int32_t i = g.slot; i < a.rows; i += g.slots) {
    double2 acc = make_double2(0.0, 0.0);
    double abs_sum = 0.0;
    for (int32_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
      const double2 xj = zload<double>(x, a.col[p], ld, g.col);
      acc = zadd(acc, zmul(a.val[p], xj));
      if (ax) abs_sum = abs_sum + zabs(a.val[p]) * zabs(xj);
    }
    const double2 ri = zsub(zload<double>(b, i, ld, g.col), acc);
    zstore<double>(r, i, ld, g.col, ri);
    m = fold_max(m, zabs(ri));
    axm = fold_max(axm, abs_sum);
  }
  if (ax) *ax = axm;
  return m;
}

// FP32 complex sweeps refined in FP64 by refine_normwise's rule (accuracy.rs:83-122), one tile
// of TW columns per workgroup: x0 = solve(b), then while a column is active, d = solve(r) and
// x + d into the other iterate buffer; a column stays active while each step at least halves
// its ratio ||b - A x||inf / ||b||inf (|z| its hypotenuse), and its best iterate ends in x0.
template <int TW>
__global__ void __launch_bounds__(1024) refine_complex_kernel(ComplexSweepDev fwd, ComplexSweepDev bwd, ZCsrDev a,
                                                              const double* __restrict__ b, float* __restrict__ t,
                                                              float* __restrict__ d, double* __restrict__ x0,
                                                              double* __restrict__ x1, double* __restrict__ r,
                                                              int64_t ld, int32_t ncols, double tol, int32_t max_steps,
                                                              int32_t* __restrict__ steps_out,
                                                              double* __restrict__ ratio_out, int32_t xcds,
                                                              int32_t floor_rule) {
  __shared__ double scratch[1024];
  __shared__ double s_bnorm[TW], s_rnorm[TW], s_ratio[TW], s_best[TW], s_axnorm[TW];
  __shared__ int32_t s_best_buf[TW], s_steps[TW], s_active[TW];
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lpf = reinterpret_cast<int32_t*>(smem);
  int32_t* lpb = lpf + fwd.levels + 1;
  for (int32_t l = threadIdx.x; l <= fwd.levels; l += blockDim.x) lpf[l] = fwd.level_ptr[l];
  for (int32_t l = threadIdx.x; l <= bwd.levels; l += blockDim.x) lpb[l] = bwd.level_ptr[l];
  __syncthreads();
  const Tile<TW> g(ncols, tile_of_workgroup(xcds));
  const int32_t col = g.col;
  const int32_t rows = 2 * a.rows;  // block rows: real and imaginary parts

  double m = 0.0;
  if (g.on)
    for (int32_t i = g.slot; i < a.rows; i += g.slots) m = fold_max(m, zabs(zload<double>(b, i, ld, col)));
  reduce_max<TW>(scratch, m, s_bnorm, g);

  // x0 = solve(b), its residual and ratio
  zsolve32<TW>(fwd, bwd, lpf, lpb, b, t, d, ld, g, g.on);
  if (g.on)
    for (int32_t k = g.slot; k < rows; k += g.slots) x0[k * ld + col] = (double)d[k * ld + col];
  __syncthreads();
  double ax = 0.0;
  m = zresidual<TW>(a, b, x0, r, ld, g, g.on, floor_rule ? &ax : nullptr);
  reduce_max<TW>(scratch, m, s_rnorm, g);
  if (floor_rule) reduce_max<TW>(scratch, ax, s_axnorm, g);
  if (g.slot == 0) {
    const double ratio = ratio_of(s_rnorm[g.lane], s_bnorm[g.lane]);
    s_ratio[g.lane] = ratio;
    s_best[g.lane] = ratio;
    s_best_buf[g.lane] = 0;
    s_steps[g.lane] = 0;
    s_active[g.lane] = g.on && max_steps > 0 && !(ratio <= tol) &&
                       !(ratio <= floor_of(floor_rule, s_axnorm[g.lane], s_bnorm[g.lane]));
  }
  __syncthreads();

  double* cur = x0;
  double* nxt = x1;
  for (int32_t step = 1; step <= max_steps; ++step) {
    bool any = false;
    for (int c = 0; c < TW; ++c) any = any || s_active[c];
    if (!any) break;
    const bool act = g.on && s_active[g.lane];
    zsolve32<TW>(fwd, bwd, lpf, lpb, r, t, d, ld, g, act);
    if (act)
      for (int32_t k = g.slot; k < rows; k += g.slots) nxt[k * ld + col] = cur[k * ld + col] + (double)d[k * ld + col];
    __syncthreads();
    m = zresidual<TW>(a, b, nxt, r, ld, g, act, floor_rule ? &ax : nullptr);
    reduce_max<TW>(scratch, m, s_rnorm, g);
    if (floor_rule) reduce_max<TW>(scratch, ax, s_axnorm, g);
    if (g.slot == 0 && act) {
      const double previous = s_ratio[g.lane];
      const double ratio = ratio_of(s_rnorm[g.lane], s_bnorm[g.lane]);
      s_ratio[g.lane] = ratio;
      if (ratio < s_best[g.lane]) {
        s_best[g.lane] = ratio;
        s_best_buf[g.lane] = nxt == x1 ? 1 : 0;
        s_steps[g.lane] = step;
      }
      const bool stop = ratio != ratio || ratio > 0.5 * previous || ratio <= tol || step == max_steps ||
                        ratio <= floor_of(floor_rule, s_axnorm[g.lane], s_bnorm[g.lane]);
      s_active[g.lane] = !stop;
    }
    __syncthreads();
    double* swap = cur;
    cur = nxt;
    nxt = swap;
  }

  // the best iterate into x0
  if (g.on && s_best_buf[g.lane] == 1)
    for (int32_t k = g.slot; k < rows; k += g.slots) x0[k * ld + col] = x1[k * ld + col];
  if (g.on && g.slot == 0) {
    steps_out[col] = s_steps[g.lane];
    ratio_out[col] = s_best[g.lane];
  }
}

// ---- partitioned FP64 solves: a sweep's independent parts side by side ----
//
// Only Cholesky factors are partitioned, and their rows are short: every sweep here issues only
// each row's own loads (gather's Skip).

// Stages part q's level pointers (from lp_all: the sweep's, or a tile's own list over the same
// levels) into LDS at lp and returns the sweep restricted to that part's levels. Every thread
// must call; ends with a barrier.
__device__ SweepDev part_of_sweep(const SweepDev& s, const int32_t* __restrict__ lp_all, int32_t q, int32_t* lp) {
  const int32_t g0 = s.part_level[q], g1 = s.part_level[q + 1];
  for (int32_t l = threadIdx.x; l <= g1 - g0; l += blockDim.x) lp[l] = lp_all[g0 + l];
  __syncthreads();
  SweepDev v = s;
  v.levels = g1 - g0;
  return v;
}

// Copies tile `tile`'s reach bitmap into LDS (every thread must call; ends with a barrier)
__device__ void stage_live(const ReachDev& reach, int32_t tile, uint32_t* live) {
  const uint32_t* __restrict__ src = reach.live + (int64_t)tile * reach.words;
  for (int32_t w = threadIdx.x; w < reach.words; w += blockDim.x) live[w] = src[w];
  __syncthreads();
}

// The right-hand side a forward row reads: column col of the block, or of e_plus - e_minus
struct ForwardRhs {
  const double* __restrict__ block;
  int64_t ld;
  int32_t col, plus, minus;
  __device__ double operator()(int32_t src) const {
    if (block) return block[src * ld + col];
    return (src == plus ? 1.0 : 0.0) - (src == minus ? 1.0 : 0.0);
  }
};

__device__ __forceinline__ ForwardRhs forward_rhs(const RhsDev& rhs, int64_t ld, int32_t col, bool on) {
  ForwardRhs f{rhs.block, ld, col, -1, -1};
  if (!rhs.block && on) {
    f.plus = rhs.plus[col];
    f.minus = rhs.minus[col];
  }
  return f;
}

// A bitmap of live rows in LDS
struct LiveBits {
  const uint32_t* bits;
  __device__ __forceinline__ bool operator()(int32_t row) const { return (bits[row >> 5] >> (row & 31)) & 1u; }
};
struct ReachAt {
  const int32_t* __restrict__ pos;
  __device__ __forceinline__ int32_t operator()(int32_t k) const { return pos[k]; }
};

// Phase 1: workgroup (tile, part) sweeps the part's rows forward into t. Split: the sweeps are
// split (blocks of vectors only), so a segment with src < 0 continues from its row's slot of the
// block it writes.
template <int TW, bool Split>
__global__ void __launch_bounds__(1024) parts_forward_kernel(SweepDev fwd, RhsDev rhs, ReachDev reach,
                                                             double* __restrict__ t, int64_t ld, int32_t ncols,
                                                             int32_t lp_words) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lp = reinterpret_cast<int32_t*>(smem);
  uint32_t* live = reinterpret_cast<uint32_t*>(lp + lp_words);
  const bool pairs_reach = reach.lp != nullptr;
  const int32_t* lp_all = pairs_reach ? reach.lp + (int64_t)blockIdx.x * (fwd.levels + 1) : fwd.level_ptr;
  const SweepDev s = part_of_sweep(fwd, lp_all, (int32_t)blockIdx.y, lp);
  const Tile<TW> g(ncols);
  const ForwardRhs from = forward_rhs(rhs, ld, g.col, g.on);
  auto store = [&](int32_t dst, double v) { t[dst * ld + g.col] = v; };
  if (pairs_reach) {
    stage_live(reach, (int32_t)blockIdx.x, live);
    wide_sweep<double, TW>(s, lp, t, ld, g, g.on, from, store, ReachAt{reach.pos}, LiveBits{live}, std::true_type());
  } else {
    auto rhs_of = [&](int32_t src) { return Split && src < 0 ? t[(~src) * ld + g.col] : from(src); };
    wide_sweep<double, TW>(s, lp, t, ld, g, g.on, rhs_of, store, Identity(), AllLive(), std::true_type());
  }
}

// Phase 2: per tile, the top part forward (into t) and then backward (into x)
template <int TW, bool Split>
__global__ void __launch_bounds__(1024) top_kernel(SweepDev fwd, SweepDev bwd, RhsDev rhs, ReachDev reach,
                                                   double* __restrict__ t, double* __restrict__ x, int64_t ld,
                                                   int32_t ncols, int32_t lpf_words, int32_t lpb_words) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lpf = reinterpret_cast<int32_t*>(smem);
  int32_t* lpb = lpf + lpf_words;
  uint32_t* live = reinterpret_cast<uint32_t*>(lpb + lpb_words);
  const bool pairs_reach = reach.lp != nullptr;
  const int32_t* lp_all = pairs_reach ? reach.lp + (int64_t)blockIdx.x * (fwd.levels + 1) : fwd.level_ptr;
  const SweepDev sf = part_of_sweep(fwd, lp_all, fwd.parts, lpf);
  const SweepDev sb = part_of_sweep(bwd, bwd.level_ptr, bwd.parts, lpb);
  const Tile<TW> g(ncols);
  const int32_t col = g.col;
  const ForwardRhs from = forward_rhs(rhs, ld, col, g.on);
  auto store_t = [&](int32_t dst, double v) { t[dst * ld + col] = v; };
  auto store_x = [&](int32_t dst, double v) { x[dst * ld + col] = v; };
  if (pairs_reach) {
    stage_live(reach, (int32_t)blockIdx.x, live);
    const LiveBits in_reach{live};
    wide_sweep<double, TW>(sf, lpf, t, ld, g, g.on, from, store_t, ReachAt{reach.pos}, in_reach, std::true_type());
    const int32_t pad = fwd.n;
    auto from_t = [&](int32_t src) -> double {
      const bool on = in_reach(src);
      const double v = t[(on ? src : pad) * ld + col];
      return on ? v : 0.0;
    };
    wide_sweep<double, TW>(sb, lpb, x, ld, g, g.on, from_t, store_x, Identity(), AllLive(), std::true_type());
  } else {
    auto rhs_t = [&](int32_t src) { return Split && src < 0 ? t[(~src) * ld + col] : from(src); };
    auto rhs_x = [&](int32_t src) { return Split && src < 0 ? x[(~src) * ld + col] : t[src * ld + col]; };
    wide_sweep<double, TW>(sf, lpf, t, ld, g, g.on, rhs_t, store_t, Identity(), AllLive(), std::true_type());
    wide_sweep<double, TW>(sb, lpb, x, ld, g, g.on, rhs_x, store_x, Identity(), AllLive(), std::true_type());
  }
}

// Phase 2 with the top in LDS (TopSlotsDev): the top is a few hundred rows over tens of levels,
// each level a handful of segments, so a level's cost is its latency. Here every value a top
// segment reads is a slot in LDS. The workgroup first loads the top rows' right-hand sides, the
// halo (the parts' rows the forward top reads, final after phase 1) and the metadata image, every
// thread issuing all its loads before its stores. Then each column of the tile is swept by one
// wavefront alone (TW columns over min(TW, waves) wavefronts), so the levels need no workgroup
// barrier: LDS operations of one wavefront complete in order, and a wavefront fence keeps the
// compiler from moving a level's reads above the last one's writes. The forward top sweeps the
// slots, the backward top sweeps them again in place (a backward row starts from its own forward
// value and reads only rows it follows, whose slots already hold their solutions), and the
// workgroup stores the solutions. Each segment subtracts its terms in the sweep's order and
// divides as the sweep does, so the bits are those of top_kernel.
struct TopStage {
  int32_t levels, segs;
  const TopRecord* rec;
  const double* diag;
  const int32_t* lp;
  const double* ov_val;
  const int32_t* ov_dep;
};

__device__ __forceinline__ TopStage top_stage(const TopPartDev& p, const unsigned char* meta,
                                              const TopLdsLayout::Part& at) {
  return TopStage{p.levels, p.segs,
                  reinterpret_cast<const TopRecord*>(meta + at.rec),
                  reinterpret_cast<const double*>(meta + at.diag),
                  reinterpret_cast<const int32_t*>(meta + at.lp),
                  reinterpret_cast<const double*>(meta + at.ov_val),
                  reinterpret_cast<const int32_t*>(meta + at.ov_dep)};
}

__device__ __forceinline__ void wavefront_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A segment's record and diagonal, as registers
struct TopSeg {
  TopRecord r;
  double diag;
};
__device__ __forceinline__ TopSeg top_seg(const TopStage& st, int32_t k) {
  k = k < st.segs ? k : st.segs - 1;
  return TopSeg{st.rec[k], st.diag[k]};
}

// One segment of column `col`: its slot's value minus its terms in order (its first four in the
// record, the rest from ov_val / ov_dep), over its diagonal
template <int TW>
__device__ __forceinline__ void top_segment(const double* ov_val, const int32_t* ov_dep, const TopSeg& g, bool unit,
                                            double* v, int32_t col) {
  const int32_t slot = g.r.head.x, terms = g.r.head.y;
  const double x0 = v[g.r.dep.x * TW + col], x1 = v[g.r.dep.y * TW + col], x2 = v[g.r.dep.z * TW + col],
               x3 = v[g.r.dep.w * TW + col];
  double a = v[slot * TW + col];
  if (terms > 0) a = a - g.r.v01.x * x0;
  if (terms > 1) a = a - g.r.v01.y * x1;
  if (terms > 2) a = a - g.r.v23.x * x2;
  if (terms > 3) a = a - g.r.v23.y * x3;
  for (int32_t q = g.r.head.z, end = g.r.head.z + terms - RECORD_ENTRIES; q < end; ++q)
    a = a - ov_val[q] * v[ov_dep[q] * TW + col];
  v[slot * TW + col] = unit ? a : a / g.diag;
}

// One sweep's top over the slots of column `col` of the tile (TW columns side by side): this
// lane takes segments first, first + step, ... of each level. A level's first record is read
// before the ordering point that ends the level before it (records do not change), so after
// that point a level waits only for the values it reads.
template <int TW>
__device__ void top_slots_sweep(const TopStage& st, bool unit, double* v, int32_t col, int32_t first, int32_t step) {
  if (st.levels == 0 || st.segs == 0) return;
  int32_t lo = st.lp[0], hi = st.lp[1];
  TopSeg g = top_seg(st, lo + first < hi ? lo + first : lo);
  for (int32_t l = 0; l < st.levels; ++l) {
    const int32_t next_hi = st.lp[l + 2 <= st.levels ? l + 2 : st.levels];
    for (int32_t k = lo + first; k < hi;) {
      top_segment<TW>(st.ov_val, st.ov_dep, g, unit, v, col);
      k += step;
      if (k < hi) g = top_seg(st, k);
    }
    lo = hi;
    hi = next_hi;
    if (l + 1 < st.levels) g = top_seg(st, lo + first < hi ? lo + first : lo);
    wavefront_order();
  }
}

template <int TW>
__global__ void __launch_bounds__(1024) top_lds_kernel(SweepDev fwd, SweepDev bwd, TopSlotsDev top, RhsDev rhs,
                                                       const double* __restrict__ t, double* __restrict__ x,
                                                       int64_t ld, int32_t ncols) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int U = 8;  // loads in flight per thread
  const TopLdsLayout o = top_lds_layout(top, TW);
  double* v = reinterpret_cast<double*>(smem + o.values);
  unsigned char* meta = smem + o.meta;
  const int32_t tid = threadIdx.x, threads = blockDim.x;
  const int32_t col0 = (int32_t)blockIdx.x * TW;
  // the slots' values: item i is slot i / TW of column i % TW
  {
    const int32_t items = (top.rows + top.halo) * TW;
    for (int32_t base = 0; base < items; base += U * threads) {
      double r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = base + u * threads + tid;
        r[u] = 0.0;
        if (i < items) {
          const int32_t sl = i / TW, col = col0 + i % TW;
          if (col < ncols) {
            if (sl < top.rows) {
              const ForwardRhs from = forward_rhs(rhs, ld, col, true);
              r[u] = from(top.rhs_row[sl]);
            } else {
              r[u] = t[top.halo_row[sl - top.rows] * ld + col];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = base + u * threads + tid;
        if (i < items) v[i] = r[u];
      }
    }
  }
  // the metadata image
  {
    int4* dst = reinterpret_cast<int4*>(meta);
    for (int32_t base = 0; base < top.meta_units; base += U * threads) {
      int4 r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = base + u * threads + tid;
        r[u] = i < top.meta_units ? top.meta[i] : make_int4(0, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = base + u * threads + tid;
        if (i < top.meta_units) dst[i] = r[u];
      }
    }
  }
  __syncthreads();
  // wavefront w < sweepers takes columns w * per .. (w + 1) * per - 1 of the tile
  const int32_t wave = tid / 64, lane = tid % 64;
  int32_t sweepers = 1;
  while (sweepers * 2 <= TW && sweepers * 2 <= threads / 64) sweepers *= 2;
  const int32_t per = TW / sweepers;
  if (wave < sweepers) {
    const int32_t col = wave * per + lane % per, first = lane / per, step = 64 / per;
    top_slots_sweep<TW>(top_stage(top.fwd, meta, o.fwd), fwd.unit, v, col, first, step);
    top_slots_sweep<TW>(top_stage(top.bwd, meta, o.bwd), bwd.unit, v, col, first, step);
  }
  __syncthreads();
  for (int32_t i = tid; i < top.rows * TW; i += threads) {
    const int32_t col = col0 + i % TW;
    if (col < ncols) x[top.out_row[i / TW] * ld + col] = v[i];
  }
}

// Phase 3: workgroup (tile, part) sweeps the part's rows backward from t into x
template <int TW, bool Split>
__global__ void __launch_bounds__(1024) parts_backward_kernel(SweepDev bwd, ReachDev reach, const double* __restrict__ t,
                                                              double* __restrict__ x, int64_t ld, int32_t ncols,
                                                              int32_t lp_words) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lp = reinterpret_cast<int32_t*>(smem);
  uint32_t* live = reinterpret_cast<uint32_t*>(lp + lp_words);
  const SweepDev s = part_of_sweep(bwd, bwd.level_ptr, (int32_t)blockIdx.y, lp);
  const Tile<TW> g(ncols);
  const int32_t col = g.col;
  auto store_x = [&](int32_t dst, double v) { x[dst * ld + col] = v; };
  if (reach.lp) {
    stage_live(reach, (int32_t)blockIdx.x, live);
    const LiveBits in_reach{live};
    const int32_t pad = bwd.n;
    auto from_t = [&](int32_t src) -> double {
      const bool on = in_reach(src);
      const double v = t[(on ? src : pad) * ld + col];
      return on ? v : 0.0;
    };
    wide_sweep<double, TW>(s, lp, x, ld, g, g.on, from_t, store_x, Identity(), AllLive(), std::true_type());
  } else {
    auto rhs_x = [&](int32_t src) { return Split && src < 0 ? x[(~src) * ld + col] : t[src * ld + col]; };
    wide_sweep<double, TW>(s, lp, x, ld, g, g.on, rhs_x, store_x, Identity(), AllLive(), std::true_type());
  }
}

// Phases 1 and 3 with a part's values in LDS (PartSlotsDev): workgroup w takes part w / tiles
// and tile w % tiles, consecutive w on one XCD, so a part's records are fetched into that XCD's
// L2 once. It loads the part's slots (forward: the right-hand sides; backward: the forward values
// from src and the halo, the top's solutions, from dst), its overflow terms and its rounds into
// LDS. Then each column of the tile is swept by one wavefront alone, as the LDS top is: a lane
// takes one segment a round, the levels need no workgroup barrier, and a wavefront fence after
// each round keeps the compiler from moving the next round's reads above its writes. A part's
// records (64 bytes and a diagonal a segment) are too many for LDS beside the values, so they
// stay in global memory: each lane loads its segments PARTS_LOOKAHEAD rounds before it sweeps
// them, which records allow since they do not depend on the values. Each segment subtracts its
// terms in the sweep's order and divides as the sweep does, so the bits are those of the parts
// kernels over the block. Then the workgroup stores the part's results into dst.
template <int TW, bool Backward>
__global__ void __launch_bounds__(1024) parts_lds_kernel(PartSlotsDev ps, int32_t unit, RhsDev rhs,
                                                         const double* __restrict__ src, double* dst, int64_t ld,
                                                         int32_t ncols, int32_t tiles, int32_t xcds) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int U = 8;  // items a thread stages or stores at once
  constexpr int D = PARTS_LOOKAHEAD;
  const PartsLdsLayout o = parts_lds_layout(ps, TW);
  double* v = reinterpret_cast<double*>(smem);
  double* ov_val = reinterpret_cast<double*>(smem + o.ov_val);
  int32_t* ov_dep = reinterpret_cast<int32_t*>(smem + o.ov_dep);
  int4* rt = reinterpret_cast<int4*>(smem + o.rounds);
  const int32_t w = tile_of_workgroup(xcds);
  const PartDesc d = ps.desc[w / tiles];
  const int32_t col0 = (w % tiles) * TW;
  const int32_t tid = threadIdx.x, threads = blockDim.x;
  // the slots' values: item i is slot i / TW of column i % TW, so a tile's columns of a row are
  // neighbouring lanes. Every thread first loads its U items' rows, then their values, so a batch
  // waits for two rounds of loads (an item with no row loads element 0 of its block and drops it)
  const int32_t items = (d.rows + d.halo) * TW;
  {
    const double* __restrict__ from = Backward ? src : rhs.block;  // a pair block: from plus, minus
    for (int32_t base = 0; base < items; base += U * threads) {
      int32_t row[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = base + u * threads + tid, sl = i / TW;
        row[u] = i >= items || col0 + i % TW >= ncols ? -1
                 : sl < d.rows                         ? ps.in_row[d.slot0 + sl]
                                                       : ps.halo_row[d.halo0 + sl - d.rows];
      }
      double r[U];
      if (from) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t i = base + u * threads + tid;
          const double* p = Backward && i / TW >= d.rows ? dst : from;
          const double value = p[row[u] >= 0 ? (int64_t)row[u] * ld + col0 + i % TW : 0];
          r[u] = row[u] >= 0 ? value : 0.0;
        }
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t i = base + u * threads + tid;
          r[u] = row[u] >= 0 ? forward_rhs(rhs, ld, col0 + i % TW, true)(row[u]) : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = base + u * threads + tid;
        if (i < items) v[i] = r[u];
      }
    }
  }
  for (int32_t i = tid; i < d.ov; i += threads) {
    ov_val[i] = ps.ov_val[d.ov0 + i];
    ov_dep[i] = ps.ov_dep[d.ov0 + i];
  }
  for (int32_t i = tid; i < d.rounds + D; i += threads)
    rt[i] = i < d.rounds ? ps.rounds[d.round0 + i] : make_int4(0, 0, 0, 0);
  __syncthreads();
  const int32_t wave = tid / 64, lane = tid % 64;
  if (wave < TW) {
    const TopRecord* __restrict__ rec = ps.rec + d.seg0;
    const double* __restrict__ diag = ps.diag + d.seg0;
    // this lane's segment of round R (an empty round or lane: the part's first, with head.w 0 so
    // it is skipped; a fetched segment has head.w 1)
    auto fetch = [&](int32_t R) {
      const int4 e = rt[R];
      const bool on = lane < e.y;
      TopSeg g{rec[on ? e.x + lane : 0], diag[on ? e.x + lane : 0]};
      g.r.head.w = on ? 1 : 0;
      return g;
    };
    // the fetches stay in round order (a scheduling barrier after each): the loop's first wait
    // must hold for the first pass too, so a first fetch issued last would make every pass wait
    // for all its loads
    TopSeg g[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      g[u] = fetch(u);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int32_t R0 = 0; R0 < d.rounds; R0 += D) {
#pragma unroll
      for (int u = 0; u < D; ++u) {
        if (g[u].r.head.w) top_segment<TW>(ov_val, ov_dep, g[u], unit != 0, v, wave);
        g[u] = fetch(R0 + u + D);
        __builtin_amdgcn_sched_barrier(0);
        wavefront_order();
      }
    }
  }
  __syncthreads();
  // the part's results, U items a thread at once: their rows, then their stores
  for (int32_t base = 0; base < d.rows * TW; base += U * threads) {
    int32_t row[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t i = base + u * threads + tid;
      row[u] = i < d.rows * TW && col0 + i % TW < ncols ? ps.out_row[d.slot0 + i / TW] : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t i = base + u * threads + tid;
      if (row[u] >= 0) dst[(int64_t)row[u] * ld + col0 + i % TW] = v[i];
    }
  }
}

// ---- FP32 sweeps refined in FP64 ----

// Forward FP32 sweep of one tile: t[dst] = (float(src[src_row]) - sum val * t[dep]) / diag
// (Split: a segment with src < 0 continues from t at its row)
template <int TW, bool Split>
__device__ void forward32(const SweepDev& s, const int32_t* lp, const double* __restrict__ src, float* __restrict__ t,
                          int64_t ld, const Tile<TW>& g, bool act) {
  const int32_t col = g.col;
  wide_sweep<float, TW>(s, lp, t, ld, g, act,
                        [&](int32_t r) { return Split && r < 0 ? t[(~r) * ld + col] : (float)src[r * ld + col]; },
                        [&](int32_t dst, float v) { t[dst * ld + col] = v; });
}

// Backward FP32 sweep of one tile into d, the correction (its own dependencies), and the FP64
// iterate xn = xc + double(d) (double(d) alone when xc is null). Unsplit, each row's store forms
// xn too. Split, the sweep stores d alone and combine() forms xn after it: a segment is not its
// row's last, and with many segments per level a store that loads xc waits behind the next
// row's gathers (loads retire in order); a segment with src < 0 continues from d at its row.
template <int TW, bool Split>
__device__ void backward32(const SweepDev& s, const int32_t* lp, const float* __restrict__ t, float* __restrict__ d,
                           const double* __restrict__ xc, double* __restrict__ xn, int64_t ld, const Tile<TW>& g,
                           bool act) {
  const int32_t col = g.col;
  wide_sweep<float, TW>(s, lp, d, ld, g, act,
                        [&](int32_t r) { return Split && r < 0 ? d[(~r) * ld + col] : t[r * ld + col]; },
                        [&](int32_t dst, float v) {
                          const int64_t at = dst * ld + col;
                          d[at] = v;
                          if (!Split) xn[at] = xc ? xc[at] + (double)v : (double)v;
                        });
}

// The new iterate xn = xc + double(d) (double(d) alone when xc is null) for the tile's columns:
// every row independent, four rows' loads in flight per thread
template <int TW>
__device__ void combine(const double* __restrict__ xc, const float* __restrict__ d, double* __restrict__ xn,
                        int32_t rows, int64_t ld, const Tile<TW>& g, bool act) {
  if (!act) return;
  int32_t i = g.slot;
  for (; i + 3 * g.slots < rows; i += 4 * g.slots) {
    int64_t at[4];
    double c[4];
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      at[q] = (int64_t)(i + q * g.slots) * ld + g.col;
      c[q] = xc ? xc[at[q]] : 0.0;
      v[q] = d[at[q]];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) xn[at[q]] = xc ? c[q] + (double)v[q] : (double)v[q];
  }
  for (; i < rows; i += g.slots) {
    const int64_t at = (int64_t)i * ld + g.col;
    xn[at] = xc ? xc[at] + (double)d[at] : (double)d[at];
  }
}

// r = b - A x for the tile's columns (each row summed from 0 in index order, as mul_vec's
// dot_kernel does, product.rs:17-27); returns this thread's partial max |r|. A thread's rows are
// pipelined: while one row's entries of x are in flight, the next row's first RESIDUAL_CHUNK
// columns and values and the bounds of the row after that are loaded, so a row of up to
// RESIDUAL_CHUNK terms costs one round of dependent loads. Two register sets alternate between
// the row being summed and the next one, so no value in flight is copied. The matrix's column and
// value arrays carry RESIDUAL_CHUNK padding entries past its last (column 0, value 0), so a chunk
// is loaded whole from any position up to nnz; a chunk's terms past its row's end are not summed,
// and their x loads repeat the row's first column's.
// With ax non-null (the floor stop rule) it also takes this thread's partial max of
// sum |a_ij| |x_j| over its rows into *ax.
struct ResidualRow {
  int32_t p0, p1;
  int32_t c[RESIDUAL_CHUNK];
  double v[RESIDUAL_CHUNK];
};

template <int TW>
__device__ double residual(const CsrDev& a, const double* __restrict__ b, const double* __restrict__ x,
                           double* __restrict__ r, int64_t ld, const Tile<TW>& g, bool act, double* ax = nullptr) {
  constexpr int C = RESIDUAL_CHUNK;
  double m = 0.0, axm = 0.0;
  if (ax) *ax = 0.0;
  const int32_t rows = a.rows;
  if (!act || g.slot >= rows) return m;
  auto bounds = [&](int32_t i, ResidualRow& w) {
    const int32_t k = i < rows ? i : rows;
    w.p0 = a.ptr[k];
    w.p1 = a.ptr[k < rows ? k + 1 : rows];
  };
  auto chunk = [&](ResidualRow& w) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      w.c[q] = a.col[w.p0 + q];
      w.v[q] = a.val[w.p0 + q];
    }
  };
  // row i from `cur` (its bounds and first chunk loaded); `nxt` holds row i + slots's bounds and
  // gets its first chunk, and `cur` then the bounds of row i + 2 slots
  auto row = [&](int32_t i, ResidualRow& cur, ResidualRow& nxt) {
    const int32_t p0 = cur.p0, p1 = cur.p1;
    const int64_t at = (int64_t)i * ld + g.col;
    double xx[C];
#pragma unroll
    for (int q = 0; q < C; ++q) xx[q] = x[(int64_t)(p0 + q < p1 ? cur.c[q] : cur.c[0]) * ld + g.col];
    const double bi = b[at];
    chunk(nxt);
    bounds(i + 2 * g.slots, cur);
    double sum = 0.0, abs_sum = 0.0;
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (p0 + q < p1) sum = sum + cur.v[q] * xx[q];
    if (ax) {
#pragma unroll
      for (int q = 0; q < C; ++q)
        if (p0 + q < p1) abs_sum = abs_sum + fabs(cur.v[q]) * fabs(xx[q]);
    }
    // the rest of a long row, a chunk at a time
    for (int32_t p = p0 + C; p < p1; p += C) {
      int32_t cc[C];
      double v2[C], x2[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        cc[q] = a.col[p + q];
        v2[q] = a.val[p + q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) x2[q] = x[(int64_t)(p + q < p1 ? cc[q] : cc[0]) * ld + g.col];
#pragma unroll
      for (int q = 0; q < C; ++q)
        if (p + q < p1) sum = sum + v2[q] * x2[q];
      if (ax) {
#pragma unroll
        for (int q = 0; q < C; ++q)
          if (p + q < p1) abs_sum = abs_sum + fabs(v2[q]) * fabs(x2[q]);
      }
    }
    const double v = bi - sum;
    r[at] = v;
    m = fold_max(m, fabs(v));
    axm = fold_max(axm, abs_sum);
  };
  ResidualRow s0, s1;
  bounds(g.slot, s0);
  chunk(s0);
  bounds(g.slot + g.slots, s1);
  for (int32_t i = g.slot;;) {
    row(i, s0, s1);
    i += g.slots;
    if (i >= rows) break;
    row(i, s1, s0);
    i += g.slots;
    if (i >= rows) break;
  }
  if (ax) *ax = axm;
  return m;
}


template <int TW, bool Split>
__global__ void __launch_bounds__(1024) refine_kernel(SweepDev fwd, SweepDev bwd, CsrDev a, const double* __restrict__ b,
                                                      float* __restrict__ t, float* __restrict__ d,
                                                      double* __restrict__ x0, double* __restrict__ x1,
                                                      double* __restrict__ r, int64_t ld, int32_t ncols, double tol,
                                                      int32_t max_steps, int32_t* __restrict__ steps_out,
                                                      double* __restrict__ ratio_out, int32_t xcds,
                                                      int32_t floor_rule) {
  __shared__ double scratch[1024];
  __shared__ double s_bnorm[TW], s_rnorm[TW], s_ratio[TW], s_best[TW], s_axnorm[TW];
  __shared__ int32_t s_best_buf[TW], s_steps[TW], s_active[TW];
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* lpf = reinterpret_cast<int32_t*>(smem);
  int32_t* lpb = lpf + fwd.levels + 1;
  stage_levels(fwd, lpf);
  stage_levels(bwd, lpb);
  const Tile<TW> g(ncols, tile_of_workgroup(xcds));

  // ||b||inf per column
  double m = 0.0;
  if (g.on)
    for (int32_t i = g.slot; i < a.rows; i += g.slots) m = fold_max(m, fabs(b[i * ld + g.col]));
  reduce_max<TW>(scratch, m, s_bnorm, g);

  // x0 = solve(b), its residual and ratio
  forward32<TW, Split>(fwd, lpf, b, t, ld, g, g.on);
  backward32<TW, Split>(bwd, lpb, t, d, nullptr, x0, ld, g, g.on);
  if (Split) {
    __syncthreads();
    combine<TW>(nullptr, d, x0, a.rows, ld, g, g.on);
    __syncthreads();
  }
  double ax = 0.0;
  m = residual<TW>(a, b, x0, r, ld, g, g.on, floor_rule ? &ax : nullptr);
  reduce_max<TW>(scratch, m, s_rnorm, g);
  if (floor_rule) reduce_max<TW>(scratch, ax, s_axnorm, g);
  if (g.slot == 0) {
    const double ratio = ratio_of(s_rnorm[g.lane], s_bnorm[g.lane]);
    s_ratio[g.lane] = ratio;
    s_best[g.lane] = ratio;
    s_best_buf[g.lane] = 0;
    s_steps[g.lane] = 0;
    s_active[g.lane] = g.on && max_steps > 0 && !(ratio <= tol) &&
                       !(ratio <= floor_of(floor_rule, s_axnorm[g.lane], s_bnorm[g.lane]));
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
    forward32<TW, Split>(fwd, lpf, r, t, ld, g, act);
    backward32<TW, Split>(bwd, lpb, t, d, cur, nxt, ld, g, act);
    if (Split) {
      __syncthreads();
      combine<TW>(cur, d, nxt, a.rows, ld, g, act);
      __syncthreads();
    }
    m = residual<TW>(a, b, nxt, r, ld, g, act, floor_rule ? &ax : nullptr);
    reduce_max<TW>(scratch, m, s_rnorm, g);
    if (floor_rule) reduce_max<TW>(scratch, ax, s_axnorm, g);
    if (g.slot == 0 && act) {
      const double previous = s_ratio[g.lane];
      const double ratio = ratio_of(s_rnorm[g.lane], s_bnorm[g.lane]);
      s_ratio[g.lane] = ratio;
      if (ratio < s_best[g.lane]) {
        s_best[g.lane] = ratio;
        s_best_buf[g.lane] = nxt == x1 ? 1 : 0;
        s_steps[g.lane] = step;
      }
      const bool stop = ratio != ratio || ratio > 0.5 * previous || ratio <= tol || step == max_steps ||
                        ratio <= floor_of(floor_rule, s_axnorm[g.lane], s_bnorm[g.lane]);
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

// A barrier that orders LDS only: the record prefetches in flight stay in flight across it
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One row in place: xs[dst] = (xs[dst] - its terms) / diag, the terms subtracted in the row's
// packed order, the inline four first, then the rest from the sweep's arrays.
template <typename T, bool G = false>
__device__ __forceinline__ void lds_row(const SweepDev& s, const T* __restrict__ val, T* xs, const Record<T>& r) {
  const int32_t len = r.p1 - r.p0;
  T acc = xs[r.dst];
  const T x0 = xs[r.dep[0]], x1 = xs[r.dep[1]], x2 = xs[r.dep[2]], x3 = xs[r.dep[3]];
  if (len > 0) acc = acc - r.val[0] * x0;
  if (len > 1) acc = acc - r.val[1] * x1;
  if (len > 2) acc = acc - r.val[2] * x2;
  if (len > 3) acc = acc - r.val[3] * x3;
  // the rest in chunks: a chunk's loads are all issued before its first term is subtracted
  for (int32_t p = r.p0 + RECORD_ENTRIES; p < r.p1; p += OVERFLOW_CHUNK) {
    int32_t dd[OVERFLOW_CHUNK];
    T vv[OVERFLOW_CHUNK], xx[OVERFLOW_CHUNK];
#pragma unroll
    for (int q = 0; q < OVERFLOW_CHUNK; ++q) {
      const int32_t pq = p + q < r.p1 ? p + q : r.p1 - 1;
      dd[q] = load_at<G>(s.dep, pq);
      vv[q] = load_at<G>(val, pq);
    }
#pragma unroll
    for (int q = 0; q < OVERFLOW_CHUNK; ++q) xx[q] = xs[dd[q]];
#pragma unroll
    for (int q = 0; q < OVERFLOW_CHUNK; ++q)
      if (p + q < r.p1) acc = acc - vv[q] * xx[q];
  }
  xs[r.dst] = s.unit ? acc : acc / r.diag;
}

// Levels a column sweep fetches its records ahead of use
constexpr int LDS_AHEAD = 4;

// One in-place sweep over xs (factor order), a barrier after each level; lp is the sweep's level
// pointers, staged in LDS. Thread `slot` takes rows lp[l] + slot, + slots, ... of level l. The
// record of its first row of each level is fetched LDS_AHEAD levels before it is used, into one
// of LDS_AHEAD register slots that the loop (unrolled by LDS_AHEAD) never copies; the fetch is
// issued whether or not the thread has a row there (the position is clamped to the sweep's
// last), so the number of loads in flight, and the waits the compiler places, are the same on
// every path. A narrow level (the common case) then waits only on LDS and its arithmetic; the
// further rows of a wide level are fetched as they come.
template <typename T, bool G = false>
__device__ void lds_sweep(const SweepDev& s, const int32_t* lp, T* xs) {
  const T* __restrict__ val = values_of<T>(s);
  const int32_t slot = threadIdx.x, slots = blockDim.x;
  const int32_t levels = s.levels;
  if (levels == 0) return;
  const int32_t last = lp[levels] - 1;
  auto ahead = [&](int32_t m, Record<T>& r) {
    const int32_t k = m < levels ? lp[m] + slot : last;
    fetch<T, G>(s, k < last ? k : last, r);
  };
  // level l from the record slot holding it; `refill` then fetches level l + LDS_AHEAD into it
  auto level = [&](int32_t l, Record<T>& r, bool refill) {
    const int32_t end = lp[l + 1];
    int32_t k = lp[l] + slot;
    if (k < end) {
      lds_row<T, G>(s, val, xs, r);
      for (k += slots; k < end; k += slots) {
        Record<T> extra;
        fetch<T, G>(s, k, extra);
        lds_row<T, G>(s, val, xs, extra);
      }
    }
    if (refill) {
      await_record<T>(r);
      ahead(l + LDS_AHEAD, r);
    }
    lds_barrier();
  };
  Record<T> r[LDS_AHEAD];
#pragma unroll
  for (int d = 0; d < LDS_AHEAD; ++d) ahead(d, r[d]);
  // whole rounds of LDS_AHEAD levels, with no exit but the loop's own test (so the loop's head is
  // reached only from a full round, and its waits stay LDS_AHEAD - 1 fetches deep), then the rest
  int32_t l0 = 0;
  for (; l0 + LDS_AHEAD <= levels; l0 += LDS_AHEAD) {
#pragma unroll
    for (int d = 0; d < LDS_AHEAD; ++d) level(l0 + d, r[d], true);
  }
#pragma unroll
  for (int d = 0; d < LDS_AHEAD; ++d)
    if (l0 + d < levels) level(l0 + d, r[d], false);
}

// Copies a sweep's level pointers into LDS (every thread must call; the caller's barrier ends it)
__device__ __forceinline__ void stage_level_ptr(const SweepDev& s, int32_t* lp) {
  for (int32_t l = threadIdx.x; l <= s.levels; l += blockDim.x) lp[l] = s.level_ptr[l];
}

template <typename T>
__global__ void __launch_bounds__(1024) solve_lds_kernel(SweepDev fwd, SweepDev bwd, int32_t n,
                                                         const int32_t* __restrict__ rhs_row,
                                                         const int32_t* __restrict__ sol_of, ColumnsDev rhs,
                                                         double* __restrict__ x, int64_t ldx) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* xs = reinterpret_cast<T*>(smem);
  int32_t* lpf = reinterpret_cast<int32_t*>(smem + lds_vector_bytes(n, sizeof(T)));
  int32_t* lpb = lpf + fwd.levels + 1;
  stage_level_ptr(fwd, lpf);
  stage_level_ptr(bwd, lpb);
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
  lds_sweep<T>(fwd, lpf, xs);
  lds_sweep<T>(bwd, lpb, xs);
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
// (ax non-null: also this thread's partial max of sum |a_ij| |x_j| into *ax)
__device__ double column_residual(const CsrDev& a, const double* __restrict__ b, const double* __restrict__ x,
                                  double* __restrict__ r, double* ax = nullptr) {
  double m = 0.0, axm = 0.0;
  for (int32_t u = threadIdx.x; u < a.rows; u += blockDim.x) {
    double sum = 0.0, abs_sum = 0.0;
    const int32_t pe = a.ptr[u + 1];
    for (int32_t p = a.ptr[u]; p < pe; ++p) sum = sum + a.val[p] * x[a.col[p]];
    if (ax)
      for (int32_t p = a.ptr[u]; p < pe; ++p) abs_sum = abs_sum + fabs(a.val[p]) * fabs(x[a.col[p]]);
    const double v = b[u] - sum;
    r[u] = v;
    m = fold_max(m, fabs(v));
    axm = fold_max(axm, abs_sum);
  }
  if (ax) *ax = axm;
  return m;
}

// FP32 sweeps in LDS refined in FP64 (refine_normwise, accuracy.rs:83-122), one workgroup per
// column: x = solve(b); while the column is active, r = b - A x, d = solve(r), x' = x + d in
// the other buffer. A column continues only while each step at least halves its ratio, so its
// best iterate is the current or the previous one; it ends in x0.
// One column (b, by unknown): the iterates in xa and xb, the residual in r (n values each); the
// best iterate ends in xa, its corrections and relative residual in *steps_out and *ratio_out.
// smem: blockDim.x doubles of scratch, the FP32 vector, both sweeps' level pointers. Every thread
// must call; the caller's barrier must separate it from the next use of smem.
__device__ void refine_column(const SweepDev& fwd, const SweepDev& bwd, const CsrDev& a,
                              const int32_t* __restrict__ rhs_row, const int32_t* __restrict__ sol_of,
                              const double* __restrict__ b, double* __restrict__ xa, double* __restrict__ xb,
                              double* __restrict__ r, double tol, int32_t max_steps, int32_t* __restrict__ steps_out,
                              double* __restrict__ ratio_out, unsigned char* smem, bool floor_rule) {
  double* scratch = reinterpret_cast<double*>(smem);
  const int32_t n = a.rows;
  float* xs = reinterpret_cast<float*>(smem + blockDim.x * sizeof(double));
  int32_t* lpf = reinterpret_cast<int32_t*>(smem + blockDim.x * sizeof(double) + lds_vector_bytes(n, sizeof(float)));
  int32_t* lpb = lpf + fwd.levels + 1;
  stage_level_ptr(fwd, lpf);
  stage_level_ptr(bwd, lpb);

  double m = 0.0;
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) m = fold_max(m, fabs(b[u]));
  const double bnorm = block_max(scratch, m);

  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xs[rhs_row[u]] = (float)b[u];
  __syncthreads();
  lds_sweep<float>(fwd, lpf, xs);
  lds_sweep<float>(bwd, lpb, xs);
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xa[u] = (double)xs[sol_of[u]];
  __syncthreads();
  double ax = 0.0;
  double ratio = ratio_of(block_max(scratch, column_residual(a, b, xa, r, floor_rule ? &ax : nullptr)), bnorm);
  double stop_at = floor_rule ? floor_of(true, block_max(scratch, ax), bnorm) : 0.0;
  double best = ratio;
  int32_t best_steps = 0;
  bool best_in_b = false;
  bool active = max_steps > 0 && !(ratio <= tol) && !(ratio <= stop_at);
  double* cur = xa;
  double* nxt = xb;
  for (int32_t step = 1; active; ++step) {
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xs[rhs_row[u]] = (float)r[u];
    __syncthreads();
    lds_sweep<float>(fwd, lpf, xs);
    lds_sweep<float>(bwd, lpb, xs);
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) nxt[u] = cur[u] + (double)xs[sol_of[u]];
    __syncthreads();
    const double previous = ratio;
    ratio = ratio_of(block_max(scratch, column_residual(a, b, nxt, r, floor_rule ? &ax : nullptr)), bnorm);
    if (floor_rule) stop_at = floor_of(true, block_max(scratch, ax), bnorm);
    if (ratio < best) {
      best = ratio;
      best_steps = step;
      best_in_b = nxt == xb;
    }
    active = !(ratio != ratio || ratio > 0.5 * previous || ratio <= tol || step == max_steps || ratio <= stop_at);
    double* swap = cur;
    cur = nxt;
    nxt = swap;
  }
  if (best_in_b)
    for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xa[u] = xb[u];
  if (threadIdx.x == 0) {
    *steps_out = best_steps;
    *ratio_out = best;
  }
}

__global__ void __launch_bounds__(1024) refine_lds_kernel(SweepDev fwd, SweepDev bwd, CsrDev a,
                                                          const int32_t* __restrict__ rhs_row,
                                                          const int32_t* __restrict__ sol_of,
                                                          const double* __restrict__ bblock, double* __restrict__ x0b,
                                                          double* __restrict__ x1b, double* __restrict__ rblock,
                                                          int64_t ld, double tol, int32_t max_steps,
                                                          int32_t* __restrict__ steps_out,
                                                          double* __restrict__ ratio_out, int32_t floor_rule) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t col = blockIdx.x;
  refine_column(fwd, bwd, a, rhs_row, sol_of, bblock + col * ld, x0b + col * ld, x1b + col * ld, rblock + col * ld, tol,
                max_steps, steps_out + col, ratio_out + col, smem, floor_rule != 0);
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

// ybus_cuda/src/kernels.cu:24-58, over a block by unknown, in two passes. Per column: w =
// inverse_weight - c^T y and q = c^T x / w with x = base + alpha y, or the column flagged singular
// when w == 0. Then every entry: out = x + q y (x alone for a singular column).
__global__ void pair_coefficient_kernel(const double* __restrict__ y, int32_t ncols, int64_t rs, int64_t cs,
                                        const double* __restrict__ base, const int32_t* __restrict__ plus,
                                        const int32_t* __restrict__ minus, const double* __restrict__ alpha,
                                        const double* __restrict__ inverse_weight, double* __restrict__ coef,
                                        int32_t* __restrict__ singular) {
  const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= ncols) return;
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
  singular[k] = w == 0.0 ? 1 : 0;
  coef[k] = w == 0.0 ? 0.0 : cx / w;
}

// Row-interleaved blocks (entry (i, c) at i ld + c): lanes over columns, each thread a run of
// rows_per_block rows of its column
__global__ void pair_apply_rows_kernel(const double* __restrict__ y, double* __restrict__ out, int32_t n, int32_t ncols,
                                       int64_t ld, int32_t rows_per_block, const double* __restrict__ base,
                                       const double* __restrict__ alpha, const double* __restrict__ coef,
                                       const int32_t* __restrict__ singular) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= ncols) return;
  const double a = alpha[c], q = coef[c];
  const bool flat = singular[c] != 0;
  const int32_t i0 = (int32_t)blockIdx.y * rows_per_block, i1 = min(n, i0 + rows_per_block);
  for (int32_t i = i0; i < i1; ++i) {
    const double v = y[i * ld + c];
    const double x = base[i] + a * v;
    out[i * ld + c] = flat ? x : x + q * v;
  }
}

// Column-major blocks (entry (i, c) at i + c ld): threads over rows, columns over grid.y
__global__ void pair_apply_columns_kernel(const double* __restrict__ y, double* __restrict__ out, int32_t n,
                                          int32_t ncols, int64_t ld, const double* __restrict__ base,
                                          const double* __restrict__ alpha, const double* __restrict__ coef,
                                          const int32_t* __restrict__ singular) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  for (int32_t c = (int32_t)blockIdx.y; c < ncols; c += (int32_t)gridDim.y) {
    const double v = y[i + c * ld];
    const double x = base[i] + alpha[c] * v;
    out[i + c * ld] = singular[c] ? x : x + coef[c] * v;
  }
}

// ---- outages of sets of branches: rank-m updates (dc/outage.rs:126-241) ----

// x_p = base_p + alpha_k Y(p, k) for the set's pairs k in order (x_i += alpha y_i)
__device__ __forceinline__ double outage_x(const double* __restrict__ y, int64_t rs, int64_t cs,
                                           const double* __restrict__ base, const double* __restrict__ alpha,
                                           int32_t k0, int32_t k1, int32_t p) {
  double x = base[p];
  for (int32_t k = k0; k < k1; ++k) x = x + alpha[k] * y[p * rs + k * cs];
  return x;
}

// c^T v for pair k: +v[plus] - v[minus] (0.0 for a missing end), as pair_update sums it
__device__ __forceinline__ double pair_dot(int32_t p, int32_t m, double vp, double vm) {
  double d = 0.0;
  if (p >= 0) d += vp;
  if (m >= 0) d -= vm;
  return d;
}

// Per set s (workgroup s): W = D^-1 - C^T Y (m x m, row-major at w_at[s]) and C^T x (m, at
// c_at[s]), where the set's pairs are set_ptr[s] .. set_ptr[s + 1] of the block's columns.
__global__ void outage_system_kernel(const double* __restrict__ y, int64_t rs, int64_t cs,
                                     const double* __restrict__ base, const int32_t* __restrict__ set_ptr,
                                     const int32_t* __restrict__ plus, const int32_t* __restrict__ minus,
                                     const double* __restrict__ alpha, const double* __restrict__ inverse_weight,
                                     const int32_t* __restrict__ w_at, double* __restrict__ w,
                                     double* __restrict__ ctx) {
  const int32_t s = (int32_t)blockIdx.x;
  const int32_t k0 = set_ptr[s], k1 = set_ptr[s + 1], m = k1 - k0;
  double* __restrict__ ws = w + w_at[s];
  for (int32_t e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int32_t i = k0 + e / m, j = k0 + e % m;
    const int32_t p = plus[i], q = minus[i];
    const double vp = p >= 0 ? y[p * rs + j * cs] : 0.0, vq = q >= 0 ? y[q * rs + j * cs] : 0.0;
    double v = -pair_dot(p, q, vp, vq);
    if (i == j) v += inverse_weight[i];
    ws[e] = v;
  }
  for (int32_t i = k0 + threadIdx.x; i < k1; i += blockDim.x) {
    const int32_t p = plus[i], q = minus[i];
    const double vp = p >= 0 ? outage_x(y, rs, cs, base, alpha, k0, k1, p) : 0.0;
    const double vq = q >= 0 ? outage_x(y, rs, cs, base, alpha, k0, k1, q) : 0.0;
    ctx[set_ptr[s] + (i - k0)] = pair_dot(p, q, vp, vq);
  }
}

// out(i, s) = x_i + z_j Y(i, j) over the set's pairs j in order (x alone for a singular set).
// Row-interleaved blocks: lanes over sets, each thread a run of rows_per_block rows; column-major
// blocks: threads over rows, sets over grid.y.
__device__ __forceinline__ double outage_value(const double* __restrict__ y, int64_t rs, int64_t cs,
                                               const double* __restrict__ base, const double* __restrict__ alpha,
                                               const double* __restrict__ z, bool flat, int32_t k0, int32_t k1,
                                               int32_t i) {
  double x = outage_x(y, rs, cs, base, alpha, k0, k1, i);
  if (!flat)
    for (int32_t j = k0; j < k1; ++j) x = x + z[j] * y[i * rs + j * cs];
  return x;
}

__global__ void outage_apply_rows_kernel(const double* __restrict__ y, int64_t ld, int32_t n, int32_t rows_per_block,
                                         const double* __restrict__ base, const int32_t* __restrict__ set_ptr,
                                         const double* __restrict__ alpha, const double* __restrict__ z,
                                         const int32_t* __restrict__ singular, double* __restrict__ out,
                                         int32_t sets) {
  const int32_t s = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= sets) return;
  const int32_t k0 = set_ptr[s], k1 = set_ptr[s + 1];
  const bool flat = singular[s] != 0;
  const int32_t i0 = (int32_t)blockIdx.y * rows_per_block, i1 = min(n, i0 + rows_per_block);
  for (int32_t i = i0; i < i1; ++i) out[i * ld + s] = outage_value(y, ld, 1, base, alpha, z, flat, k0, k1, i);
}

__global__ void outage_apply_columns_kernel(const double* __restrict__ y, int64_t ld, int32_t n,
                                            const double* __restrict__ base, const int32_t* __restrict__ set_ptr,
                                            const double* __restrict__ alpha, const double* __restrict__ z,
                                            const int32_t* __restrict__ singular, double* __restrict__ out,
                                            int32_t sets) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  for (int32_t s = (int32_t)blockIdx.y; s < sets; s += (int32_t)gridDim.y)
    out[i + s * ld] = outage_value(y, 1, ld, base, alpha, z, singular[s] != 0, set_ptr[s], set_ptr[s + 1], i);
}

// A column's running maximum of |v| * weight (kernels.cu:78-128): rows visited in increasing
// order, so a strictly larger product wins; a value that isn't finite sets `bad` and is skipped.
// merge() folds in another run's result, the smaller row winning a tie.
struct MaxAcc {
  double value = 0.0;
  int32_t row = -1, bad = 0;
  __device__ void visit(double v, double w, int32_t i) {
    if (!isfinite(v)) {
      bad = 1;
      return;
    }
    const double product = fabs(v) * w;
    if (product > value) {
      value = product;
      row = i;
    }
  }
  __device__ void merge(double v, int32_t r, int32_t b) {
    if (r >= 0 && (v > value || (v == value && (row < 0 || r < row)))) {
      value = v;
      row = r;
    }
    bad |= b;
  }
};

// Entry (i, c) of a block
struct BlockEntry {
  const double* __restrict__ b;
  int64_t rs, cs;
  __device__ double operator()(int32_t i, int32_t c) const { return b[i * rs + c * cs]; }
};

// Entry (r, c) of M x + offset with row zero[c] of column c zeroed, summed from 0 in index
// order as the host's M.mul_vec(theta) then + offset (pairs.rs:312-366)
struct ProductEntry {
  CsrDev m;
  const double* __restrict__ x;
  int64_t rs, cs;
  const double* __restrict__ offset;
  const int32_t* __restrict__ zero;
  __device__ double operator()(int32_t r, int32_t c) const {
    double sum = 0.0;
    const int32_t pe = m.ptr[r + 1];
    for (int32_t p = m.ptr[r]; p < pe; ++p) sum = sum + m.val[p] * x[m.col[p] * rs + c * cs];
    return zero[c] == r ? 0.0 : sum + offset[r];
  }
};

// Row-interleaved blocks: a workgroup takes 64 columns (one per lane) and one chunk of
// chunk_rows rows, its waves interleaving the rows; the waves' results combine in LDS and each
// (chunk, column) result goes to the partial arrays, chunk-major.
template <class Entry>
__global__ void __launch_bounds__(256) max_rows_kernel(Entry entry, int32_t rows, int32_t ncols, int32_t chunk_rows,
                                                       const double* __restrict__ weight, double* __restrict__ pv,
                                                       int32_t* __restrict__ pr, int32_t* __restrict__ pb) {
  __shared__ double s_value[256];
  __shared__ int32_t s_row[256], s_bad[256];
  const int lane = threadIdx.x % 64, waves = blockDim.x / 64;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int32_t col = (int32_t)blockIdx.x * 64 + lane;
  const int32_t r0 = (int32_t)blockIdx.y * chunk_rows, r1 = min(rows, r0 + chunk_rows);
  MaxAcc acc;
  if (col < ncols)
    for (int32_t i = r0 + wave; i < r1; i += waves) acc.visit(entry(i, col), weight[i], i);
  s_value[threadIdx.x] = acc.value;
  s_row[threadIdx.x] = acc.row;
  s_bad[threadIdx.x] = acc.bad;
  __syncthreads();
  if (wave == 0 && col < ncols) {
    for (int w = 1; w < waves; ++w) acc.merge(s_value[w * 64 + lane], s_row[w * 64 + lane], s_bad[w * 64 + lane]);
    const int64_t at = (int64_t)blockIdx.y * ncols + col;
    pv[at] = acc.value;
    pr[at] = acc.row;
    pb[at] = acc.bad;
  }
}

// Each column's chunks combined in chunk order
__global__ void max_combine_kernel(int32_t ncols, int32_t chunks, const double* __restrict__ pv,
                                   const int32_t* __restrict__ pr, const int32_t* __restrict__ pb,
                                   double* __restrict__ best, int32_t* __restrict__ best_row,
                                   int32_t* __restrict__ not_finite) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= ncols) return;
  MaxAcc acc;
  for (int32_t k = 0; k < chunks; ++k) {
    const int64_t at = (int64_t)k * ncols + c;
    acc.merge(pv[at], pr[at], pb[at]);
  }
  best[c] = acc.value;
  best_row[c] = acc.row;
  not_finite[c] = acc.bad;
}

// Column-major blocks: one workgroup per column, threads over rows (each visiting its rows in
// increasing order), combined in LDS with the smaller row winning a tie
template <class Entry>
__global__ void __launch_bounds__(256) max_columns_kernel(Entry entry, int32_t rows, const double* __restrict__ weight,
                                                          double* __restrict__ best, int32_t* __restrict__ best_row,
                                                          int32_t* __restrict__ not_finite) {
  __shared__ double s_value[256];
  __shared__ int32_t s_row[256], s_bad[256];
  const int32_t col = (int32_t)blockIdx.x;
  MaxAcc acc;
  for (int32_t i = threadIdx.x; i < rows; i += blockDim.x) acc.visit(entry(i, col), weight[i], i);
  s_value[threadIdx.x] = acc.value;
  s_row[threadIdx.x] = acc.row;
  s_bad[threadIdx.x] = acc.bad;
  __syncthreads();
  for (int stride = blockDim.x / 2; stride > 0; stride /= 2) {
    if ((int)threadIdx.x < stride) {
      MaxAcc mine;
      mine.value = s_value[threadIdx.x];
      mine.row = s_row[threadIdx.x];
      mine.bad = s_bad[threadIdx.x];
      mine.merge(s_value[threadIdx.x + stride], s_row[threadIdx.x + stride], s_bad[threadIdx.x + stride]);
      s_value[threadIdx.x] = mine.value;
      s_row[threadIdx.x] = mine.row;
      s_bad[threadIdx.x] = mine.bad;
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
// at i * rs + c * cs and of out at r * ors + c * ocs. Row-interleaved output (ocs = 1): lanes over
// columns and a run of rows per thread; column-major output: threads over rows.
__global__ void operator_rows_kernel(ProductEntry e, int32_t ncols, int32_t rows_per_block, double* __restrict__ out,
                                     int64_t ors) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= ncols) return;
  const int32_t r0 = (int32_t)blockIdx.y * rows_per_block, r1 = min(e.m.rows, r0 + rows_per_block);
  for (int32_t r = r0; r < r1; ++r) out[r * ors + c] = e(r, c);
}

__global__ void operator_columns_kernel(ProductEntry e, int32_t ncols, double* __restrict__ out, int64_t ocs) {
  const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= e.m.rows) return;
  for (int32_t c = (int32_t)blockIdx.y; c < ncols; c += (int32_t)gridDim.y) out[r + c * ocs] = e(r, c);
}

// ---- batches of factors: persistent column solves fed by a device queue ----

// One right-hand side b (contiguous, by unknown) of one factor solved with its solution in LDS:
// smem holds the vector in factor order, then both sweeps' level pointers; out gets x by unknown.
// Every thread must call; the caller's barrier must separate it from the next use of smem.
template <bool G>
__device__ void column_solve(const SweepDev& fwd, const SweepDev& bwd, int32_t n, const int32_t* __restrict__ rhs_row,
                             const int32_t* __restrict__ sol_of, const double* __restrict__ b,
                             double* __restrict__ out, unsigned char* smem) {
  double* xs = reinterpret_cast<double*>(smem);
  int32_t* lpf = reinterpret_cast<int32_t*>(smem + lds_vector_bytes(n, sizeof(double)));
  int32_t* lpb = lpf + fwd.levels + 1;
  stage_level_ptr(fwd, lpf);
  stage_level_ptr(bwd, lpb);
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) xs[rhs_row[u]] = b[u];
  __syncthreads();
  lds_sweep<double, G>(fwd, lpf, xs);
  lds_sweep<double, G>(bwd, lpb, xs);
  for (int32_t u = threadIdx.x; u < n; u += blockDim.x) out[u] = xs[sol_of[u]];
}

// A persistent kernel over a queue of (factor, column) items: each workgroup takes the next item
// (one atomic add on the queue's counter by one thread), solves it, and takes another, until the
// queue is empty; every workgroup then sees an item past the end and returns. The items are
// ordered longest first, so the last ones to start are the shortest. G: the records are loaded
// as global memory (global_at), which keeps more of them in flight and so more registers: up to
// 512 threads, for the classes of one or two workgroups a compute unit; the others keep the loads
// the compiler chooses, in up to 128 registers, so up to 16 small workgroups share a unit.
template <bool G>
__global__ void __launch_bounds__(G ? 512 : 1024) batch_kernel(const BatchEntryDev* __restrict__ entries,
                                                     const int2* __restrict__ items, int32_t total,
                                                     int32_t* __restrict__ queue, const double* __restrict__ rhs,
                                                     double* __restrict__ x) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_item;
  for (;;) {
    __syncthreads();  // the previous item is finished with smem and s_item
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int32_t it = s_item;
    if (it >= total) return;
    const int2 w = items[it];
    const BatchEntryDev e = entries[w.x];
    const int64_t at = e.offset + (int64_t)w.y * e.n;
    column_solve<G>(e.fwd, e.bwd, e.n, e.rhs_row, e.sol_of, rhs + at, x + at, smem);
  }
}

// The batch's FP32-refined solve: the same queue, each item refined like refine_lds_kernel's
// columns (the factor's matrix for the residual, its own iterate and residual buffers, and its
// corrections and ratio at its column's place among all the batch's columns)
__global__ void __launch_bounds__(1024) batch_refine_kernel(const BatchEntryDev* __restrict__ entries,
                                                            const int2* __restrict__ items, int32_t total,
                                                            int32_t* __restrict__ queue, const double* __restrict__ rhs,
                                                            double* __restrict__ x, double* __restrict__ x1,
                                                            double* __restrict__ r, double tol, int32_t max_steps,
                                                            int32_t* __restrict__ steps, double* __restrict__ ratio,
                                                            int32_t floor_rule) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int32_t it = s_item;
    if (it >= total) return;
    const int2 w = items[it];
    const BatchEntryDev e = entries[w.x];
    const int64_t at = e.offset + (int64_t)w.y * e.n;
    const int64_t column = e.first_column + w.y;
    refine_column(e.fwd, e.bwd, e.a, e.rhs_row, e.sol_of, rhs + at, x + at, x1 + at, r + at, tol, max_steps,
                  steps + column, ratio + column, smem, floor_rule != 0);
  }
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int32_t value, int32_t count) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < count) p[i] = value;
}

// ---- adjoint: the gradient with respect to A's stored values (SDDMM on A's pattern) ----

// For each stored entry q = (row[q], col[q]) of A: acc[q] = (first ? 0 : acc[q]) + the sum over
// the block's columns c, in ascending order, of bbar(row, c) * x(col, c). Blocks of columns taken
// one after another therefore give each entry's sum over all columns in one fixed order, with no
// atomics: the result is the same on every run.
__global__ void sddmm_kernel(const int32_t* __restrict__ row, const int32_t* __restrict__ col, int64_t nnz,
                             const double* __restrict__ bbar, const double* __restrict__ x, int64_t rs, int64_t cs,
                             int32_t ncols, double* __restrict__ acc, int32_t first) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const double* __restrict__ bi = bbar + (int64_t)row[q] * rs;
  const double* __restrict__ xj = x + (int64_t)col[q] * rs;
  double sum = first ? 0.0 : acc[q];
  int32_t c = 0;
  if (cs == 1 && rs % 2 == 0) {
    // a row-interleaved block (rows 16-byte aligned): eight columns of each row per round, their
    // loads issued together so a row's line is fetched once rather than once per column, then
    // summed in order
    constexpr int RUN = 8;
    for (; c + RUN <= ncols; c += RUN) {
      double2 b2[RUN / 2], x2[RUN / 2];
#pragma unroll
      for (int k = 0; k < RUN / 2; ++k) {
        b2[k] = *reinterpret_cast<const double2*>(bi + c + 2 * k);
        x2[k] = *reinterpret_cast<const double2*>(xj + c + 2 * k);
      }
#pragma unroll
      for (int k = 0; k < RUN / 2; ++k) {
        sum = sum + b2[k].x * x2[k].x;
        sum = sum + b2[k].y * x2[k].y;
      }
    }
  }
  for (; c < ncols; ++c) sum = sum + bi[c * cs] * xj[c * cs];
  acc[q] = sum;
}

__global__ void scatter_negated_kernel(const double* __restrict__ acc, const int32_t* __restrict__ order, int64_t nnz,
                                       int32_t any, double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nnz) out[order[q]] = -(any ? acc[q] : 0.0);
}


}  // namespace

// ---- launchers ----

// How many XCDs the one-dimensional wide launches spread runs of consecutive tiles over
// (tile_of_workgroup): MI355X has 8. YBUS_HIP_XCDS=1 keeps workgroup i on tile i.
static int32_t xcd_groups() {
  static const int32_t groups = [] {
    const char* e = std::getenv("YBUS_HIP_XCDS");
    const int v = e ? std::atoi(e) : 8;
    return (int32_t)(v >= 1 && v <= 64 ? v : 8);
  }();
  return groups;
}

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
  const size_t bytes = (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t);
#define S(TWV, SPLIT)                                                                                            \
  {                                                                                                              \
    hipError_t e = hipFuncSetAttribute((const void*)solve64_kernel<TWV, SPLIT>,                                  \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);                 \
    if (e != hipSuccess) return e;                                                                               \
    hipLaunchKernelGGL((solve64_kernel<TWV, SPLIT>), dim3(grid), dim3(threads), bytes, stream, fwd, bwd, rhs, t, x, \
                       ld, ncols, xcd_groups());                                                                 \
  }
#define L(TWV)         \
  if (split)           \
    S(TWV, true)       \
  else                 \
    S(TWV, false)
  const bool split = fwd.split || bwd.split;
  DISPATCH_TILE(tile, L)
#undef L
#undef S
  return hipGetLastError();
}

size_t window_lds_bytes(const SweepDev& fwd, const SweepDev& bwd, int tile, int32_t cap) {
  return (size_t)cap * (size_t)tile * sizeof(double) + (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t);
}

int32_t window_cap_for(const SweepDev& fwd, const SweepDev& bwd, int tile, size_t lds_cap) {
  const size_t levels = (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t);
  if (tile < 1 || lds_cap <= levels) return 0;
  return (int32_t)((lds_cap - levels) / ((size_t)tile * sizeof(double)));
}

hipError_t launch_solve64_window(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const WindowDev& wf,
                                 const SweepDev& bwd, const WindowDev& wb, const RhsDev& rhs, double* t, double* x,
                                 int64_t ld, int32_t ncols) {
  if (threads < tile || threads > 512 || threads % 64 || wf.cap < 1 || wb.cap != wf.cap || !wf.head || !wb.head)
    return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
  const size_t bytes = window_lds_bytes(fwd, bwd, tile, wf.cap);
#define L(TWV)                                                                                                  \
  {                                                                                                             \
    hipError_t e = hipFuncSetAttribute((const void*)solve64_window_kernel<TWV>,                                 \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);                \
    if (e != hipSuccess) return e;                                                                              \
    hipLaunchKernelGGL((solve64_window_kernel<TWV>), dim3(grid), dim3(threads), bytes, stream, fwd, wf, bwd, wb, \
                       rhs, t, x, ld, ncols, xcd_groups());                                                     \
  }
  DISPATCH_TILE(tile, L)
#undef L
  return hipGetLastError();
}

hipError_t launch_solve64_parts(hipStream_t stream, int tile, int threads, int top_tile, int top_threads,
                                const SweepDev& fwd, const SweepDev& bwd, const PartLevels& levels, const RhsDev& rhs,
                                const ReachDev& reach, const TopSlotsDev* top_lds, const PartsLds& parts_lds,
                                double* t, double* x, int64_t ld, int32_t ncols) {
  auto valid = [](int tw, int th) {
    return tw >= 1 && tw <= 64 && (tw & (tw - 1)) == 0 && th >= tw && th <= 1024 && th % 64 == 0;
  };
  if (!valid(tile, threads) || !valid(top_tile, top_threads) || fwd.parts < 1 || bwd.parts < 1)
    return hipErrorInvalidValue;
  // the LDS parts: a wavefront per column of the tile, both sweeps' parts, no reach
  const int lds_tile = parts_lds.tile;
  if (lds_tile && (!parts_lds.fwd || !parts_lds.bwd || reach.lp || lds_tile > 16 || !valid(lds_tile, parts_lds.threads) ||
                   parts_lds.threads < 64 * lds_tile || parts_lds.fwd->parts != fwd.parts ||
                   parts_lds.bwd->parts != bwd.parts))
    return hipErrorInvalidValue;
  const int32_t lds_tiles = lds_tile ? (ncols + lds_tile - 1) / lds_tile : 0;
  // split sweeps come with blocks of vectors only (a pair block's reach is over the unsplit
  // ones), and a pair block's reach bitmaps are per tile, so its top kernel keeps the parts' tile
  const bool split = fwd.split || bwd.split;
  if (reach.lp && (split || top_tile != tile || top_lds)) return hipErrorInvalidValue;
  if (top_lds && top_lds->rows < 1) return hipErrorInvalidValue;
  const unsigned tiles = (unsigned)((ncols + tile - 1) / tile), top_tiles = (unsigned)((ncols + top_tile - 1) / top_tile);
  // LDS: level pointers rounded to 4 words (so the bitmap after them stays 16-byte aligned), and
  // the tile's reach bitmap for a pair block
  auto words = [](int32_t levels) { return (levels + 1 + 3) / 4 * 4; };
  const size_t live = reach.lp ? (size_t)reach.words * sizeof(uint32_t) : 0;
  const int32_t wf = words(levels.fwd_parts), wt = words(levels.fwd_top), wb = words(levels.bwd_top),
                wp = words(levels.bwd_parts);
  const size_t b1 = wf * sizeof(int32_t) + live, b2 = (size_t)(wt + wb) * sizeof(int32_t) + live,
               b3 = wp * sizeof(int32_t) + live;
#define LAUNCH(KERNEL, TWV, SPLIT, BYTES, GRID, THREADS, ...)                                                    \
  {                                                                                                             \
    const hipError_t e = hipFuncSetAttribute((const void*)KERNEL<TWV, SPLIT>,                                    \
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)(BYTES));         \
    if (e != hipSuccess) return e;                                                                              \
    hipLaunchKernelGGL((KERNEL<TWV, SPLIT>), GRID, dim3(THREADS), BYTES, stream, __VA_ARGS__);                   \
  }
#define LAUNCH_LDS(TWV)                                                                                          \
  {                                                                                                             \
    const size_t bytes = top_lds_layout(*top_lds, TWV).total;                                                   \
    const hipError_t e = hipFuncSetAttribute((const void*)top_lds_kernel<TWV>,                                  \
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);           \
    if (e != hipSuccess) return e;                                                                              \
    hipLaunchKernelGGL((top_lds_kernel<TWV>), dim3(top_tiles), dim3(top_threads), bytes, stream, fwd, bwd,       \
                       *top_lds, rhs, t, x, ld, ncols);                                                         \
  }
#define LAUNCH_PARTS(TWV, BACKWARD, PS, UNIT, SRC, DST)                                                          \
  {                                                                                                             \
    const size_t bytes = parts_lds_layout(PS, TWV).total;                                                       \
    const hipError_t e = hipFuncSetAttribute((const void*)parts_lds_kernel<TWV, BACKWARD>,                       \
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);           \
    if (e != hipSuccess) return e;                                                                              \
    hipLaunchKernelGGL((parts_lds_kernel<TWV, BACKWARD>), dim3((unsigned)(lds_tiles * (PS).parts)),              \
                       dim3(parts_lds.threads), bytes, stream, PS, UNIT, rhs, SRC, DST, ld, ncols, lds_tiles,    \
                       xcd_groups());                                                                           \
  }
#define FWD_LDS(TWV) LAUNCH_PARTS(TWV, false, *parts_lds.fwd, (int32_t)fwd.unit, nullptr, t)
#define BWD_LDS(TWV) LAUNCH_PARTS(TWV, true, *parts_lds.bwd, (int32_t)bwd.unit, t, x)
#define FWD(TWV)                                                                                                 \
  if (split)                                                                                                    \
    LAUNCH(parts_forward_kernel, TWV, true, b1, dim3(tiles, (unsigned)fwd.parts), threads, fwd, rhs, reach, t,  \
           ld, ncols, wf)                                                                                       \
  else                                                                                                          \
    LAUNCH(parts_forward_kernel, TWV, false, b1, dim3(tiles, (unsigned)fwd.parts), threads, fwd, rhs, reach, t, \
           ld, ncols, wf)
#define TOP(TWV)                                                                                                 \
  if (top_lds)                                                                                                  \
    LAUNCH_LDS(TWV)                                                                                             \
  else if (split)                                                                                               \
    LAUNCH(top_kernel, TWV, true, b2, dim3(top_tiles), top_threads, fwd, bwd, rhs, reach, t, x, ld, ncols, wt,  \
           wb)                                                                                                  \
  else                                                                                                          \
    LAUNCH(top_kernel, TWV, false, b2, dim3(top_tiles), top_threads, fwd, bwd, rhs, reach, t, x, ld, ncols, wt, \
           wb)
#define BWD(TWV)                                                                                                 \
  if (split)                                                                                                    \
    LAUNCH(parts_backward_kernel, TWV, true, b3, dim3(tiles, (unsigned)bwd.parts), threads, bwd, reach, t, x,   \
           ld, ncols, wp)                                                                                       \
  else                                                                                                          \
    LAUNCH(parts_backward_kernel, TWV, false, b3, dim3(tiles, (unsigned)bwd.parts), threads, bwd, reach, t, x,  \
           ld, ncols, wp)
  // the LDS parts take tiles of up to 16 columns (a wavefront each)
#define DISPATCH_LDS_TILE(F)             \
  switch (lds_tile) {                    \
    case 1: F(1); break;                 \
    case 2: F(2); break;                 \
    case 4: F(4); break;                 \
    case 8: F(8); break;                 \
    case 16: F(16); break;               \
    default: return hipErrorInvalidValue; \
  }
  if (lds_tile) {
    DISPATCH_LDS_TILE(FWD_LDS)
  } else {
    DISPATCH_TILE(tile, FWD)
  }
  DISPATCH_TILE(top_tile, TOP)
  if (lds_tile) {
    DISPATCH_LDS_TILE(BWD_LDS)
  } else {
    DISPATCH_TILE(tile, BWD)
  }
#undef BWD
#undef TOP
#undef FWD
#undef DISPATCH_LDS_TILE
#undef BWD_LDS
#undef FWD_LDS
#undef LAUNCH_PARTS
#undef LAUNCH_LDS
#undef LAUNCH
  return hipGetLastError();
}

size_t reach_lds_bytes(const SweepDev& fwd, const SweepDev& bwd) {
  return (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t) + (size_t)(fwd.n / 32 + 1) * sizeof(uint32_t);
}

hipError_t launch_solve64_reach(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                                const RhsDev& rhs, const ReachDev& reach, double* t, double* x, int64_t ld,
                                int32_t ncols, bool skip_loads) {
  if (threads < tile || threads > 1024 || threads % 64) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
  const size_t bytes = reach_lds_bytes(fwd, bwd);
#define S(TWV, SKIP)                                                                                            \
  {                                                                                                             \
    hipError_t e = hipFuncSetAttribute((const void*)solve64_reach_kernel<TWV, SKIP>,                            \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);                 \
    if (e != hipSuccess) return e;                                                                              \
    hipLaunchKernelGGL((solve64_reach_kernel<TWV, SKIP>), dim3(grid), dim3(threads), bytes, stream, fwd, bwd, rhs, \
                       reach, t, x, ld, ncols, xcd_groups());                                                   \
  }
#define L(TWV)         \
  if (skip_loads)      \
    S(TWV, true)       \
  else                 \
    S(TWV, false)
  DISPATCH_TILE(tile, L)
#undef L
#undef S
  return hipGetLastError();
}

hipError_t launch_refine(hipStream_t stream, int tile, int threads, const SweepDev& fwd, const SweepDev& bwd,
                         const CsrDev& a, const double* b, float* t, float* d, double* x0, double* x1, double* r,
                         int64_t ld, int32_t ncols, double tol, int32_t max_steps, int32_t* steps, double* ratio,
                         int32_t floor_rule) {
  if (threads < tile || threads > 1024 || threads % 64 || (threads & (threads - 1))) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
  const size_t bytes = (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t);
#define S(TWV, SPLIT)                                                                                            \
  {                                                                                                              \
    hipError_t e = hipFuncSetAttribute((const void*)refine_kernel<TWV, SPLIT>,                                   \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);                 \
    if (e != hipSuccess) return e;                                                                               \
    hipLaunchKernelGGL((refine_kernel<TWV, SPLIT>), dim3(grid), dim3(threads), bytes, stream, fwd, bwd, a, b, t, d, \
                       x0, x1, r, ld, ncols, tol, max_steps, steps, ratio, xcd_groups(), floor_rule);            \
  }
#define L(TWV)         \
  if (split)           \
    S(TWV, true)       \
  else                 \
    S(TWV, false)
  const bool split = fwd.split || bwd.split;
  DISPATCH_TILE(tile, L)
#undef L
#undef S
  return hipGetLastError();
}

hipError_t launch_solve_complex(hipStream_t stream, int tile, int threads, const ComplexSweepDev& fwd,
                                const ComplexSweepDev& bwd, const double* b, double* t, double* x, int64_t ld,
                                int32_t ncols, bool conjugate) {
  if (threads < tile || threads > 1024 || threads % 64 || ld < ncols) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
  const size_t bytes = (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t);
#define L(TWV)                                                                                             \
  {                                                                                                          \
    hipError_t e = hipFuncSetAttribute((const void*)solve_complex_kernel<TWV>,                               \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);              \
    if (e != hipSuccess) return e;                                                                           \
    hipLaunchKernelGGL((solve_complex_kernel<TWV>), dim3(grid), dim3(threads), bytes, stream, fwd, bwd, b, t, x, \
                       ld, ncols, (int32_t)conjugate, xcd_groups());                                         \
  }
  DISPATCH_TILE(tile, L)
#undef L
  return hipGetLastError();
}

hipError_t launch_refine_complex(hipStream_t stream, int tile, int threads, const ComplexSweepDev& fwd,
                                 const ComplexSweepDev& bwd, const ZCsrDev& a, const double* b, float* t, float* d,
                                 double* x0, double* x1, double* r, int64_t ld, int32_t ncols, double tol,
                                 int32_t max_steps, int32_t* steps, double* ratio, int32_t floor_rule) {
  if (threads < tile || threads > 1024 || threads % 64 || (threads & (threads - 1)) || ld < ncols ||
      !fwd.val32 || !bwd.val32)
    return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((ncols + tile - 1) / tile);
  const size_t bytes = (size_t)(fwd.levels + bwd.levels + 2) * sizeof(int32_t);
#define L(TWV)                                                                                                   \
  {                                                                                                              \
    hipError_t e = hipFuncSetAttribute((const void*)refine_complex_kernel<TWV>,                                   \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);                  \
    if (e != hipSuccess) return e;                                                                               \
    hipLaunchKernelGGL((refine_complex_kernel<TWV>), dim3(grid), dim3(threads), bytes, stream, fwd, bwd, a, b, t, \
                       d, x0, x1, r, ld, ncols, tol, max_steps, steps, ratio, xcd_groups(), floor_rule);         \
  }
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

// Rows per thread in the row-interleaved elementwise kernels: enough that grid.y stays within
// its limit, at least 16 so each thread streams a run of rows
static int32_t rows_per_thread(int32_t rows) { return std::max<int32_t>(16, (rows + 65534) / 65535); }

hipError_t launch_pair_update(hipStream_t stream, const double* y, double* out, int32_t n, int32_t ncols, int64_t rs,
                              int64_t cs, const double* base, const int32_t* plus, const int32_t* minus,
                              const double* alpha, const double* inverse_weight, int32_t* singular, double* coef) {
  if (n <= 0 || ncols <= 0) return hipSuccess;
  if (cs != 1 && rs != 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pair_coefficient_kernel, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, stream, y, ncols,
                     rs, cs, base, plus, minus, alpha, inverse_weight, coef, singular);
  if (cs == 1) {
    const int32_t per = rows_per_thread(n);
    const dim3 grid((unsigned)((ncols + 255) / 256), (unsigned)((n + per - 1) / per));
    hipLaunchKernelGGL(pair_apply_rows_kernel, grid, dim3(256), 0, stream, y, out, n, ncols, rs, per, base, alpha,
                       coef, singular);
  } else {
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)std::min<int32_t>(ncols, 65535));
    hipLaunchKernelGGL(pair_apply_columns_kernel, grid, dim3(256), 0, stream, y, out, n, ncols, cs, base, alpha, coef,
                       singular);
  }
  return hipGetLastError();
}

hipError_t launch_outage_system(hipStream_t stream, const double* y, int64_t rs, int64_t cs, const double* base,
                                const int32_t* set_ptr, int32_t sets, const int32_t* plus, const int32_t* minus,
                                const double* alpha, const double* inverse_weight, const int32_t* w_at, double* w,
                                double* ctx) {
  if (sets <= 0) return hipSuccess;
  hipLaunchKernelGGL(outage_system_kernel, dim3((unsigned)sets), dim3(256), 0, stream, y, rs, cs, base, set_ptr, plus,
                     minus, alpha, inverse_weight, w_at, w, ctx);
  return hipGetLastError();
}

hipError_t launch_outage_apply(hipStream_t stream, const double* y, int64_t rs, int64_t cs, int32_t n,
                               const double* base, const int32_t* set_ptr, const double* alpha, const double* z,
                               const int32_t* singular, double* out, int64_t ors, int64_t ocs, int32_t sets) {
  if (sets <= 0 || n <= 0) return hipSuccess;
  // the output shares the block's layout
  if (ors != rs || ocs != cs) return hipErrorInvalidValue;
  if (cs == 1) {
    const int32_t per = rows_per_thread(n);
    const dim3 grid((unsigned)((sets + 255) / 256), (unsigned)((n + per - 1) / per));
    hipLaunchKernelGGL(outage_apply_rows_kernel, grid, dim3(256), 0, stream, y, rs, n, per, base, set_ptr, alpha, z,
                       singular, out, sets);
  } else if (rs == 1) {
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)std::min<int32_t>(sets, 65535));
    hipLaunchKernelGGL(outage_apply_columns_kernel, grid, dim3(256), 0, stream, y, cs, n, base, set_ptr, alpha, z,
                       singular, out, sets);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int32_t max_chunks(int32_t rows, int32_t ncols, int cus) {
  const int64_t groups = (ncols + 63) / 64;
  const int64_t want = (8 * (int64_t)cus + groups - 1) / groups;
  return (int32_t)std::clamp<int64_t>(want, 1, std::max<int64_t>(1, (rows + 255) / 256));
}

template <class Entry>
static hipError_t launch_max(hipStream_t stream, const Entry& e, int32_t rows, int32_t ncols, bool column_major,
                             const double* weight, int32_t chunks, double* pv, int32_t* pr, int32_t* pb, double* best,
                             int32_t* best_row, int32_t* not_finite) {
  if (ncols <= 0) return hipSuccess;
  if (column_major) {
    hipLaunchKernelGGL((max_columns_kernel<Entry>), dim3((unsigned)ncols), dim3(256), 0, stream, e, rows, weight, best,
                       best_row, not_finite);
    return hipGetLastError();
  }
  if (chunks < 1) return hipErrorInvalidValue;
  const int32_t chunk_rows = (rows + chunks - 1) / chunks;
  const dim3 grid((unsigned)((ncols + 63) / 64), (unsigned)chunks);
  hipLaunchKernelGGL((max_rows_kernel<Entry>), grid, dim3(256), 0, stream, e, rows, ncols, chunk_rows, weight, pv, pr,
                     pb);
  hipLaunchKernelGGL(max_combine_kernel, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, stream, ncols, chunks, pv,
                     pr, pb, best, best_row, not_finite);
  return hipGetLastError();
}

hipError_t launch_column_max(hipStream_t stream, const double* block, int32_t rows, int32_t ncols, int64_t rs,
                             int64_t cs, const double* weight, int32_t chunks, double* pv, int32_t* pr, int32_t* pb,
                             double* best, int32_t* best_row, int32_t* not_finite) {
  if (rs != 1 && cs != 1) return hipErrorInvalidValue;
  return launch_max(stream, BlockEntry{block, rs, cs}, rows, ncols, rs == 1, weight, chunks, pv, pr, pb, best, best_row,
                    not_finite);
}

hipError_t launch_product_max(hipStream_t stream, const CsrDev& m, const double* x, int64_t rs, int64_t cs,
                              const double* offset, const int32_t* zero, const double* weight, int32_t ncols,
                              int32_t chunks, double* pv, int32_t* pr, int32_t* pb, double* best, int32_t* best_row,
                              int32_t* not_finite) {
  if (rs != 1 && cs != 1) return hipErrorInvalidValue;
  return launch_max(stream, ProductEntry{m, x, rs, cs, offset, zero}, m.rows, ncols, rs == 1, weight, chunks, pv, pr,
                    pb, best, best_row, not_finite);
}

// The batch kernel a class takes: global record loads for one or two workgroups a compute unit
static const void* batch_kernel_for(int per_cu) {
  return per_cu <= 2 ? (const void*)batch_kernel<true> : (const void*)batch_kernel<false>;
}

hipError_t prepare_batch(size_t lds_bytes, int per_cu) {
  return hipFuncSetAttribute(batch_kernel_for(per_cu), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

hipError_t launch_batch(hipStream_t stream, int grid, int threads, size_t lds_bytes, int per_cu,
                        const BatchEntryDev* entries, const int2* items, int32_t total, int32_t* queue,
                        const double* rhs, double* x) {
  if (grid <= 0 || threads < 64 || threads > (per_cu <= 2 ? 512 : 1024) || threads % 64) return hipErrorInvalidValue;
  if (per_cu <= 2)
    hipLaunchKernelGGL(batch_kernel<true>, dim3(grid), dim3(threads), lds_bytes, stream, entries, items, total, queue,
                       rhs, x);
  else
    hipLaunchKernelGGL(batch_kernel<false>, dim3(grid), dim3(threads), lds_bytes, stream, entries, items, total, queue,
                       rhs, x);
  return hipGetLastError();
}

hipError_t prepare_batch_refine(size_t lds_bytes) {
  return hipFuncSetAttribute((const void*)batch_refine_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds_bytes);
}

hipError_t launch_batch_refine(hipStream_t stream, int grid, int threads, size_t lds_bytes,
                               const BatchEntryDev* entries, const int2* items, int32_t total, int32_t* queue,
                               const double* rhs, double* x, double* x1, double* r, double tol, int32_t max_steps,
                               int32_t* steps, double* ratio, int32_t floor_rule) {
  if (grid <= 0 || threads < 64 || threads > 1024 || (threads & (threads - 1))) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_refine_kernel, dim3(grid), dim3(threads), lds_bytes, stream, entries, items, total, queue,
                     rhs, x, x1, r, tol, max_steps, steps, ratio, floor_rule);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t stream, int32_t* p, int32_t value, int32_t count) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, p, value, count);
  return hipGetLastError();
}

hipError_t launch_sddmm(hipStream_t stream, const int32_t* row, const int32_t* col, int64_t nnz, const double* bbar,
                        const double* x, int64_t rs, int64_t cs, int32_t ncols, double* acc, bool first) {
  if (nnz <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((nnz + 255) / 256);
  hipLaunchKernelGGL(sddmm_kernel, dim3(grid), dim3(256), 0, stream, row, col, nnz, bbar, x, rs, cs, ncols, acc,
                     first ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_scatter_negated(hipStream_t stream, const double* acc, const int32_t* order, int64_t nnz, bool any,
                                  double* out) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_negated_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, stream, acc, order, nnz,
                     any ? 1 : 0, out);
  return hipGetLastError();
}

hipError_t launch_operator(hipStream_t stream, const CsrDev& m, const double* x, int64_t rs, int64_t cs,
                           const double* offset, const int32_t* zero, double* out, int64_t ors, int64_t ocs,
                           int32_t ncols) {
  if (m.rows <= 0 || ncols <= 0) return hipSuccess;
  const ProductEntry e{m, x, rs, cs, offset, zero};
  if (ocs == 1) {
    const int32_t per = rows_per_thread(m.rows);
    const dim3 grid((unsigned)((ncols + 255) / 256), (unsigned)((m.rows + per - 1) / per));
    hipLaunchKernelGGL(operator_rows_kernel, grid, dim3(256), 0, stream, e, ncols, per, out, ors);
  } else if (ors == 1) {
    const dim3 grid((unsigned)((m.rows + 255) / 256), (unsigned)std::min<int32_t>(ncols, 65535));
    hipLaunchKernelGGL(operator_columns_kernel, grid, dim3(256), 0, stream, e, ncols, out, ocs);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t solve_lds_bytes(int32_t n, bool fp32, int32_t levels) {
  return lds_vector_bytes(n, fp32 ? sizeof(float) : sizeof(double)) + (size_t)(levels + 2) * sizeof(int32_t);
}

size_t refine_lds_bytes(int32_t n, int threads, int32_t levels) {
  return (size_t)threads * sizeof(double) + lds_vector_bytes(n, sizeof(float)) + (size_t)(levels + 2) * sizeof(int32_t);
}

hipError_t launch_solve_lds(hipStream_t stream, int threads, bool fp32, const SweepDev& fwd, const SweepDev& bwd,
                            int32_t n, const int32_t* rhs_row, const int32_t* sol_of, const ColumnsDev& rhs,
                            double* x, int64_t ldx, int32_t ncols) {
  if (threads < 64 || threads > 1024 || threads % 64) return hipErrorInvalidValue;
  const size_t bytes = solve_lds_bytes(n, fp32, fwd.levels + bwd.levels);
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
                             int32_t max_steps, int32_t* steps, double* ratio, int32_t floor_rule) {
  if (threads < 64 || threads > 1024 || (threads & (threads - 1))) return hipErrorInvalidValue;
  const size_t bytes = refine_lds_bytes(a.rows, threads, fwd.levels + bwd.levels);
  hipError_t e = hipFuncSetAttribute((const void*)refine_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(refine_lds_kernel, dim3(ncols), dim3(threads), bytes, stream, fwd, bwd, a, rhs_row, sol_of, b, x0,
                     x1, r, ld, tol, max_steps, steps, ratio, floor_rule);
  return hipGetLastError();
}

// ---- LU refactorization (lu.rs:662-757) ----

// LDS of one wavefront's column: x (x_cap values), the staged L values and their target slots,
// and each staged step's first entry and L position
size_t refactor_lds_bytes(int32_t x_cap, int32_t stage, bool complex) {
  const size_t value = complex ? 2 * sizeof(double) : sizeof(double);
  return (size_t)x_cap * value + (size_t)stage * (value + sizeof(int32_t)) + 2 * 64 * sizeof(int32_t);
}

int32_t refactor_stage_for(int32_t longest_l) {
  int32_t stage = 64;
  while (stage < longest_l && stage < REFACTOR_STAGE) stage *= 2;
  return stage;
}

// Columns hand their L values over between workgroups anywhere on the device (other XCDs, other
// L2s) through device-coherent accesses: the values are stored and loaded as device-scope atomics,
// which bypass the caches a device-scope fence would have to write back or invalidate. A value is
// its own flag: every L value holds REFACTOR_PENDING until its column writes it, so a reading
// column needs no other signal and sees each value as soon as it lands.
__device__ inline double shared_value(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void share_value(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool pending(double v) { return (uint64_t)__double_as_longlong(v) == REFACTOR_PENDING; }

// The refactor's arithmetic on its element type. Real: x - l u and x / p in FP64, |x| the absolute
// value. Complex (num-complex 0.4's Complex<f64>, as the reference's Lu<Complex64>): the product
// then the difference, its quotient formula, and |x| = hypot(re, im) (ComplexFloat::abs is the
// norm); a complex value crosses between wavefronts as its two parts, each its own device-scope
// access, and counts as written only once both are.
struct RealValues {
  using V = double;
  static __device__ V zero() { return 0.0; }
  static __device__ V one() { return 1.0; }
  static __device__ V sub_mul(V x, V l, V u) { return x - l * u; }
  static __device__ V quotient(V x, V p) { return x / p; }
  static __device__ double magnitude(V x) { return fabs(x); }
  static __device__ bool is_zero(V x) { return x == 0.0; }
  static __device__ V load(const V* p) { return shared_value(p); }
  static __device__ void store(V* p, V v) { share_value(p, v); }
  static __device__ bool unwritten(V v) { return pending(v); }
};
struct ComplexValues {
  using V = double2;
  static __device__ V zero() { return make_double2(0.0, 0.0); }
  static __device__ V one() { return make_double2(1.0, 0.0); }
  static __device__ V sub_mul(V x, V l, V u) { return zsub(x, zmul(l, u)); }
  static __device__ V quotient(V x, V p) { return zdiv(x, p); }
  static __device__ double magnitude(V x) { return hypot(x.x, x.y); }
  static __device__ bool is_zero(V x) { return x.x == 0.0 && x.y == 0.0; }
  static __device__ V load(const V* p) {
    const double* q = reinterpret_cast<const double*>(p);
    return make_double2(shared_value(q), shared_value(q + 1));
  }
  static __device__ void store(V* p, V v) {
    double* q = reinterpret_cast<double*>(p);
    share_value(q, v.x);
    share_value(q + 1, v.y);
  }
  static __device__ bool unwritten(V v) { return pending(v.x) || pending(v.y); }
};

// One wavefront per column, columns taken from control[0] in level order, so every column a
// wavefront waits on was taken earlier by a wavefront that is running (each wait ends). The steps
// of a column go in batches: the L values and target slots of the next (up to 64) rows of U(:, k)
// that fit the stage are loaded together; the steps whose values are all written run, in order, on
// x in LDS, each a broadcast read of u_jk and one subtract per entry of L(:, j) (distinct slots,
// all after j's), the next step's entries read while this one's update is made; the others are
// loaded again in the next batch. (A single workgroup is one wavefront, so __syncthreads orders
// its LDS accesses and costs no barrier.) A wait that outlasts spin_limit polls gives the column
// up (counted in control[2]); its L values are still written (zeros), so no wavefront waits
// forever.
template <int R, class Val>
__global__ __launch_bounds__(64) void refactor_kernel(RefactorDev d) {
  using V = typename Val::V;
  constexpr int32_t STAGE = 64 * R;
  extern __shared__ double refactor_lds[];
  V* xs = reinterpret_cast<V*>(refactor_lds);
  V* sval = xs + d.x_cap;
  int32_t* stgt = reinterpret_cast<int32_t*>(sval + STAGE);
  int32_t* s_off = stgt + STAGE;
  int32_t* s_first = s_off + 64;
  const V* a_val = reinterpret_cast<const V*>(d.a_val);
  V* l_val = reinterpret_cast<V*>(d.l_val);
  V* u_val = reinterpret_cast<V*>(d.u_val);
  const int lane = threadIdx.x;
  for (;;) {
    int32_t c = 0;
    if (lane == 0) c = atomicAdd(&d.control[0], 1);
    c = __shfl(c, 0, 64);
    if (c >= d.n) break;
    const int4 head = d.head[c], span = d.span[c];
    const int32_t k = head.x, u0 = head.y, nu = head.z, l0 = head.w, below = span.x;
    const int32_t steps = nu - 1, m = nu + below;
    for (int32_t i = lane; i < m; i += 64) xs[i] = Val::zero();
    __syncthreads();
    for (int32_t p = span.y + lane; p < span.z; p += 64) xs[d.a_slot[p]] = a_val[p];
    __syncthreads();
    int32_t e = span.w;
    int32_t t = 0;
    bool abandoned = false;
    while (t < steps) {
      // the candidates: the next steps whose entries fit the stage
      const int32_t s = t + lane;
      int32_t len = 0, first = 0;
      if (s < steps) {
        const int4 st = d.step[u0 + s];
        first = st.y;
        len = st.z;
      }
      int32_t incl = len;
      for (int off = 1; off < 64; off <<= 1) {
        const int32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
      }
      const uint64_t fits = __ballot(s < steps && incl <= STAGE);
      const int32_t cand = fits == ~0ull ? 64 : (int32_t)__builtin_ctzll(~fits);
      if (cand == 0) {
        // longer than the stage (the host admits no such factor)
        abandoned = true;
        break;
      }
      const int32_t total = __shfl(incl, cand - 1, 64);
      if (lane < cand) {
        s_off[lane] = incl - len;
        s_first[lane] = first;
      }
      __syncthreads();
      V v[R];
      int32_t g[R], at[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int32_t f = lane + 64 * r;
        at[r] = 0;
        if (f < total) {
          // the step entry f belongs to: the last one whose offset is at most f
          int32_t lo = 0, hi = cand - 1;
          while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= f)
              lo = mid;
            else
              hi = mid - 1;
          }
          at[r] = s_first[lo] + (f - s_off[lo]);
          v[r] = Val::load(&l_val[at[r]]);
          g[r] = d.target[e + f];
        }
      }
      // until the first step's values are all written: the steps whose entries are all written
      // (a prefix, incl growing with the step), the unwritten values loaded again while none is
      int32_t count = 0;
      for (uint32_t polls = 0;; ++polls) {
        int32_t waiting = total;  // the first entry whose value is not written yet
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int32_t f = lane + 64 * r;
          if (f < total && Val::unwritten(v[r]) && f < waiting) waiting = f;
        }
        for (int off = 32; off > 0; off >>= 1) waiting = min(waiting, __shfl_xor(waiting, off, 64));
        count = __popcll(__ballot(lane < cand && incl <= waiting));
        if (count > 0 || polls > d.spin_limit) break;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int32_t f = lane + 64 * r;
          if (f < total && Val::unwritten(v[r])) v[r] = Val::load(&l_val[at[r]]);
        }
      }
      if (count == 0) {
        abandoned = true;
        break;
      }
      const int32_t used = __shfl(incl, count - 1, 64);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int32_t f = lane + 64 * r;
        if (f < used) {
          sval[f] = v[r];
          stgt[f] = g[r];
        }
      }
      __syncthreads();
      // step q's first 64 entries are read one step ahead; entries past 64 (a long column of L)
      // are read in their own step
      int32_t o0 = s_off[0], o1 = count > 1 ? s_off[1] : used;
      bool has = o0 + lane < o1;
      int32_t slot = has ? stgt[o0 + lane] : 0;
      V lv = has ? sval[o0 + lane] : Val::zero();
      for (int32_t q = 0; q < count; ++q) {
        const V u = xs[t + q];
        const V xv = has ? xs[slot] : Val::zero();
        const int32_t n0 = o1, n1 = q + 2 < count ? s_off[q + 2] : used;
        const bool next_has = q + 1 < count && n0 + lane < n1;
        const int32_t next_slot = next_has ? stgt[n0 + lane] : 0;
        const V next_lv = next_has ? sval[n0 + lane] : Val::zero();
        if (has) xs[slot] = Val::sub_mul(xv, lv, u);
        for (int32_t f = o0 + 64 + lane; f < o1; f += 64) {
          const int32_t sl = stgt[f];
          xs[sl] = Val::sub_mul(xs[sl], sval[f], u);
        }
        __syncthreads();
        o0 = n0;
        o1 = n1;
        has = next_has;
        slot = next_slot;
        lv = next_lv;
      }
      t += count;
      e += used;
    }
    if (!abandoned) {
      // lu.rs:717-733: the reused pivot against the largest magnitude below it
      const V pivot = xs[nu - 1];
      double big = -1.0;
      for (int32_t i = lane; i < below; i += 64) {
        const double b = Val::magnitude(xs[nu + i]);
        if (b > big) big = b;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(big, off, 64);
        if (o > big) big = o;
      }
      double largest = Val::magnitude(pivot);
      if (big > largest) largest = big;
      // the values other columns wait for go first
      for (int32_t i = lane; i < below; i += 64) Val::store(&l_val[l0 + 1 + i], Val::quotient(xs[nu + i], pivot));
      if (lane == 0) {
        if (Val::is_zero(pivot) || Val::magnitude(pivot) < d.tau * largest) atomicMin(&d.control[1], k);
        l_val[l0] = Val::one();
      }
      // x's U slots hold u_jk (no step writes a slot at or before its own), the last one the pivot
      for (int32_t i = lane; i < nu; i += 64) u_val[u0 + i] = xs[i];
    } else {
      if (lane == 0) atomicAdd(&d.control[2], 1);
      for (int32_t i = lane; i < below; i += 64) Val::store(&l_val[l0 + 1 + i], Val::zero());
    }
    __syncthreads();
  }
}

__global__ void fill_u64_kernel(uint64_t* __restrict__ p, uint64_t value, int64_t count) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < count) p[i] = value;
}

__global__ void gather_kernel(const double* __restrict__ src, const int32_t* __restrict__ idx, int64_t count,
                              double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[idx[i]];
}

hipError_t launch_gather(hipStream_t stream, const double* src, const int32_t* idx, int64_t count, double* dst) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, src, idx, count, dst);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t stream, uint64_t* p, uint64_t value, int64_t count) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, p, value, count);
  return hipGetLastError();
}

template <class Val>
static hipError_t launch_refactor_as(hipStream_t stream, const RefactorDev& d, int grid, int32_t stage, size_t bytes) {
  switch (stage) {
    case 64: hipLaunchKernelGGL((refactor_kernel<1, Val>), dim3((unsigned)grid), dim3(64), bytes, stream, d); break;
    case 128: hipLaunchKernelGGL((refactor_kernel<2, Val>), dim3((unsigned)grid), dim3(64), bytes, stream, d); break;
    case 256: hipLaunchKernelGGL((refactor_kernel<4, Val>), dim3((unsigned)grid), dim3(64), bytes, stream, d); break;
    case 512: hipLaunchKernelGGL((refactor_kernel<8, Val>), dim3((unsigned)grid), dim3(64), bytes, stream, d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_refactor(hipStream_t stream, const RefactorDev& d, int grid, int32_t stage, bool complex) {
  if (d.n <= 0) return hipSuccess;
  const size_t bytes = refactor_lds_bytes(d.x_cap, stage, complex);
  if (grid <= 0 || bytes > 65536 - 256) return hipErrorInvalidValue;
  return complex ? launch_refactor_as<ComplexValues>(stream, d, grid, stage, bytes)
                 : launch_refactor_as<RealValues>(stream, d, grid, stage, bytes);
}

__global__ void refresh_entries_kernel(RefreshDev r, const double* __restrict__ src) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= r.entries) return;
  const double v = src[r.vpos[p]];
  r.val64[p] = v;
  r.val32[p] = (float)v;
}

__global__ void refresh_rows_kernel(RefreshDev r, const double* __restrict__ src) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= r.n) return;
  const int32_t p0 = r.row_ptr[k], p1 = r.row_ptr[k + 1];
  double v[RECORD_ENTRIES];
#pragma unroll
  for (int q = 0; q < RECORD_ENTRIES; ++q) v[q] = p0 + q < p1 ? src[r.vpos[p0 + q]] : 0.0;
  r.val01[k] = make_double2(v[0], v[1]);
  r.val23[k] = make_double2(v[2], v[3]);
  r.v32[k] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
  if (r.dpos) {
    // a split sweep's continuation segment divides by 1 (no source position)
    const double dg = r.dpos[k] >= 0 ? src[r.dpos[k]] : 1.0;
    r.diag64[k] = dg;
    r.diag32[k] = (float)dg;
    r.rdiag64[k] = dg;
    r.rdiag32[k] = (float)dg;
  }
}

// A complex sweep's values (and their FP32 copies, where kept) from the refactored ones
__global__ void refresh_complex_kernel(int32_t n, int64_t entries, const int32_t* __restrict__ vpos,
                                       const int32_t* __restrict__ dpos, const double2* __restrict__ src,
                                       double2* __restrict__ val, double2* __restrict__ diag,
                                       float2* __restrict__ val32, float2* __restrict__ diag32) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < entries) {
    const double2 v = src[vpos[i]];
    val[i] = v;
    if (val32) val32[i] = make_float2((float)v.x, (float)v.y);
  }
  if (dpos && i < n) {
    const double2 v = src[dpos[i]];
    diag[i] = v;
    if (diag32) diag32[i] = make_float2((float)v.x, (float)v.y);
  }
}

hipError_t launch_refresh_complex(hipStream_t stream, int32_t n, int64_t entries, const int32_t* vpos,
                                  const int32_t* dpos, const double2* src, double2* val, double2* diag, float2* val32,
                                  float2* diag32) {
  const int64_t count = std::max<int64_t>(entries, dpos ? n : 0);
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(refresh_complex_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, n, entries,
                     vpos, dpos, src, val, diag, val32, diag32);
  return hipGetLastError();
}

hipError_t launch_refresh(hipStream_t stream, const RefreshDev& r, const double* src) {
  if (r.entries > 0)
    hipLaunchKernelGGL(refresh_entries_kernel, dim3((unsigned)((r.entries + 255) / 256)), dim3(256), 0, stream, r,
                       src);
  if (r.n > 0)
    hipLaunchKernelGGL(refresh_rows_kernel, dim3((unsigned)((r.n + 255) / 256)), dim3(256), 0, stream, r, src);
  return hipGetLastError();
}

// The no-allocation workspace: one array per device, its address looked up once per device
static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

hipError_t scratch_address(int device, void** out) {
  static std::mutex lock;
  static std::map<int, void*> address;
  std::lock_guard<std::mutex> hold(lock);
  auto it = address.find(device);
  if (it != address.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  address[device] = p;
  *out = p;
  return hipSuccess;
}

}  // namespace ybus_hip
