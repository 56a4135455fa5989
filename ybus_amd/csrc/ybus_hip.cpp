// The C ABI of ybus_hip (include/ybus_hip.h): contexts, factor upload and analysis, device
// workspaces, and the solve entry points, over the kernels in kernels.hip. Every entry point
// catches its own failures and reports them as a status and a thread-local message.
#include "ybus_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <optional>
#include <thread>
#include <vector>

#include "analysis.hpp"
#include "kernels.hpp"

using namespace ybus_hip;

namespace {

thread_local std::string last_error;

struct Failure : std::runtime_error {
  int32_t status;
  Failure(int32_t s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void fail(int32_t status, const std::string& message) { throw Failure(status, message); }

void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(YBUS_HIP_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  fail(YBUS_HIP_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int32_t guard(F f) {
  try {
    f();
    return YBUS_HIP_OK;
  } catch (const Failure& e) {
    last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    last_error = "host allocation failed";
    return YBUS_HIP_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    last_error = e.what();
    return YBUS_HIP_INTERNAL;
  }
}

// Device memory that frees itself, or (lent) a view of memory it does not own
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t len = 0;
  bool lent = false;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && !lent) (void)hipFree(p);
    p = nullptr;
    len = 0;
    lent = false;
  }
  // At least `n` elements (contents not kept when it grows); a lent view never grows
  void reserve(size_t n) {
    if (n <= len && p) return;
    if (lent) fail(YBUS_HIP_INTERNAL, "a lent workspace buffer was asked to grow");
    release();
    check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    len = n;
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    reserve(v.size());
    if (!v.empty()) check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
};

int64_t host_threads();

// A host-to-host copy split among the library's host threads (at least 4 MB each): a large copy
// into memory the caller has not touched yet is bound by page faults, which threads take in
// parallel
void parallel_copy(void* dst, const void* src, size_t bytes) {
  const int64_t threads = std::min<int64_t>(host_threads(), (int64_t)(bytes >> 22));
  if (threads <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes + (size_t)threads - 1) / (size_t)threads;
  std::vector<std::thread> pool;
  for (int64_t t = 1; t < threads; ++t) {
    const size_t lo = (size_t)t * per, hi = std::min(bytes, lo + per);
    if (lo < hi)
      pool.emplace_back([=] { std::memcpy((char*)dst + lo, (const char*)src + lo, hi - lo); });
  }
  std::memcpy(dst, src, std::min(bytes, per));
  for (std::thread& th : pool) th.join();
}

// Page-locked host staging, grown as needed (factor.rs:63-89). Large transfers go through it in
// two alternating chunks: the DMA of one chunk runs while the host threads copy the other
// between the chunk and the caller's memory, so a transfer costs about the longer of the two
// rather than their sum, and the pinned buffer stays at two chunks.
struct Staging {
  static constexpr size_t CHUNK = size_t(32) << 20;  // bytes
  double* p = nullptr;
  size_t len = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Staging() {
    if (p) (void)hipHostFree(p);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  double* get(size_t n) {
    if (n > len || !p) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      len = 0;
      check(hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
      len = n;
    }
    return p;
  }
  char* chunks(size_t bytes) {
    for (hipEvent_t& e : ev)
      if (!e) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    const size_t need = bytes > CHUNK ? 2 * CHUNK : bytes;
    return reinterpret_cast<char*>(get((need + sizeof(double) - 1) / sizeof(double)));
  }
  // `bytes` from device memory at src into any host memory at dst, after the stream's earlier work
  void to_host(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    char* pin = chunks(bytes);
    const size_t count = (bytes + CHUNK - 1) / CHUNK;
    for (size_t i = 0; i <= count; ++i) {
      if (i < count) {
        const size_t off = i * CHUNK;
        check(hipMemcpyAsync(pin + (i % 2) * CHUNK, (const char*)src + off, std::min(CHUNK, bytes - off),
                             hipMemcpyDeviceToHost, s),
              "D2H");
        check(hipEventRecord(ev[i % 2], s), "hipEventRecord");
      }
      if (i > 0) {
        const size_t j = i - 1, off = j * CHUNK;
        check(hipEventSynchronize(ev[j % 2]), "D2H");
        parallel_copy((char*)dst + off, pin + (j % 2) * CHUNK, std::min(CHUNK, bytes - off));
      }
    }
  }
  // `bytes` from any host memory at src into device memory at dst; returns with the copy done
  void to_device(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    char* pin = chunks(bytes);
    const size_t count = (bytes + CHUNK - 1) / CHUNK;
    for (size_t i = 0; i < count; ++i) {
      const size_t off = i * CHUNK, n = std::min(CHUNK, bytes - off);
      // the DMA that last read this chunk is done before the host writes it again
      if (i >= 2) check(hipEventSynchronize(ev[i % 2]), "H2D");
      parallel_copy(pin + (i % 2) * CHUNK, (const char*)src + off, n);
      check(hipMemcpyAsync((char*)dst + off, pin + (i % 2) * CHUNK, n, hipMemcpyHostToDevice, s), "H2D");
      check(hipEventRecord(ev[i % 2], s), "hipEventRecord");
    }
    check(hipStreamSynchronize(s), "H2D");
  }
};

// A sweep on the device: its level pointers, its rows' records (analysis order, field by
// field), and the CSR arrays holding every entry (the records inline the first few of each row)
// A sweep's sources encoded for the windowed solve with a ring of `cap` positions (window_code)
struct DevWindow {
  DevBuf<int4> head, dep4;
  DevBuf<int32_t> dep;
  WindowDev dev;
  void build(const Sweep& s, int32_t cap, hipStream_t stream) {
    const WindowCode c = window_code(s, cap);
    const int32_t count = s.positions();
    std::vector<int4> hd(std::max(count, 1)), d4(std::max(count, 1));
    for (int32_t k = 0; k < count; ++k) {
      const int32_t p0 = s.row_ptr[k], p1 = s.row_ptr[k + 1];
      hd[k] = make_int4(s.dst[k], c.src[k], p0, p1);
      int32_t d[RECORD_ENTRIES];
      for (int q = 0; q < RECORD_ENTRIES; ++q) d[q] = p0 + q < p1 ? c.dep[p0 + q] : s.n;  // unused: the pad row
      d4[k] = make_int4(d[0], d[1], d[2], d[3]);
    }
    head.upload(hd, stream);
    dep4.upload(d4, stream);
    dep.upload(c.dep, stream);
    check(hipStreamSynchronize(stream), "upload window");
    dev = WindowDev{head.p, dep4.p, dep.p, cap};
  }
};

struct DevSweep {
  DevBuf<int32_t> level_ptr, src, dst, row_ptr, dep, part_level;
  // the windowed solve's encodings of this sweep, by ring size (built on first use)
  std::map<int32_t, std::unique_ptr<DevWindow>> windows;
  const DevWindow* window(int32_t cap, hipStream_t stream, bool build) {
    auto it = windows.find(cap);
    if (it != windows.end()) return it->second.get();
    if (!build) return nullptr;
    auto w = std::make_unique<DevWindow>();
    w->build(host, cap, stream);
    return (windows[cap] = std::move(w)).get();
  }
  DevBuf<double> val64, diag64;
  DevBuf<float> val32, diag32;
  DevBuf<int4> head, dep4;
  DevBuf<double2> val01, val23;
  DevBuf<float4> v32;
  DevBuf<double> rdiag64;
  DevBuf<float> rdiag32;
  Sweep host;
  // `wide`: the sweep reads its dependencies from a row-interleaved block with a pad row after
  // its n rows; a record's unused inline slots point there (one line every workgroup keeps in
  // cache) rather than at a row of the block, so a row with fewer than four terms issues no load
  // that misses. A column sweep's unused slots point at the row's own slot in LDS.
  void upload(Sweep s, hipStream_t stream, bool wide) {
    host = std::move(s);
    const int32_t n = host.n;
    // one record per position (a split sweep has a position per segment of a row)
    const int32_t count = host.positions();
    std::vector<int4> hd(count), dp(count);
    std::vector<double2> a(count), b(count);
    std::vector<float4> c(count);
    std::vector<double> rd(count, 1.0);
    std::vector<float> rd32(count, 1.0f);
    for (int32_t k = 0; k < count; ++k) {
      if (!host.unit) {
        rd[k] = host.diag[k];
        rd32[k] = (float)host.diag[k];
      }
      const int32_t p0 = host.row_ptr[k], p1 = host.row_ptr[k + 1];
      hd[k] = make_int4(host.dst[k], host.src[k], p0, p1);
      int32_t d[RECORD_ENTRIES];
      double v[RECORD_ENTRIES];
      for (int q = 0; q < RECORD_ENTRIES; ++q) {
        // unused slots: value 0, and an index that is valid and cheap to read
        const bool inside = p0 + q < p1;
        d[q] = inside ? host.dep[p0 + q] : (wide ? n : host.dst[k]);
        v[q] = inside ? host.val[p0 + q] : 0.0;
      }
      dp[k] = make_int4(d[0], d[1], d[2], d[3]);
      a[k] = make_double2(v[0], v[1]);
      b[k] = make_double2(v[2], v[3]);
      c[k] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    }
    head.upload(hd, stream);
    dep4.upload(dp, stream);
    val01.upload(a, stream);
    val23.upload(b, stream);
    v32.upload(c, stream);
    rdiag64.upload(rd, stream);
    rdiag32.upload(rd32, stream);
    level_ptr.upload(host.level_ptr, stream);
    if (host.parts) part_level.upload(host.part_level, stream);
    src.upload(host.src, stream);
    dst.upload(host.dst, stream);
    row_ptr.upload(host.row_ptr, stream);
    dep.upload(host.dep, stream);
    val64.upload(host.val, stream);
    std::vector<float> v32v(host.val.begin(), host.val.end());
    val32.upload(v32v, stream);
    if (!host.unit) {
      diag64.upload(host.diag, stream);
      std::vector<float> d32(host.diag.begin(), host.diag.end());
      diag32.upload(d32, stream);
    }
    check(hipStreamSynchronize(stream), "upload factor");
  }
  SweepDev view() const {
    SweepDev v;
    v.n = host.n;
    v.levels = host.levels;
    v.level_ptr = level_ptr.p;
    v.src = src.p;
    v.dst = dst.p;
    v.row_ptr = row_ptr.p;
    v.dep = dep.p;
    v.val64 = val64.p;
    v.val32 = val32.p;
    v.diag64 = host.unit ? nullptr : diag64.p;
    v.diag32 = host.unit ? nullptr : diag32.p;
    v.rec.head = head.p;
    v.rec.dep = dep4.p;
    v.rec.val01 = val01.p;
    v.rec.val23 = val23.p;
    v.rec.val32 = v32.p;
    v.rec.diag64 = rdiag64.p;
    v.rec.diag32 = rdiag32.p;
    v.unit = host.unit;
    v.split = host.split;
    v.parts = host.parts;
    v.part_level = host.parts ? part_level.p : nullptr;
    return v;
  }
  // The most levels of one of the independent parts, and the top part's
  int32_t most_part_levels() const {
    int32_t m = 0;
    for (int32_t q = 0; q < host.parts; ++q) m = std::max(m, host.part_level[q + 1] - host.part_level[q]);
    return m;
  }
  int32_t top_levels() const { return host.parts ? host.part_level[host.parts + 1] - host.part_level[host.parts] : 0; }
  size_t entries() const { return host.dep.size(); }
};

// The top parts of a partitioned pair of sweeps renumbered into LDS slots (top_slots), with the
// metadata image the LDS top kernel copies (laid out by top_lds_layout)
struct DevTopSlots {
  DevBuf<int32_t> rhs_row, halo_row, out_row;
  DevBuf<int4> meta;
  TopSlotsDev dev;
  void upload(const TopSlots& h, const Sweep& f, const Sweep& b, hipStream_t stream) {
    TopSlotsDev d;
    d.rows = h.rows;
    d.halo = h.halo;
    // terms past a segment's fourth go to the overflow arrays
    auto overflow = [](const TopSlots::Part& p, const Sweep& s) {
      int32_t m = 0;
      for (int32_t k = 0; k < p.segs; ++k)
        m += std::max(0, s.row_ptr[p.seg0 + k + 1] - s.row_ptr[p.seg0 + k] - RECORD_ENTRIES);
      return m;
    };
    d.fwd = TopPartDev{h.fwd.levels, h.fwd.segs, overflow(h.fwd, f)};
    d.bwd = TopPartDev{h.bwd.levels, h.bwd.segs, overflow(h.bwd, b)};
    const TopLdsLayout o = top_lds_layout(d, 1);
    std::vector<int4> image((o.meta_bytes + 15) / 16, make_int4(0, 0, 0, 0));
    unsigned char* base = reinterpret_cast<unsigned char*>(image.data());
    auto fill = [&](const TopSlots::Part& p, const Sweep& s, const TopLdsLayout::Part& at) {
      TopRecord* rec = reinterpret_cast<TopRecord*>(base + at.rec);
      double* diag = reinterpret_cast<double*>(base + at.diag);
      int32_t* lp = reinterpret_cast<int32_t*>(base + at.lp);
      double* ov_val = reinterpret_cast<double*>(base + at.ov_val);
      int32_t* ov_dep = reinterpret_cast<int32_t*>(base + at.ov_dep);
      int32_t ov = 0;
      for (int32_t k = 0; k < p.segs; ++k) {
        const int32_t q0 = s.row_ptr[p.seg0 + k] - p.term0, q1 = s.row_ptr[p.seg0 + k + 1] - p.term0;
        int32_t dep[RECORD_ENTRIES];
        double val[RECORD_ENTRIES];
        for (int e = 0; e < RECORD_ENTRIES; ++e) {
          const bool inside = q0 + e < q1;
          dep[e] = inside ? p.dep[q0 + e] : p.slot[k];
          val[e] = inside ? s.val[p.term0 + q0 + e] : 0.0;
        }
        rec[k].head = make_int4(p.slot[k], q1 - q0, ov, 0);
        rec[k].dep = make_int4(dep[0], dep[1], dep[2], dep[3]);
        rec[k].v01 = make_double2(val[0], val[1]);
        rec[k].v23 = make_double2(val[2], val[3]);
        for (int32_t q = q0 + RECORD_ENTRIES; q < q1; ++q, ++ov) {
          ov_val[ov] = s.val[p.term0 + q];
          ov_dep[ov] = p.dep[q];
        }
        diag[k] = s.unit ? 1.0 : s.diag[p.seg0 + k];
      }
      for (int32_t l = 0; l <= p.levels; ++l) lp[l] = s.level_ptr[p.level0 + l] - p.seg0;
    };
    fill(h.fwd, f, o.fwd);
    fill(h.bwd, b, o.bwd);
    rhs_row.upload(h.rhs_row, stream);
    halo_row.upload(h.halo_row, stream);
    out_row.upload(h.out_row, stream);
    meta.upload(image, stream);
    check(hipStreamSynchronize(stream), "upload top slots");
    d.rhs_row = rhs_row.p;
    d.halo_row = halo_row.p;
    d.out_row = out_row.p;
    d.meta = meta.p;
    d.meta_units = (int32_t)image.size();
    dev = d;
  }
};

// One sweep's independent parts renumbered into LDS slots (part_slots), as the LDS parts kernel
// reads them: per part a descriptor, the segments' records and diagonals, the slots' rows, the
// halo's rows, the rounds (a wavefront's pass over up to 64 of a level's segments, padded to a
// multiple of PARTS_LOOKAHEAD with empty ones) and the terms past each segment's fourth
struct DevPartSlots {
  DevBuf<PartDesc> desc;
  DevBuf<TopRecord> rec;
  DevBuf<double> diag, ov_val;
  DevBuf<int32_t> in_row, out_row, halo_row, ov_dep;
  DevBuf<int4> rounds;
  PartSlotsDev dev;
  // false (nothing uploaded) where an index would leave its part's slots, terms or segments
  bool upload(const PartSlots::Dir& h, int32_t parts, const Sweep& s, hipStream_t stream) {
    constexpr int32_t lanes = 64;
    std::vector<PartDesc> descs(parts);
    std::vector<TopRecord> recs(h.slot.size());
    std::vector<double> diags(h.slot.size()), ovv;
    std::vector<int32_t> ovd;
    std::vector<int4> rnd;
    PartSlotsDev d;
    d.parts = parts;
    for (int32_t q = 0; q < parts; ++q) {
      PartDesc& p = descs[q];
      p = PartDesc{};
      p.seg0 = h.seg_begin[q];
      p.segs = h.seg_begin[q + 1] - h.seg_begin[q];
      p.slot0 = h.slot_begin[q];
      p.rows = h.slot_begin[q + 1] - h.slot_begin[q];
      p.halo0 = h.halo_begin[q];
      p.halo = h.halo_begin[q + 1] - h.halo_begin[q];
      p.round0 = (int32_t)rnd.size();
      p.ov0 = (int32_t)ovv.size();
      for (int32_t l = s.part_level[q]; l < s.part_level[q + 1]; ++l) {
        const int32_t lo = s.level_ptr[l] - p.seg0, width = s.level_ptr[l + 1] - s.level_ptr[l];
        for (int32_t j = 0; j < width; j += lanes)
          rnd.push_back(make_int4(lo + j, std::min(lanes, width - j), j + lanes >= width ? 1 : 0, 0));
      }
      while ((rnd.size() - (size_t)p.round0) % PARTS_LOOKAHEAD) rnd.push_back(make_int4(0, 0, 0, 0));
      p.rounds = (int32_t)rnd.size() - p.round0;
      for (int32_t k = h.seg_begin[q]; k < h.seg_begin[q + 1]; ++k) {
        const int32_t q0 = s.row_ptr[k], q1 = s.row_ptr[k + 1];
        TopRecord& r = recs[k];
        int32_t dep[RECORD_ENTRIES];
        double val[RECORD_ENTRIES];
        for (int e = 0; e < RECORD_ENTRIES; ++e) {
          const bool inside = q0 + e < q1;
          dep[e] = inside ? h.dep[q0 + e] : h.slot[k];
          val[e] = inside ? s.val[q0 + e] : 0.0;
        }
        r.head = make_int4(h.slot[k], q1 - q0, (int32_t)ovv.size() - p.ov0, 0);
        r.dep = make_int4(dep[0], dep[1], dep[2], dep[3]);
        r.v01 = make_double2(val[0], val[1]);
        r.v23 = make_double2(val[2], val[3]);
        for (int32_t e = q0 + RECORD_ENTRIES; e < q1; ++e) {
          ovv.push_back(s.val[e]);
          ovd.push_back(h.dep[e]);
        }
        diags[k] = s.unit ? 1.0 : s.diag[k];
      }
      p.ov = (int32_t)ovv.size() - p.ov0;
      // every slot a record or overflow term names is one of the part's, and every round's
      // segments are the part's
      const int32_t segs = h.seg_begin[q + 1] - h.seg_begin[q], slots = p.rows + p.halo;
      for (int32_t k = h.seg_begin[q]; k < h.seg_begin[q + 1]; ++k) {
        const TopRecord& r = recs[k];
        for (int32_t x : {r.head.x, r.dep.x, r.dep.y, r.dep.z, r.dep.w})
          if (x < 0 || x >= slots) return false;
        if (r.head.x >= p.rows || r.head.y < 0 || r.head.z < 0 ||
            r.head.z + std::max(0, r.head.y - RECORD_ENTRIES) > p.ov)
          return false;
      }
      for (int32_t e = p.ov0; e < p.ov0 + p.ov; ++e)
        if (ovd[e] < 0 || ovd[e] >= slots) return false;
      for (int32_t e = p.round0; e < p.round0 + p.rounds; ++e)
        if (rnd[e].x < 0 || rnd[e].y < 0 || rnd[e].x + rnd[e].y > segs) return false;
      if (segs < 1 || p.rows < 1) return false;
      for (int32_t e = p.slot0; e < p.slot0 + p.rows; ++e)
        if (h.in_row[e] < 0 || h.in_row[e] >= s.n || h.out_row[e] < 0 || h.out_row[e] >= s.n) return false;
      for (int32_t e = p.halo0; e < p.halo0 + p.halo; ++e)
        if (h.halo_row[e] < 0 || h.halo_row[e] >= s.n) return false;
      d.most_slots = std::max(d.most_slots, p.rows + p.halo);
      d.most_ov = std::max(d.most_ov, p.ov);
      d.most_rounds = std::max(d.most_rounds, p.rounds);
    }
    desc.upload(descs, stream);
    rec.upload(recs, stream);
    diag.upload(diags, stream);
    ov_val.upload(ovv, stream);
    ov_dep.upload(ovd, stream);
    in_row.upload(h.in_row, stream);
    out_row.upload(h.out_row, stream);
    halo_row.upload(h.halo_row, stream);
    rounds.upload(rnd, stream);
    check(hipStreamSynchronize(stream), "upload part slots");
    d.desc = desc.p;
    d.rec = rec.p;
    d.diag = diag.p;
    d.in_row = in_row.p;
    d.out_row = out_row.p;
    d.halo_row = halo_row.p;
    d.rounds = rounds.p;
    d.ov_val = ov_val.p;
    d.ov_dep = ov_dep.p;
    dev = d;
    return true;
  }
};

// A complex factor's sweep on the device: level pointers, the rows' CSR arrays, and the values
// as (re, im) pairs
struct DevSweepC {
  DevBuf<int32_t> level_ptr, src, dst, row_ptr, dep;
  DevBuf<int4> head;
  DevBuf<double2> val, diag;
  DevBuf<float2> val32, diag32;  // the values rounded to FP32, for refined solves
  int32_t n = 0, levels = 0;
  bool unit = false;
  void upload(const Sweep& s, hipStream_t stream) {
    n = s.n;
    levels = s.levels;
    unit = s.unit;
    std::vector<double2> v(s.val.size()), d(unit ? 0 : (size_t)n);
    for (size_t p = 0; p < v.size(); ++p) v[p] = make_double2(s.val[p], s.val_im[p]);
    for (size_t k = 0; k < d.size(); ++k) d[k] = make_double2(s.diag[k], s.diag_im[k]);
    std::vector<float2> v32(v.size()), d32(d.size());
    for (size_t p = 0; p < v.size(); ++p) v32[p] = make_float2((float)v[p].x, (float)v[p].y);
    for (size_t k = 0; k < d.size(); ++k) d32[k] = make_float2((float)d[k].x, (float)d[k].y);
    val32.upload(v32, stream);
    if (!unit) diag32.upload(d32, stream);
    std::vector<int4> h(n);
    for (int32_t k = 0; k < n; ++k) h[k] = make_int4(s.dst[k], s.src[k], s.row_ptr[k], s.row_ptr[k + 1]);
    head.upload(h, stream);
    level_ptr.upload(s.level_ptr, stream);
    src.upload(s.src, stream);
    dst.upload(s.dst, stream);
    row_ptr.upload(s.row_ptr, stream);
    dep.upload(s.dep, stream);
    val.upload(v, stream);
    if (!unit) diag.upload(d, stream);
  }
  ComplexSweepDev view() const {
    ComplexSweepDev c;
    c.n = n;
    c.levels = levels;
    c.level_ptr = level_ptr.p;
    c.head = head.p;
    c.src = src.p;
    c.dst = dst.p;
    c.row_ptr = row_ptr.p;
    c.dep = dep.p;
    c.val = val.p;
    c.diag = unit ? nullptr : diag.p;
    c.val32 = val32.p;
    c.diag32 = unit ? nullptr : diag32.p;
    return c;
  }
};

struct DevCsr {
  int32_t rows = 0, cols = 0;
  DevBuf<int32_t> ptr, col;
  DevBuf<double> val;
  size_t nnz = 0;
  DevBuf<int32_t> row;         // each entry's row (the adjoint's SDDMM takes entries one by one)
  std::vector<int64_t> order;  // each entry's position in the matrix as the caller stored it
  DevBuf<int32_t> order_dev;   // the same on the device
  // the matrix as the caller gave it (its orientation and pattern), and for the transpose each
  // entry's position in it: a refactor of the same matrix rewrites both values in place
  bool src_row_major = true;
  std::vector<uint64_t> src_ptr, src_idx;
  DevBuf<int32_t> t_order;
  CsrDev view() const { return CsrDev{rows, cols, ptr.p, col.p, val.p}; }
};

constexpr uint64_t MAX_DEFAULT_WIDTH = 32768;
constexpr uint64_t SINGLE_ROW_WIDTH = 64;
constexpr int64_t INDEX_LIMIT = 0x7fffffff;

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// v with `pad` value-initialized entries after its last
template <typename T>
std::vector<T> padded(const std::vector<T>& v, size_t pad) {
  std::vector<T> out(v);
  out.resize(v.size() + pad, T());
  return out;
}

// The most branches one outage set may hold (its system is solved densely on the host)
constexpr int32_t MAX_OUTAGE_SET = 64;

// Host threads for the library's own host work: the machine's, at most 16
int64_t host_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::clamp<int64_t>(hw ? hw : 1, 1, 16);
}

// Runs independent host tasks together, each on its own thread (the first on the calling one),
// and rethrows the first task's failure once all have ended
void run_together(const std::vector<std::function<void()>>& tasks) {
  std::vector<std::exception_ptr> errors(tasks.size());
  std::vector<std::thread> pool;
  for (size_t i = 1; i < tasks.size(); ++i)
    pool.emplace_back([&, i] {
      try {
        tasks[i]();
      } catch (...) {
        errors[i] = std::current_exception();
      }
    });
  try {
    if (!tasks.empty()) tasks[0]();
  } catch (...) {
    errors[0] = std::current_exception();
  }
  for (std::thread& t : pool) t.join();
  for (const std::exception_ptr& e : errors)
    if (e) std::rethrow_exception(e);
}

double seconds_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

std::vector<int32_t> to_i32(const uint64_t* v, size_t n, size_t entries) {
  std::vector<int32_t> out(n);
  for (size_t i = 0; i < n; ++i) {
    if (v[i] > (uint64_t)INDEX_LIMIT)
      fail(YBUS_HIP_TOO_LARGE, "a matrix with " + std::to_string(entries) + " entries is too large for 32-bit indices");
    out[i] = (int32_t)v[i];
  }
  return out;
}

void check_permutation(const uint64_t* perm, uint64_t n, const char* what) {
  if (!perm) fail(YBUS_HIP_INVALID_ARGUMENT, std::string(what) + " is null");
  std::vector<bool> seen(n, false);
  for (uint64_t i = 0; i < n; ++i) {
    if (perm[i] >= n || seen[perm[i]]) fail(YBUS_HIP_INVALID_ARGUMENT, std::string(what) + " is not a permutation");
    seen[perm[i]] = true;
  }
}

}  // namespace

struct ybus_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string name;
  int cus = 256;
  // this device's no-allocation workspace (SCRATCH_BYTES), from scratch_address at creation
  void* scratch = nullptr;
  ~ybus_hip_ctx() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct ybus_hip_block;

// How a block is laid out and solved. Columns: each right-hand side contiguous (the host's own
// layout, ld = n), solved one workgroup per column with the vector in LDS; for a factor whose
// solution fits in LDS and a block of few enough columns that one or two rounds of workgroups
// take them all. Wide: row-interleaved (n rows of ld columns), solved by column tiles in global
// memory; for everything else.
struct Layout {
  bool columns = true;
  int64_t ld = 0;
  int64_t rs() const { return columns ? 1 : ld; }
  int64_t cs() const { return columns ? ld : 1; }
};

// A forward sweep's dependency graph, for finding the reach of pair blocks: each row's dependents
// (rows that read it), its position, the row reading each right-hand side row, and each row's
// first dependent in processing order. When every other dependent of a row is also a dependent
// of that first one (a Cholesky factor's columns nest this way: the elimination tree), a row's
// reach is its path to the root through first dependents (is_path).
struct ReachIndex {
  std::vector<int32_t> dep_ptr, dependents, pos_of_row, row_of_src, first_dependent;
  bool is_path = false;
  void build(const Sweep& sw) {
    const int32_t rows = sw.n;
    dep_ptr.assign(rows + 1, 0);
    pos_of_row.assign(rows, 0);
    row_of_src.assign(rows, 0);
    for (int32_t k = 0; k < rows; ++k) {
      pos_of_row[sw.dst[k]] = k;
      row_of_src[sw.src[k]] = sw.dst[k];
      for (int32_t q = sw.row_ptr[k]; q < sw.row_ptr[k + 1]; ++q) dep_ptr[sw.dep[q] + 1] += 1;
    }
    for (int32_t r = 0; r < rows; ++r) dep_ptr[r + 1] += dep_ptr[r];
    dependents.assign(dep_ptr[rows], 0);
    std::vector<int32_t> cursor(dep_ptr.begin(), dep_ptr.end() - 1);
    for (int32_t k = 0; k < rows; ++k)
      for (int32_t q = sw.row_ptr[k]; q < sw.row_ptr[k + 1]; ++q) dependents[cursor[sw.dep[q]]++] = sw.dst[k];
    first_dependent.assign(rows, -1);
    for (int32_t r = 0; r < rows; ++r)
      for (int32_t e = dep_ptr[r]; e < dep_ptr[r + 1]; ++e)
        if (first_dependent[r] < 0 || pos_of_row[dependents[e]] < pos_of_row[first_dependent[r]])
          first_dependent[r] = dependents[e];
    std::vector<int32_t> mark(rows, -1);
    is_path = true;
    for (int32_t r = 0; r < rows && is_path; ++r) {
      const int32_t up = first_dependent[r];
      if (up < 0) continue;
      for (int32_t e = dep_ptr[up]; e < dep_ptr[up + 1]; ++e) mark[dependents[e]] = r;
      for (int32_t e = dep_ptr[r]; e < dep_ptr[r + 1]; ++e)
        if (dependents[e] != up && mark[dependents[e]] != r) is_path = false;
    }
  }
};

// An LU factor's pattern as uploaded (CSC, pivot-step numbering), kept for a later refactor or
// value update
struct LuPattern {
  std::vector<uint64_t> l_ptr, l_idx, u_ptr, u_idx, row_perm, col_perm;
  std::vector<double> values;  // the uploaded values: L's, then U's ((re, im) pairs when complex)
};

// One sweep's value arrays and where each value comes from in the new factor values (L's, then
// U's, in the uploaded CSC order)
struct SweepRefresh {
  DevSweep* sweep = nullptr;     // a real factor's sweep
  DevSweepC* zsweep = nullptr;   // or a complex factor's
  DevBuf<int32_t> vpos, dpos;
  SweepRefresh() = default;
  SweepRefresh(const SweepRefresh&) = delete;
};

// New values for an uploaded LU factor: the sweeps' refresh maps, and (once a matrix pattern is
// given) the device refactorization's arrays for that pattern
struct RefactorPlan {
  std::deque<SweepRefresh> refresh;
  // the factor's values (L's, then U's), and the next ones being made
  DevBuf<double> vals, next;
  bool numeric = false;
  std::vector<uint64_t> a_ptr, a_idx;  // the matrix pattern the numeric plan was built for
  DevBuf<int32_t> a_slot, target, control;
  DevBuf<int4> head, span, step;
  DevBuf<double> a_val;
  int32_t x_cap = 0;
  int grid = 0;
  int32_t stage = REFACTOR_STAGE;  // entries staged per batch
  ybus_hip_refactor_stats stats{};
};

struct ybus_hip_factor {
  ybus_hip_ctx* ctx = nullptr;
  int64_t n = 0;
  bool lu = false;
  uint64_t width = 1;
  int64_t nnz_factor = 0;  // S: entries swept by one forward and backward solve
  bool fits = false;       // one column's solution fits in LDS (the column path is built)
  int force = 0;           // 0: choose the path per block, 1: always wide, 2: columns where it fits
  int64_t column_limit = 0;  // the widest block the column path takes by default
  size_t lds_cap = 0;
  // wide path: schedules reading and writing permuted blocks
  DevSweep fwd, bwd, tfwd, tbwd;
  // an LU factor's wide sweeps with their long rows split into chained segments (split_rows),
  // which the wide solves of vectors and the FP32 refinement take (0: not built)
  int32_t split_cap = 0;
  DevSweep sfwd, sbwd, stfwd, stbwd;
  // the column path's sweeps split the same way (in place in LDS, a segment continues from its
  // row's own slot without any change to the kernel); the cap they were split with (0: none)
  int32_t column_split_cap = 0;
  // wide path, partitioned (a Cholesky factor whose elimination tree splits): the same sweeps
  // with their rows grouped into independent parts and a top part; FP64 solves take these
  bool partitioned = false;
  bool force_parts = false;  // YBUS_HIP_PARTS named a part count: untuned blocks take the parts
  DevSweep pfwd, pbwd;
  PartLevels part_levels;
  // the partitioned sweeps with long rows split within their parts (blocks of vectors)
  bool parts_split = false;
  DevSweep spfwd, spbwd;
  // the top parts of the sweeps blocks of vectors take, in LDS slots (rows 0: none);
  // YBUS_HIP_TOP_LDS=0 keeps the top kernel over the block in global memory
  DevTopSlots top_slots;
  bool use_top_lds = true;
  // the same sweeps' independent parts in LDS slots (parts_lds.parts 0: none), which blocks of
  // vectors take at the tuned tile width (Tuned::parts_lds_tile); YBUS_HIP_PARTS_LDS=tile[:threads]
  // sets it for every block of vectors, 0 keeps the parts kernels over the block in global memory
  struct {
    int32_t parts = 0;
    DevPartSlots fwd, bwd;
  } parts_lds;
  int parts_lds_env = -1, parts_lds_env_threads = 0;
  // the one-schedule wide FP64 solve with each tile's last positions in an LDS ring
  // (solve64_window_kernel), where tuned or YBUS_HIP_WINDOW=1 (0: never)
  int window_env = -1;
  // inside ybus_hip_factor_solve_device_noalloc: nothing may be allocated, so a ring encoding not
  // yet built is not built (the barrier solve runs instead)
  bool no_alloc = false;
  // column path: schedules in factor order, and each unknown's factor row for the right-hand
  // side (rhs_row) and the solution (sol_of), for the solve and the transposed solve
  DevSweep cfwd, cbwd, ctfwd, ctbwd;
  DevBuf<int32_t> rhs_row, sol_of, t_rhs_row, t_sol_of;
  bool has_transpose = false;
  // a complex factor: its sweeps (solve and transposed solve) and its blocks of vectors; none
  // of the real arrays above is built
  bool complex = false;
  DevSweepC zfwd, zbwd, ztfwd, ztbwd;
  DevBuf<double> zb, zt, zx, zvec;
  // a complex factor's refined solves: the second iterate, the residuals, the FP32 blocks, and
  // the matrix the residuals are taken against (CSR by rows, (re, im) values)
  DevBuf<double> zx1, zr;
  DevBuf<float> zt32, zd32;
  struct {
    int32_t rows = 0;
    size_t nnz = 0;
    DevBuf<int32_t> ptr, col;
    DevBuf<double2> val;
  } za;
  // a real LU factor's pattern, and the plan for new values (built by the first refactor or
  // value update)
  LuPattern lu_host;
  std::unique_ptr<RefactorPlan> refactor;
  DevCsr a;   // the matrix itself, by user rows, for the FP64 residual
  DevCsr at;  // its transpose, by rows (A's columns), for the residual of a transposed refinement
  // workspace for `cols` columns
  int64_t cols = 0;
  Layout layout;
  DevBuf<double> b, t, x, x1, r, y2, vec;
  DevBuf<float> t32, d32;
  DevBuf<int32_t> plus, minus, isteps, singular;
  DevBuf<double> dratio;
  DevBuf<double> abar;  // the adjoint's running sums, one per stored entry of the matrix
  // pair updates' coefficients, and the maxima's partial and final results (per column)
  DevBuf<double> coef, max_pv, max_best, ones, upd_base, upd_alpha, upd_weight;
  DevBuf<int32_t> max_pr, max_pb, max_row, max_bad;
  // the staged pair block's reach, per tile of `reach_tile` columns (0: none staged), and the
  // host index it is computed from (built on first use)
  bool use_reach = true;
  int reach_tile = 0;
  DevBuf<int32_t> reach_pos, reach_lp;
  DevBuf<uint32_t> reach_live;  // partitioned sweeps: per tile, a bit per forward row in its reach
  bool reach_parts = false;  // the staged reach is over the partitioned sweep's positions
  ReachIndex reach_index[2];  // for the wide forward sweep, and for its partitioned form
  int64_t staged = 0;
  bool staged_pairs = false;  // the staged block is pairs (plus, minus), not vectors (b)
  ybus_hip_block* borrowed = nullptr;
  Staging staging;
  ybus_hip_timings timings{};
  ybus_hip_kernel_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int tile_override = 0, threads_wide = 512, threads_columns = 512;
  // the refined solves' stop rule (YBUS_HIP_STOP_NORMWISE or YBUS_HIP_STOP_AT_FLOOR)
  int32_t refine_stop = YBUS_HIP_STOP_NORMWISE;
  bool env_threads = false;  // YBUS_HIP_THREADS set: it overrides the tuned workgroup size
  // YBUS_HIP_TOP=tile[:threads]: the partitioned top kernel's configuration for blocks of vectors
  int top_override = 0, top_override_threads = 0;
  // kernel configurations chosen by ybus_hip_factor_autotune, per block width (keyed by the
  // power of two at or above it)
  struct Tuned {
    bool columns = false;
    int threads = 512, tile = 0;
    int parts = -1;  // the partitioned sweeps: 1 on, 0 off, -1 by the default rule
    int top_tile = 0, top_threads = 0;  // the partitioned top kernel's own (0: the parts')
    int parts_lds_tile = 0, parts_lds_threads = 0;  // the parts in LDS slots (0: over the block)
    bool window = false;  // the one-schedule wide solve with the last positions in LDS
  };
  std::map<int, Tuned> tuned;
  const Tuned* trial = nullptr;  // a configuration being timed by the autotuner

  static int bucket(int64_t ncols) {
    int b = 0;
    while ((int64_t(1) << b) < ncols) ++b;
    return b;
  }
  // pairs: a pair block (whose forward sweep takes each tile's reach), else a block of vectors
  const Tuned* tuned_for(int64_t ncols, bool pairs) const {
    if (trial) return trial;
    auto it = tuned.find(2 * bucket(ncols) + (pairs ? 1 : 0));
    return it == tuned.end() ? nullptr : &it->second;
  }
  int wide_threads(int64_t ncols, bool pairs) const {
    const Tuned* t = tuned_for(ncols, pairs);
    return t && !t->columns && !env_threads ? t->threads : threads_wide;
  }
  int column_threads(int64_t ncols, bool pairs) const {
    const Tuned* t = tuned_for(ncols, pairs);
    return t && t->columns && !env_threads ? t->threads : threads_columns;
  }
  // Whether a one-schedule wide block of `ncols` columns takes the windowed solve
  bool window_for(int64_t ncols, bool pairs) const {
    if (window_env >= 0) return window_env == 1;
    const Tuned* t = tuned_for(ncols, pairs);
    return t && t->window;
  }
  // The LDS parts' configuration a partitioned block of vectors takes (tile 0: the parts over the
  // block): YBUS_HIP_PARTS_LDS's, else the tuned one, where both sweeps' parts fit LDS
  PartsLds lds_parts_for(int64_t ncols, bool pairs) const {
    PartsLds pl;
    if (pairs || parts_lds.parts == 0) return pl;
    int tile = 0, threads = 0;
    if (parts_lds_env >= 0) {
      tile = parts_lds_env;
      threads = parts_lds_env_threads ? parts_lds_env_threads : std::max(256, 64 * tile);
    } else if (const Tuned* c = tuned_for(ncols, pairs)) {
      tile = c->parts_lds_tile;
      threads = c->parts_lds_threads;
    }
    if (tile <= 0 || parts_lds_layout(parts_lds.fwd.dev, tile).total > lds_cap ||
        parts_lds_layout(parts_lds.bwd.dev, tile).total > lds_cap)
      return pl;
    pl.fwd = &parts_lds.fwd.dev;
    pl.bwd = &parts_lds.bwd.dev;
    pl.tile = tile;
    pl.threads = threads;
    return pl;
  }

  ~ybus_hip_factor() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }

  hipStream_t stream() const { return ctx->stream; }

  // Sizes the workspace for `ncols` columns; whatever was staged is dropped, since the solves
  // that call this overwrite the staged buffers
  void ensure_workspace(int64_t ncols, bool refine, bool pairs = false) {
    staged = 0;
    staged_pairs = false;
    reach_tile = 0;
    const int64_t nn = std::max<int64_t>(n, 1);
    if (ncols > cols) {
      cols = ncols;
      for (DevBuf<double>* d : {&b, &t, &x, &x1, &r, &y2, &vec}) d->release();
      t32.release();
      d32.release();
    }
    layout.columns = use_columns(ncols, pairs);
    const bool columns = layout.columns;
    layout.ld = columns ? nn : round_up(std::max<int64_t>(cols, 1), 16);
    const size_t elems = (size_t)nn * (size_t)(columns ? std::max<int64_t>(cols, 1) : layout.ld);
    // the blocks the wide sweeps read dependencies from carry a pad row (see DevSweep::upload)
    const size_t padded = columns ? elems : elems + (size_t)layout.ld;
    b.reserve(elems);
    x.reserve(padded);
    y2.reserve(elems);
    const size_t per = (size_t)std::max<int64_t>(cols, 16);
    plus.reserve(per);
    minus.reserve(per);
    singular.reserve(per);
    isteps.reserve(per);
    dratio.reserve(per);
    if (!columns) {
      t.reserve(padded);
      vec.reserve(elems);
    }
    if (refine) {
      x1.reserve(elems);
      r.reserve(elems);
      if (!columns) {
        t32.reserve(padded);
        d32.reserve(padded);
      }
    }
  }

  // Whether a block of `ncols` columns takes the column path
  bool use_columns(int64_t ncols, bool pairs = false) const {
    if (!fits || force == 1) return false;
    if (force == 2) return true;
    if (const Tuned* t = tuned_for(ncols, pairs)) return t->columns;
    return ncols <= column_limit;
  }

  // Whether a wide block of `ncols` columns sweeps the partitioned form: as tuned, else when its
  // tiles alone would leave compute units idle
  bool use_parts(int64_t ncols, bool pairs) const {
    if (!partitioned) return false;
    if (const Tuned* t = tuned_for(ncols, pairs))
      if (!t->columns && t->parts >= 0) return t->parts == 1;
    if (force_parts) return true;
    const int tile = tile_for(ncols, pairs);
    return (ncols + tile - 1) / tile < (int64_t)ctx->cus;
  }

  int tile_for(int64_t ncols, bool pairs) const {
    if (tile_override) return tile_override;
    if (const Tuned* t = tuned_for(ncols, pairs))
      if (!t->columns && t->tile) return t->tile;
    return tile_width_for(ncols, ctx->cus);
  }

  // The reach of each tile of a pair block (plus/minus: the block's unknowns, -1 for none) in the
  // wide forward sweep: the rows that depend, directly or through others, on a row whose
  // right-hand side is one of the tile's unknowns, as positions in level order, with each level's
  // start per tile (a breadth-first walk over the sweep's dependents from the tile's rows).
  void stage_reach(const std::vector<int32_t>& p, const std::vector<int32_t>& m, int tile, bool parts) {
    const Sweep& sw = (parts ? pfwd : fwd).host;
    const int32_t rows = sw.n, levels = sw.levels;
    ReachIndex& ix = reach_index[parts ? 1 : 0];
    if (ix.dep_ptr.empty()) ix.build(sw);
    const std::vector<int32_t>&dep_ptr = ix.dep_ptr, &dependents = ix.dependents, &pos_of_row = ix.pos_of_row,
                         &row_of_src = ix.row_of_src, &first_dependent = ix.first_dependent;
    const bool reach_is_path = ix.is_path;
    const int64_t k = (int64_t)p.size();
    const int64_t tiles = (k + tile - 1) / tile;
    // each tile's reach, sorted by position, found by workers that each take a run of tiles
    std::vector<std::vector<int32_t>> reach(tiles);
    auto work = [&](int64_t t0, int64_t t1) {
      std::vector<int32_t> stamp(rows, -1), queue;
      queue.reserve(rows);
      for (int64_t t = t0; t < t1; ++t) {
        queue.clear();
        auto visit = [&](int32_t r) {
          if (stamp[r] != (int32_t)t) {
            stamp[r] = (int32_t)t;
            queue.push_back(r);
          }
        };
        for (int64_t c = t * tile; c < std::min<int64_t>(k, (t + 1) * tile); ++c) {
          if (p[c] >= 0) visit(row_of_src[p[c]]);
          if (m[c] >= 0) visit(row_of_src[m[c]]);
        }
        if (reach_is_path) {
          // walk each start up its path until it meets rows already taken
          const size_t starts = queue.size();
          for (size_t qi = 0; qi < starts; ++qi)
            for (int32_t r = first_dependent[queue[qi]]; r >= 0 && stamp[r] != (int32_t)t; r = first_dependent[r])
              visit(r);
        } else {
          for (size_t qi = 0; qi < queue.size(); ++qi)
            for (int32_t e = dep_ptr[queue[qi]]; e < dep_ptr[queue[qi] + 1]; ++e) visit(dependents[e]);
        }
        std::vector<int32_t>& out = reach[t];
        out.resize(queue.size());
        for (size_t qi = 0; qi < queue.size(); ++qi) out[qi] = pos_of_row[queue[qi]];
        std::sort(out.begin(), out.end());
      }
    };
    const int64_t workers = std::clamp<int64_t>(tiles / 8, 1, host_threads());
    if (workers == 1) {
      work(0, tiles);
    } else {
      std::vector<std::thread> pool;
      std::vector<std::exception_ptr> errors(workers);
      for (int64_t w = 0; w < workers; ++w)
        pool.emplace_back([&, w] {
          try {
            work(tiles * w / workers, tiles * (w + 1) / workers);
          } catch (...) {
            errors[w] = std::current_exception();
          }
        });
      for (std::thread& t : pool) t.join();
      for (std::exception_ptr& e : errors)
        if (e) std::rethrow_exception(e);
    }
    size_t total = 0;
    for (const std::vector<int32_t>& r : reach) total += r.size();
    if (total > (size_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "the pair block's reach is too large");
    std::vector<int32_t> pos, lp((size_t)tiles * (levels + 1));
    pos.reserve(total);
    for (int64_t t = 0; t < tiles; ++t) {
      const size_t base = pos.size();
      pos.insert(pos.end(), reach[t].begin(), reach[t].end());
      for (int32_t l = 0; l <= levels; ++l)
        lp[(size_t)t * (levels + 1) + l] =
            (int32_t)(std::lower_bound(pos.begin() + base, pos.end(), sw.level_ptr[l]) - pos.begin());
    }
    if (pos.empty()) pos.push_back(0);
    reach_pos.upload(pos, stream());
    reach_lp.upload(lp, stream());
    if (parts) {
      // each tile's reach as a bitmap over forward rows (and the pad row, always clear)
      const int64_t words = rows / 32 + 1;
      std::vector<uint32_t> bits((size_t)tiles * words, 0u);
      for (int64_t t = 0; t < tiles; ++t)
        for (int32_t q : reach[t]) {
          const int32_t r = sw.dst[q];
          bits[(size_t)t * words + (r >> 5)] |= 1u << (r & 31);
        }
      reach_live.upload(bits, stream());
    }
    check(hipStreamSynchronize(stream()), "stage reach");
    reach_tile = tile;
    reach_parts = parts;
  }

  void begin_kernel() { check(hipEventRecord(ev0, stream()), "hipEventRecord"); }
  void end_kernel(uint64_t bytes) {
    check(hipEventRecord(ev1, stream()), "hipEventRecord");
    check(hipEventSynchronize(ev1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    stats.launches += 1;
    stats.milliseconds += ms;
    stats.bytes += bytes;
  }

  // Algorithmic bytes of one FP64 block solve of `ncols` columns: the factor's values and
  // indices swept once, its row pointers, and each column's right-hand side in and solution out
  // (for pairs, two indices in and the solution out)
  uint64_t bytes64(int64_t ncols, bool pairs) const {
    const uint64_t factor = (uint64_t)nnz_factor * 12 + 8 * (uint64_t)(n + 1);
    const uint64_t columns_bytes =
        pairs ? (uint64_t)ncols * ((uint64_t)n * 8 + 8) : (uint64_t)ncols * (uint64_t)n * 16;
    return factor + columns_bytes;
  }

  // A complex block of up to `width` columns (2n rows of ld doubles): b, t, x and the vectors'
  // staging, and for refinement the second iterate, the residuals, the FP32 blocks and the
  // per-column results
  void reserve_complex(int64_t width, int64_t ld, bool refine) {
    const size_t elems = (size_t)(2 * n) * (size_t)ld;
    zb.reserve(elems);
    zt.reserve(elems);
    zx.reserve(elems);
    zvec.reserve((size_t)(2 * n) * (size_t)width);
    if (refine) {
      zx1.reserve(elems);
      zr.reserve(elems);
      zt32.reserve(elems);
      zd32.reserve(elems);
      isteps.reserve((size_t)std::max<int64_t>(width, 16));
      dratio.reserve((size_t)std::max<int64_t>(width, 16));
    }
  }

  // Solves the complex block in zb (`chunk` columns, ld doubles a row) into zx: FP64, transposed
  // (an LU factor's U^T, L^T sweeps; for a Cholesky factor conj(A^-1 conj(b))), or FP32 sweeps
  // refined in FP64 against the attached complex matrix, each column's corrections and ratio into
  // isteps / dratio
  void zsolve(int mode, int64_t chunk, int64_t ld, double tol, int32_t max_steps) {
    const int tile = tile_override ? tile_override : tile_width_for(chunk, ctx->cus);
    const uint64_t nn = (uint64_t)n;
    begin_kernel();
    if (mode == YBUS_HIP_FP32_REFINED) {
      int th = 64;
      while (th * 2 <= threads_wide) th *= 2;
      check(launch_refine_complex(stream(), tile, th, zfwd.view(), zbwd.view(),
                                  ZCsrDev{za.rows, za.ptr.p, za.col.p, za.val.p}, zb.p, zt32.p, zd32.p, zx.p, zx1.p,
                                  zr.p, ld, (int32_t)chunk, tol, max_steps, isteps.p, dratio.p, refine_stop),
            "complex refine kernel");
      // the FP32 factor's values (8 bytes) and indices once per solve, the residual's matrix once
      // per residual (16-byte values, 4-byte indices), and each column's vectors; counted for one
      // solve and residual plus three corrections, the kernel's usual work
      const uint64_t solve = (uint64_t)nnz_factor * 12 + 8 * (nn + 1), resid = (uint64_t)za.nnz * 20 + 4 * (nn + 1);
      end_kernel(4 * (solve + resid) + 32 * nn * (uint64_t)chunk * 4);
      return;
    }
    const bool transpose = mode == YBUS_HIP_TRANSPOSE;
    const bool lu_t = transpose && lu;
    check(launch_solve_complex(stream(), tile, threads_wide, (lu_t ? ztfwd : zfwd).view(), (lu_t ? ztbwd : zbwd).view(),
                               zb.p, zt.p, zx.p, ld, (int32_t)chunk, transpose && !lu),
          "complex solve kernel");
    // the factor's (re, im) values and 4-byte indices once, row pointers, and a 16-byte value
    // in and out per unknown and column
    end_kernel((uint64_t)nnz_factor * 20 + 8 * (nn + 1) + 32 * nn * (uint64_t)chunk);
  }

  // Whether the wide solves take the split sweeps (an LU factor with long rows)
  bool split_sweeps() const { return lu && !complex && split_cap > 0; }

  // FP64 solve (or transposed solve) of the block: from b, or from pairs, into x
  void solve64(bool transpose, bool pairs, int64_t ncols) {
    if (transpose && !has_transpose) transpose = false;  // A^T = A for a Cholesky factor
    begin_kernel();
    if (layout.columns) {
      const DevSweep& f = transpose ? ctfwd : cfwd;
      const DevSweep& bw = transpose ? ctbwd : cbwd;
      const DevBuf<int32_t>& rr = transpose ? t_rhs_row : rhs_row;
      const DevBuf<int32_t>& so = transpose ? t_sol_of : sol_of;
      ColumnsDev rhs{pairs ? nullptr : b.p, layout.ld, pairs ? plus.p : nullptr, pairs ? minus.p : nullptr};
      check(launch_solve_lds(stream(), column_threads(ncols, pairs), false, f.view(), bw.view(), (int32_t)n, rr.p, so.p, rhs, x.p,
                             layout.ld, (int32_t)ncols),
            "solve kernel");
    } else {
      const DevSweep& f = transpose ? tfwd : fwd;
      const DevSweep& bw = transpose ? tbwd : bwd;
      RhsDev rhs{pairs ? nullptr : b.p, pairs ? plus.p : nullptr, pairs ? minus.p : nullptr};
      const int tile = tile_for(ncols, pairs);
      if (use_parts(ncols, pairs)) {
        // a Cholesky factor (transpose is the same solve): its parts side by side
        ReachDev reach{};
        if (pairs && reach_tile == tile && reach_parts)
          reach = ReachDev{reach_pos.p, reach_lp.p, reach_live.p, (int32_t)(n / 32 + 1)};
        const bool sp = parts_split && !pairs;
        const int threads = wide_threads(ncols, pairs);
        int top_tile = tile, top_threads = threads;
        if (!pairs && top_override) {
          top_tile = top_override;
          if (top_override_threads) top_threads = top_override_threads;
        } else if (const Tuned* c = tuned_for(ncols, pairs)) {
          if (!pairs && c->top_tile && !tile_override && !env_threads) {
            top_tile = c->top_tile;
            top_threads = c->top_threads;
          }
        }
        // blocks of vectors: the top in LDS where its slots fit a workgroup's LDS at the top's tile
        const bool lds_top = !pairs && use_top_lds && top_slots.dev.rows > 0 &&
                             top_lds_layout(top_slots.dev, top_tile).total <= lds_cap;
        check(launch_solve64_parts(stream(), tile, threads, top_tile, top_threads, (sp ? spfwd : pfwd).view(),
                                   (sp ? spbwd : pbwd).view(), part_levels, rhs, reach,
                                   lds_top ? &top_slots.dev : nullptr, lds_parts_for(ncols, pairs), t.p, x.p,
                                   layout.ld, (int32_t)ncols),
              "solve kernel");
      } else if (pairs && !transpose && reach_tile == tile && !reach_parts) {
        check(launch_solve64_reach(stream(), tile, wide_threads(ncols, pairs), f.view(), bw.view(), rhs,
                                   ReachDev{reach_pos.p, reach_lp.p}, t.p, x.p, layout.ld, (int32_t)ncols, !lu),
              "solve kernel");
      } else {
        // an LU factor's split sweeps, where built
        const bool sp = split_sweeps();
        DevSweep& fs = const_cast<DevSweep&>(sp ? (transpose ? stfwd : sfwd) : f);
        DevSweep& bs = const_cast<DevSweep&>(sp ? (transpose ? stbwd : sbwd) : bw);
        const DevWindow* wf = nullptr;
        const DevWindow* wb = nullptr;
        if (window_for(ncols, pairs) && wide_threads(ncols, pairs) <= 512) {
          const int32_t cap = window_cap_for(fs.view(), bs.view(), tile, lds_cap);
          if (cap >= 1) {
            wf = fs.window(cap, stream(), !no_alloc);
            wb = bs.window(cap, stream(), !no_alloc);
          }
        }
        if (wf && wb)
          check(launch_solve64_window(stream(), tile, wide_threads(ncols, pairs), fs.view(), wf->dev, bs.view(),
                                      wb->dev, rhs, t.p, x.p, layout.ld, (int32_t)ncols),
                "solve kernel");
        else
          check(launch_solve64(stream(), tile, wide_threads(ncols, pairs), fs.view(), bs.view(), rhs, t.p, x.p,
                               layout.ld, (int32_t)ncols),
                "solve kernel");
      }
    }
    end_kernel(bytes64(ncols, pairs));
  }

  // FP32 sweeps refined in FP64 (transpose: A^T x = b, with A^T's sweeps and residual)
  void refine(int64_t ncols, double tol, int32_t max_steps, bool transpose = false) {
    if (!a.ptr.p) fail(YBUS_HIP_INVALID_ARGUMENT, "FP32 refinement needs the matrix (ybus_hip_factor_set_matrix)");
    const bool t_sweeps = transpose && has_transpose;  // A^T = A for a Cholesky factor
    const CsrDev m = transpose ? at.view() : a.view();
    begin_kernel();
    if (layout.columns) {
      int th = 1;
      while (th * 2 <= column_threads(ncols, false)) th *= 2;
      check(launch_refine_lds(stream(), th, (t_sweeps ? ctfwd : cfwd).view(), (t_sweeps ? ctbwd : cbwd).view(), m,
                              (t_sweeps ? t_rhs_row : rhs_row).p, (t_sweeps ? t_sol_of : sol_of).p, b.p, x.p, x1.p,
                              r.p, layout.ld, (int32_t)ncols, tol, max_steps, isteps.p, dratio.p, refine_stop),
            "refine kernel");
    } else {
      // the refinement sweeps one schedule: a configuration tuned for the partitioned solve does
      // not carry over, and the default rule picks its tile and workgroup size instead
      const Tuned* tuned_cfg = tuned_for(ncols, false);
      const bool own = tuned_cfg && !tuned_cfg->columns && tuned_cfg->parts == 1;
      const int tile = own ? (tile_override ? tile_override : tile_width_for(ncols, ctx->cus)) : tile_for(ncols, false);
      int th = 64;
      while (th * 2 <= (own ? threads_wide : wide_threads(ncols, false))) th *= 2;
      const bool sp = split_sweeps();
      const DevSweep& fs = sp ? (t_sweeps ? stfwd : sfwd) : (t_sweeps ? tfwd : fwd);
      const DevSweep& bs = sp ? (t_sweeps ? stbwd : sbwd) : (t_sweeps ? tbwd : bwd);
      check(launch_refine(stream(), tile, th, fs.view(), bs.view(), m, b.p, t32.p, d32.p, x.p, x1.p, r.p, layout.ld,
                          (int32_t)ncols, tol, max_steps, isteps.p, dratio.p, refine_stop),
            "refine kernel");
    }
    // bytes: per column, its own corrections k_c; the factor and the matrix once per sweep and
    // residual of the longest column
    std::vector<int32_t> steps(ncols);
    check(hipEventRecord(ev1, stream()), "hipEventRecord");
    check(hipMemcpyAsync(steps.data(), isteps.p, ncols * sizeof(int32_t), hipMemcpyDeviceToHost, stream()), "steps");
    check(hipStreamSynchronize(stream()), "refine");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    uint64_t kmax = 0, ksum = 0;
    for (int32_t s : steps) {
      kmax = std::max<uint64_t>(kmax, (uint64_t)s);
      ksum += (uint64_t)s;
    }
    const uint64_t nn = (uint64_t)n;
    const uint64_t sweep = (uint64_t)nnz_factor * 8 + 8 * (nn + 1);
    const uint64_t resid = (uint64_t)a.nnz * 12 + 4 * (nn + 1);
    stats.launches += 1;
    stats.milliseconds += ms;
    stats.bytes += (kmax + 1) * sweep + (kmax + 1) * resid + (uint64_t)ncols * nn * 16 + ksum * nn * 24;
  }
};

struct ybus_hip_block {
  ybus_hip_factor* f = nullptr;
  int64_t columns = 0;
  const double* data = nullptr;
  Layout layout;
  std::vector<uint8_t> singular;
  std::vector<uint64_t> steps;
  std::vector<double> ratio;
};

struct ybus_hip_operator {
  ybus_hip_factor* f = nullptr;
  DevCsr m;
  DevBuf<double> offset, weight, out, by_column;
  DevBuf<int32_t> zero;
  int64_t out_cols = 0;
};

static std::unique_ptr<ybus_hip_block> new_block(ybus_hip_factor* f, int64_t columns, const double* data) {
  auto block = std::make_unique<ybus_hip_block>();
  block->f = f;
  block->columns = columns;
  block->data = data;
  block->layout = f->layout;
  block->singular.assign(columns, 0);
  block->steps.assign(columns, 0);
  block->ratio.assign(columns, 0.0);
  return block;
}

// ---- C ABI ----

extern "C" {

size_t ybus_hip_last_error_message(char* buf, size_t len) {
  if (buf && len) {
    size_t n = std::min(len - 1, last_error.size());
    std::memcpy(buf, last_error.data(), n);
    buf[n] = 0;
  }
  return last_error.size() + 1;
}

int32_t ybus_hip_ctx_create(ybus_hip_ctx** out) {
  return guard([&] {
    if (!out) fail(YBUS_HIP_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) fail(YBUS_HIP_NO_DEVICE, "no HIP device");
    auto ctx = std::make_unique<ybus_hip_ctx>();
    ctx->device = 0;
    check(hipSetDevice(0), "hipSetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
    ctx->name = prop.name;
    // the marketing name comes from a table the driver may lack; the ISA name is always there
    if (ctx->name.empty()) ctx->name = prop.gcnArchName;
    ctx->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(scratch_address(ctx->device, &ctx->scratch), "hipGetSymbolAddress (no-allocation workspace)");
    *out = ctx.release();
  });
}

void ybus_hip_ctx_free(ybus_hip_ctx* ctx) { delete ctx; }

int32_t ybus_hip_ctx_name(ybus_hip_ctx* ctx, char* buf, size_t len, size_t* needed) {
  return guard([&] {
    if (!ctx) fail(YBUS_HIP_INVALID_ARGUMENT, "ctx is null");
    if (buf && len) {
      size_t n = std::min(len - 1, ctx->name.size());
      std::memcpy(buf, ctx->name.data(), n);
      buf[n] = 0;
    }
    if (needed) *needed = ctx->name.size() + 1;
  });
}

int32_t ybus_hip_ctx_memory(ybus_hip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  return guard([&] {
    if (!ctx) fail(YBUS_HIP_INVALID_ARGUMENT, "ctx is null");
    size_t f = 0, t = 0;
    check(hipSetDevice(ctx->device), "hipSetDevice");
    check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
  });
}

// The largest dynamic LDS a workgroup may use on the context's device
static size_t lds_capacity(const ybus_hip_ctx* ctx) {
  int bytes = 0;
  if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess || bytes <= 0)
    bytes = 64 * 1024;
  return (size_t)bytes;
}

// Polls of a column flag before the device refactor gives a wait up: far beyond any wait of a
// working refactor (each poll sleeps); YBUS_HIP_REFACTOR_SPIN sets it
static uint32_t refactor_spin_limit() {
  const char* v = std::getenv("YBUS_HIP_REFACTOR_SPIN");
  if (v) return (uint32_t)std::max(1L, std::atol(v));
  return 1u << 22;
}

// Chooses the path (the solution of one column in LDS, or wide tiles) and sets the defaults
static void begin_factor(ybus_hip_factor* f, ybus_hip_ctx* ctx, int64_t n) {
  f->ctx = ctx;
  f->n = n;
  const char* thr = std::getenv("YBUS_HIP_THREADS");
  if (thr) {
    f->threads_wide = f->threads_columns = std::clamp(std::atoi(thr), 64, 1024) / 64 * 64;
    f->env_threads = true;
  }
  const char* tile = std::getenv("YBUS_HIP_TILE");
  if (tile) f->tile_override = std::atoi(tile);
  if (const char* lds = std::getenv("YBUS_HIP_TOP_LDS")) f->use_top_lds = std::string(lds) != "0";
  if (const char* pl = std::getenv("YBUS_HIP_PARTS_LDS")) {
    f->parts_lds_env = std::max(0, std::atoi(pl));
    if (const char* colon = std::strchr(pl, ':')) f->parts_lds_env_threads = std::atoi(colon + 1);
  }
  if (const char* win = std::getenv("YBUS_HIP_WINDOW")) f->window_env = std::string(win) != "0" ? 1 : 0;
  if (const char* top = std::getenv("YBUS_HIP_TOP")) {
    f->top_override = std::atoi(top);
    if (const char* colon = std::strchr(top, ':')) f->top_override_threads = std::atoi(colon + 1);
  }
  const char* reach = std::getenv("YBUS_HIP_REACH");
  if (reach && std::string(reach) == "0") f->use_reach = false;
  const char* path = std::getenv("YBUS_HIP_PATH");
  if (path) f->force = std::string(path) == "wide" ? 1 : std::string(path) == "columns" ? 2 : 0;
  // an LU factor's rows longer than this many terms are split into segments for the wide solves:
  // segments of the four inline terms, which wait for one round of loads each (measured on the
  // 25,000-bus Jacobian, FP64: 4.57 ms unsplit, 2.98 ms in segments of 4, 3.27 of 12, 3.58 of 20)
  f->split_cap = RECORD_ENTRIES;
  const char* split = std::getenv("YBUS_HIP_SPLIT");
  if (split) f->split_cap = std::max(0, std::atoi(split));
  f->column_limit = 2 * (int64_t)ctx->cus;
  f->lds_cap = lds_capacity(ctx);
}

// Whether one column's solution fits in LDS, for FP64 solves and for FP32 refinement; the column
// path is built only then (YBUS_HIP_PATH=wide skips it)
static bool fits_columns(const ybus_hip_factor* f, int32_t levels) {
  if (f->force == 1) return false;
  const int32_t n = (int32_t)std::max<int64_t>(f->n, 1);
  int th = 1;
  while (th * 2 <= f->threads_columns) th *= 2;
  return solve_lds_bytes(n, false, levels) <= f->lds_cap && refine_lds_bytes(n, th, levels) <= f->lds_cap;
}

static void finish_factor(ybus_hip_factor* f) {
  check(hipEventCreate(&f->ev0), "hipEventCreate");
  check(hipEventCreate(&f->ev1), "hipEventCreate");
  // factor.rs:158-160: two blocks of n values per column in a quarter of the free memory;
  // here clamped at 32768 columns, enough workgroups to fill every compute unit many times over
  size_t free_b = 0, total_b = 0;
  check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  const uint64_t value = f->complex ? 16 : 8;
  uint64_t w = free_b / 4 / (2 * value * (uint64_t)std::max<int64_t>(f->n, 1));
  f->width = std::clamp<uint64_t>(w, 1, MAX_DEFAULT_WIDTH);
  ybus_hip_factor_set_block_width(f, f->width);
}

// Each unknown's factor row: the inverse of a gather (or scatter) permutation
static std::vector<int32_t> inverse_i32(const std::vector<int64_t>& perm) {
  std::vector<int32_t> inv(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = (int32_t)i;
  return inv;
}

// Checks a Cholesky factor's arrays (L in CSR, each row ending with its diagonal, and the
// permutation) and returns its number of entries
static uint64_t check_cholesky(uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val,
                               const uint64_t* perm) {
  if (!l_ptr || (!l_idx && l_ptr[n]) || (!l_val && l_ptr[n])) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  check_permutation(perm, n, "perm");
  std::string why;
  if (!valid_compressed((int64_t)n, (int64_t)n, l_ptr, l_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "L: " + why);
  const uint64_t entries = l_ptr[n];
  if (entries > (uint64_t)INDEX_LIMIT || n > (uint64_t)INDEX_LIMIT)
    fail(YBUS_HIP_TOO_LARGE, "a factor with " + std::to_string(entries) + " entries is too large for 32-bit indices");
  for (uint64_t i = 0; i < n; ++i) {
    if (l_ptr[i + 1] == l_ptr[i] || l_idx[l_ptr[i + 1] - 1] != i)
      fail(YBUS_HIP_INVALID_ARGUMENT, "each row of L must end with its diagonal");
  }
  return entries;
}

// Checks an LU factor's arrays (L and U in CSC, L's unit diagonal first and U's diagonal last in
// each column, and both permutations) and returns its number of entries
static uint64_t check_lu(uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const double* l_val,
                         const uint64_t* u_ptr, const uint64_t* u_idx, const double* u_val, const uint64_t* row_perm,
                         const uint64_t* col_perm) {
  if (!l_ptr || !u_ptr) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  if ((l_ptr[n] && (!l_idx || !l_val)) || (u_ptr[n] && (!u_idx || !u_val)))
    fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  check_permutation(row_perm, n, "row_perm");
  check_permutation(col_perm, n, "col_perm");
  std::string why;
  if (!valid_compressed((int64_t)n, (int64_t)n, l_ptr, l_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "L: " + why);
  if (!valid_compressed((int64_t)n, (int64_t)n, u_ptr, u_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "U: " + why);
  const uint64_t entries = l_ptr[n] + u_ptr[n];
  if (entries > (uint64_t)INDEX_LIMIT || n > (uint64_t)INDEX_LIMIT)
    fail(YBUS_HIP_TOO_LARGE, "a factor with " + std::to_string(entries) + " entries is too large for 32-bit indices");
  for (uint64_t j = 0; j < n; ++j) {
    if (l_ptr[j + 1] == l_ptr[j] || l_idx[l_ptr[j]] != j)
      fail(YBUS_HIP_INVALID_ARGUMENT, "each column of L must start with its unit diagonal");
    if (u_ptr[j + 1] == u_ptr[j] || u_idx[u_ptr[j + 1] - 1] != j)
      fail(YBUS_HIP_INVALID_ARGUMENT, "each column of U must end with its diagonal");
  }
  return entries;
}

// The real parts and the imaginary parts of `count` interleaved (re, im) values
static void split_complex(const double* v, uint64_t count, std::vector<double>& re, std::vector<double>& im) {
  re.resize(count);
  im.resize(count);
  for (uint64_t p = 0; p < count; ++p) {
    re[p] = v[2 * p];
    im[p] = v[2 * p + 1];
  }
}

// Entry points of real factors reject a complex one
static void require_real(const ybus_hip_factor* f) {
  if (f->complex) fail(YBUS_HIP_INVALID_ARGUMENT, "a complex factor takes the complex solves only");
}

int32_t ybus_hip_factor_upload_cholesky(ybus_hip_ctx* ctx, uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx,
                                        const double* l_val, const uint64_t* perm, ybus_hip_factor** out) {
  return guard([&] {
    if (!ctx || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t entries = check_cholesky(n, l_ptr, l_idx, l_val, perm);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto f = std::make_unique<ybus_hip_factor>();
    begin_factor(f.get(), ctx, (int64_t)n);
    f->lu = false;
    f->nnz_factor = 2 * (int64_t)entries;
    std::vector<int64_t> gather(n), ident(n);
    for (uint64_t i = 0; i < n; ++i) {
      gather[i] = (int64_t)perm[i];
      ident[i] = (int64_t)i;
    }
    // the elimination tree's independent subtrees, for FP64 solves on the wide path: a part per
    // 2,048 rows, at most 32 (YBUS_HIP_PARTS sets the target; 0 turns it off)
    const char* parts_env = std::getenv("YBUS_HIP_PARTS");
    const int32_t target =
        parts_env ? std::max(0, std::atoi(parts_env)) : (int32_t)std::clamp<int64_t>((int64_t)n / 2048, 0, 32);
    f->force_parts = parts_env && target >= 2;
    // the host analysis, its independent pieces on threads of their own
    RowForm lower, upper;
    run_together({[&] { lower = cholesky_forward(n, l_ptr, l_idx, l_val); },
                  [&] { upper = cholesky_backward(n, l_ptr, l_idx, l_val); }});
    Sweep sf, sb, wf, wb, pf, pb;
    int32_t parts = 0;
    std::vector<int32_t> part_of;
    run_together({[&] { sf = schedule(lower, ident, ident); }, [&] { sb = schedule(upper, ident, ident); },
                  [&] { wf = schedule(lower, gather, ident); }, [&] { wb = schedule(upper, ident, gather); },
                  [&] {
                    if (target >= 2) part_of = tree_parts(lower, target, parts);
                    if (parts >= 2) pf = schedule_parts(lower, gather, ident, part_of, parts, false);
                  },
                  [&] {
                    // the same parts, found again on this thread (tree_parts is deterministic)
                    int32_t p = 0;
                    if (target >= 2) {
                      const std::vector<int32_t> mine = tree_parts(lower, target, p);
                      if (p >= 2) pb = schedule_parts(upper, ident, gather, mine, p, true);
                    }
                  }});
    f->fits = fits_columns(f.get(), sf.levels + sb.levels);
    if (f->fits) {
      // cholesky.rs:570-579: y[i] = b[perm[i]], and x[perm[i]] = y[i]; long rows split into
      // segments whose terms are all inline (split_rows), so no level loads terms on demand
      if (f->split_cap > 0)
        run_together({[&] { sf = split_rows(sf, f->split_cap); }, [&] { sb = split_rows(sb, f->split_cap); }});
      f->column_split_cap = f->split_cap;
      f->cfwd.upload(std::move(sf), ctx->stream, false);
      f->cbwd.upload(std::move(sb), ctx->stream, false);
      f->rhs_row.upload(inverse_i32(gather), ctx->stream);
      f->sol_of.upload(inverse_i32(gather), ctx->stream);
    }
    f->fwd.upload(std::move(wf), ctx->stream, true);
    f->bwd.upload(std::move(wb), ctx->stream, true);
    if (parts >= 2) {
      if (pf.parts && pb.parts) {
        // the same parts with their long rows split within each part (blocks of vectors take
        // these; pair blocks keep the rows whole, as their reach is found over them)
        if (f->split_cap > 0) {
          Sweep spf, spb;
          run_together({[&] { spf = split_rows(pf, f->split_cap); }, [&] { spb = split_rows(pb, f->split_cap); }});
          if (spf.positions() > spf.n || spb.positions() > spb.n) {
            f->spfwd.upload(std::move(spf), ctx->stream, true);
            f->spbwd.upload(std::move(spb), ctx->stream, true);
            f->parts_split = true;
          }
        }
        // the top of the sweeps blocks of vectors take, renumbered into LDS slots
        const Sweep& tf = f->parts_split ? f->spfwd.host : pf;
        const Sweep& tb = f->parts_split ? f->spbwd.host : pb;
        const TopSlots ts = top_slots(tf, tb);
        if (ts.rows > 0) f->top_slots.upload(ts, tf, tb, ctx->stream);
        // and their parts, in LDS slots too
        const PartSlots ps = part_slots(tf, tb);
        if (ps.parts > 0 && f->parts_lds.fwd.upload(ps.fwd, ps.parts, tf, ctx->stream) &&
            f->parts_lds.bwd.upload(ps.bwd, ps.parts, tb, ctx->stream))
          f->parts_lds.parts = ps.parts;
        f->pfwd.upload(std::move(pf), ctx->stream, true);
        f->pbwd.upload(std::move(pb), ctx->stream, true);
        f->part_levels = PartLevels{f->pfwd.most_part_levels(), f->pfwd.top_levels(), f->pbwd.top_levels(),
                                    f->pbwd.most_part_levels()};
        f->partitioned = true;
      }
    }
    check(hipStreamSynchronize(ctx->stream), "upload factor");
    finish_factor(f.get());
    *out = f.release();
  });
}

int32_t ybus_hip_factor_upload_lu(ybus_hip_ctx* ctx, uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx,
                                  const double* l_val, const uint64_t* u_ptr, const uint64_t* u_idx, const double* u_val,
                                  const uint64_t* row_perm, const uint64_t* col_perm, ybus_hip_factor** out) {
  return guard([&] {
    if (!ctx || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t entries = check_lu(n, l_ptr, l_idx, l_val, u_ptr, u_idx, u_val, row_perm, col_perm);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto f = std::make_unique<ybus_hip_factor>();
    begin_factor(f.get(), ctx, (int64_t)n);
    f->lu = true;
    f->nnz_factor = (int64_t)entries;
    std::vector<int64_t> rp(n), cp(n), ident(n);
    for (uint64_t i = 0; i < n; ++i) {
      rp[i] = (int64_t)row_perm[i];
      cp[i] = (int64_t)col_perm[i];
      ident[i] = (int64_t)i;
    }
    // the host analysis, its independent pieces on threads of their own
    RowForm l, u, ut, lt;
    run_together({[&] { l = lu_forward(n, l_ptr, l_idx, l_val); }, [&] { u = lu_backward(n, u_ptr, u_idx, u_val); },
                  [&] { ut = lu_transpose_forward(n, u_ptr, u_idx, u_val); },
                  [&] { lt = lu_transpose_backward(n, l_ptr, l_idx, l_val); }});
    Sweep sl, su, stu, stl, wl, wu, wut, wlt;
    run_together({[&] { sl = schedule(l, ident, ident); }, [&] { su = schedule(u, ident, ident); },
                  [&] { stu = schedule(ut, ident, ident); }, [&] { stl = schedule(lt, ident, ident); },
                  [&] { wl = schedule(l, rp, ident); }, [&] { wu = schedule(u, ident, cp); },
                  [&] { wut = schedule(ut, cp, ident); }, [&] { wlt = schedule(lt, ident, rp); }});
    f->fits = fits_columns(f.get(), std::max(sl.levels + su.levels, stu.levels + stl.levels));
    if (f->fits) {
      // lu.rs:770-785: y[i] = b[P[i]], x[Q[i]] = y[i]; transposed (lu.rs:506-533): through Q, then P;
      // long rows split into segments whose terms are all inline
      if (f->split_cap > 0) {
        const int32_t cap = f->split_cap;
        run_together({[&] { sl = split_rows(sl, cap); }, [&] { su = split_rows(su, cap); },
                      [&] { stu = split_rows(stu, cap); }, [&] { stl = split_rows(stl, cap); }});
      }
      f->column_split_cap = f->split_cap;
      f->cfwd.upload(std::move(sl), ctx->stream, false);
      f->cbwd.upload(std::move(su), ctx->stream, false);
      f->ctfwd.upload(std::move(stu), ctx->stream, false);
      f->ctbwd.upload(std::move(stl), ctx->stream, false);
      f->rhs_row.upload(inverse_i32(rp), ctx->stream);
      f->sol_of.upload(inverse_i32(cp), ctx->stream);
      f->t_rhs_row.upload(inverse_i32(cp), ctx->stream);
      f->t_sol_of.upload(inverse_i32(rp), ctx->stream);
    }
    if (f->split_cap > 0) {
      Sweep sp[4];
      const int32_t cap = f->split_cap;
      run_together({[&] { sp[0] = split_rows(wl, cap); }, [&] { sp[1] = split_rows(wu, cap); },
                    [&] { sp[2] = split_rows(wut, cap); }, [&] { sp[3] = split_rows(wlt, cap); }});
      // built only when some row is long enough to be split
      bool any = false;
      for (const Sweep& q : sp) any = any || q.positions() > q.n;
      if (any) {
        f->sfwd.upload(std::move(sp[0]), ctx->stream, true);
        f->sbwd.upload(std::move(sp[1]), ctx->stream, true);
        f->stfwd.upload(std::move(sp[2]), ctx->stream, true);
        f->stbwd.upload(std::move(sp[3]), ctx->stream, true);
      } else {
        f->split_cap = 0;
      }
    }
    f->fwd.upload(std::move(wl), ctx->stream, true);
    f->bwd.upload(std::move(wu), ctx->stream, true);
    f->tfwd.upload(std::move(wut), ctx->stream, true);
    f->tbwd.upload(std::move(wlt), ctx->stream, true);
    f->has_transpose = true;
    f->lu_host.l_ptr.assign(l_ptr, l_ptr + n + 1);
    f->lu_host.l_idx.assign(l_idx, l_idx + l_ptr[n]);
    f->lu_host.u_ptr.assign(u_ptr, u_ptr + n + 1);
    f->lu_host.u_idx.assign(u_idx, u_idx + u_ptr[n]);
    f->lu_host.row_perm.assign(row_perm, row_perm + n);
    f->lu_host.col_perm.assign(col_perm, col_perm + n);
    f->lu_host.values.assign(l_val, l_val + l_ptr[n]);
    f->lu_host.values.insert(f->lu_host.values.end(), u_val, u_val + u_ptr[n]);
    check(hipStreamSynchronize(ctx->stream), "upload factor");
    finish_factor(f.get());
    *out = f.release();
  });
}

// ---- new values for an uploaded LU factor (lu.rs:662-757) ----

// An LU factor with its pattern kept, real or complex as `complex` says
static void require_lu(const ybus_hip_factor* f, bool complex) {
  if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "null factor");
  if (!f->lu || f->lu_host.l_ptr.empty()) fail(YBUS_HIP_INVALID_ARGUMENT, "new values need an LU factor");
  if (f->complex != complex)
    fail(YBUS_HIP_INVALID_ARGUMENT, complex ? "a real factor takes the real refactor and value update"
                                            : "a complex factor takes the complex refactor and value update");
}

// The sweeps' refresh maps: each sweep scheduled again from factors whose values are their own
// positions (plus one), so every value slot of the uploaded layout names its source
static RefactorPlan& refactor_plan(ybus_hip_factor* f) {
  if (f->refactor) return *f->refactor;
  auto plan = std::make_unique<RefactorPlan>();
  const LuPattern& h = f->lu_host;
  const int64_t n = f->n;
  const uint64_t nl = h.l_ptr[n], nu = h.u_ptr[n];
  std::vector<double> lpos(nl), upos(nu);
  for (uint64_t p = 0; p < nl; ++p) lpos[p] = (double)(p + 1);
  for (uint64_t p = 0; p < nu; ++p) upos[p] = (double)(nl + p + 1);
  std::vector<int64_t> rp(n), cp(n), ident(n);
  for (int64_t i = 0; i < n; ++i) {
    rp[i] = (int64_t)h.row_perm[i];
    cp[i] = (int64_t)h.col_perm[i];
    ident[i] = i;
  }
  hipStream_t stream = f->stream();
  if (f->complex) {
    // complex sweeps: the same schedules over forms whose real parts are the positions
    std::vector<double> zl(nl, 0.0), zu(nu, 0.0);
    RowForm zlf, zuf, zutf, zltf;
    run_together({[&] { zlf = complex_form(lu_forward(n, h.l_ptr.data(), h.l_idx.data(), lpos.data()),
                                           lu_forward(n, h.l_ptr.data(), h.l_idx.data(), zl.data()), false); },
                  [&] { zuf = complex_form(lu_backward(n, h.u_ptr.data(), h.u_idx.data(), upos.data()),
                                           lu_backward(n, h.u_ptr.data(), h.u_idx.data(), zu.data()), false); },
                  [&] { zutf = complex_form(lu_transpose_forward(n, h.u_ptr.data(), h.u_idx.data(), upos.data()),
                                            lu_transpose_forward(n, h.u_ptr.data(), h.u_idx.data(), zu.data()), false); },
                  [&] { zltf = complex_form(lu_transpose_backward(n, h.l_ptr.data(), h.l_idx.data(), lpos.data()),
                                            lu_transpose_backward(n, h.l_ptr.data(), h.l_idx.data(), zl.data()), false); }});
    Sweep zs[4];
    run_together({[&] { zs[0] = schedule(zlf, rp, ident); }, [&] { zs[1] = schedule(zuf, ident, cp); },
                  [&] { zs[2] = schedule(zutf, cp, ident); }, [&] { zs[3] = schedule(zltf, ident, rp); }});
    DevSweepC* targets[4] = {&f->zfwd, &f->zbwd, &f->ztfwd, &f->ztbwd};
    for (int q = 0; q < 4; ++q) {
      const Sweep& z = zs[q];
      if (z.n != targets[q]->n || z.levels != targets[q]->levels || z.unit != targets[q]->unit)
        fail(YBUS_HIP_INTERNAL, "a sweep's schedule differs from its upload");
      plan->refresh.emplace_back();
      SweepRefresh& r = plan->refresh.back();
      r.zsweep = targets[q];
      std::vector<int32_t> vp(z.val.size());
      for (size_t p = 0; p < vp.size(); ++p) vp[p] = (int32_t)z.val[p] - 1;
      r.vpos.upload(vp, stream);
      if (!z.unit) {
        std::vector<int32_t> dp(z.diag.size());
        for (size_t k = 0; k < dp.size(); ++k) dp[k] = (int32_t)z.diag[k] - 1;
        r.dpos.upload(dp, stream);
      }
      check(hipStreamSynchronize(stream), "refresh maps");
    }
    plan->vals.upload(h.values, stream);
    plan->next.reserve(h.values.size());
    check(hipStreamSynchronize(stream), "refresh maps");
    f->refactor = std::move(plan);
    return *f->refactor;
  }
  RowForm l, u, ut, lt;
  run_together({[&] { l = lu_forward(n, h.l_ptr.data(), h.l_idx.data(), lpos.data()); },
                [&] { u = lu_backward(n, h.u_ptr.data(), h.u_idx.data(), upos.data()); },
                [&] { ut = lu_transpose_forward(n, h.u_ptr.data(), h.u_idx.data(), upos.data()); },
                [&] { lt = lu_transpose_backward(n, h.l_ptr.data(), h.l_idx.data(), lpos.data()); }});
  auto add = [&](DevSweep& d, const Sweep& s) {
    if (s.row_ptr != d.host.row_ptr || s.dep != d.host.dep || s.unit != d.host.unit)
      fail(YBUS_HIP_INTERNAL, "a sweep's schedule differs from its upload");
    plan->refresh.emplace_back();
    SweepRefresh& r = plan->refresh.back();
    r.sweep = &d;
    std::vector<int32_t> vp(s.val.size());
    for (size_t p = 0; p < vp.size(); ++p) vp[p] = (int32_t)s.val[p] - 1;
    r.vpos.upload(vp, stream);
    if (!s.unit) {
      std::vector<int32_t> dp(s.diag.size());
      for (size_t k = 0; k < dp.size(); ++k) dp[k] = (int32_t)s.diag[k] - 1;
      r.dpos.upload(dp, stream);
    }
    check(hipStreamSynchronize(stream), "refresh maps");
  };
  Sweep s[8];
  run_together({[&] { s[0] = schedule(l, rp, ident); }, [&] { s[1] = schedule(u, ident, cp); },
                [&] { s[2] = schedule(ut, cp, ident); }, [&] { s[3] = schedule(lt, ident, rp); },
                [&] { if (f->fits) s[4] = schedule(l, ident, ident); },
                [&] { if (f->fits) s[5] = schedule(u, ident, ident); },
                [&] { if (f->fits) s[6] = schedule(ut, ident, ident); },
                [&] { if (f->fits) s[7] = schedule(lt, ident, ident); }});
  add(f->fwd, s[0]);
  add(f->bwd, s[1]);
  add(f->tfwd, s[2]);
  add(f->tbwd, s[3]);
  if (f->split_sweeps()) {
    // the split sweeps' maps come from splitting the position schedules the same way; a
    // continuation segment's diagonal (1) maps to no position (-1)
    Sweep sp[4];
    const int32_t cap = f->split_cap;
    run_together({[&] { sp[0] = split_rows(s[0], cap, 0.0); }, [&] { sp[1] = split_rows(s[1], cap, 0.0); },
                  [&] { sp[2] = split_rows(s[2], cap, 0.0); }, [&] { sp[3] = split_rows(s[3], cap, 0.0); }});
    add(f->sfwd, sp[0]);
    add(f->sbwd, sp[1]);
    add(f->stfwd, sp[2]);
    add(f->stbwd, sp[3]);
  }
  if (f->fits) {
    if (f->column_split_cap > 0) {
      const int32_t cap = f->column_split_cap;
      run_together({[&] { s[4] = split_rows(s[4], cap, 0.0); }, [&] { s[5] = split_rows(s[5], cap, 0.0); },
                    [&] { s[6] = split_rows(s[6], cap, 0.0); }, [&] { s[7] = split_rows(s[7], cap, 0.0); }});
    }
    add(f->cfwd, s[4]);
    add(f->cbwd, s[5]);
    add(f->ctfwd, s[6]);
    add(f->ctbwd, s[7]);
  }
  plan->vals.upload(h.values, stream);
  plan->next.reserve(h.values.size());
  check(hipStreamSynchronize(stream), "refresh maps");
  f->refactor = std::move(plan);
  return *f->refactor;
}

// Rewrites every sweep's values from plan.next, then makes next the factor's values; returns the
// HIP-event milliseconds of the rewrite
static double refresh_sweeps(ybus_hip_factor* f, RefactorPlan& plan) {
  hipStream_t stream = f->stream();
  check(hipEventRecord(f->ev0, stream), "hipEventRecord");
  for (SweepRefresh& r : plan.refresh) {
    if (r.zsweep) {
      DevSweepC& z = *r.zsweep;
      check(launch_refresh_complex(stream, z.n, (int64_t)z.val.len, r.vpos.p, z.unit ? nullptr : r.dpos.p,
                                   reinterpret_cast<const double2*>(plan.next.p), z.val.p, z.unit ? nullptr : z.diag.p,
                                   z.val32.p, z.unit ? nullptr : z.diag32.p),
            "refresh sweeps");
      continue;
    }
    DevSweep& d = *r.sweep;
    RefreshDev v;
    v.n = d.host.positions();
    v.entries = (int64_t)d.host.dep.size();
    v.row_ptr = d.row_ptr.p;
    v.vpos = r.vpos.p;
    v.dpos = d.host.unit ? nullptr : r.dpos.p;
    v.val64 = d.val64.p;
    v.val32 = d.val32.p;
    v.diag64 = d.diag64.p;
    v.diag32 = d.diag32.p;
    v.val01 = d.val01.p;
    v.val23 = d.val23.p;
    v.v32 = d.v32.p;
    v.rdiag64 = d.rdiag64.p;
    v.rdiag32 = d.rdiag32.p;
    check(launch_refresh(stream, v, plan.next.p), "refresh sweeps");
  }
  check(hipEventRecord(f->ev1, stream), "hipEventRecord");
  check(hipEventSynchronize(f->ev1), "refresh sweeps");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, f->ev0, f->ev1), "hipEventElapsedTime");
  std::swap(plan.vals.p, plan.next.p);
  std::swap(plan.vals.len, plan.next.len);
  return ms;
}

// Binary search for `row` in idx[lo, hi) (increasing); -1 when absent
static int64_t find_row(const std::vector<uint64_t>& idx, uint64_t lo, uint64_t hi, uint64_t row) {
  auto it = std::lower_bound(idx.begin() + (int64_t)lo, idx.begin() + (int64_t)hi, row);
  return it != idx.begin() + (int64_t)hi && *it == row ? (int64_t)(it - idx.begin()) : -1;
}

// The device refactorization's arrays for the matrix pattern (a_ptr, a_idx)
static void build_numeric(ybus_hip_factor* f, RefactorPlan& plan, const uint64_t* a_ptr, const uint64_t* a_idx) {
  const LuPattern& h = f->lu_host;
  const int64_t n = f->n;
  const uint64_t nnz_a = a_ptr[n];
  // each column of U: rows increasing, all above the diagonal but the last; of L: the diagonal,
  // then rows increasing below it
  int32_t longest_l = 0;
  int64_t x_cap = 1;
  for (int64_t k = 0; k < n; ++k) {
    for (uint64_t p = h.u_ptr[k]; p + 1 < h.u_ptr[k + 1]; ++p)
      if (h.u_idx[p] >= (uint64_t)k || (p > h.u_ptr[k] && h.u_idx[p] <= h.u_idx[p - 1]))
        fail(YBUS_HIP_INVALID_ARGUMENT, "refactor needs U's columns in increasing row order");
    for (uint64_t p = h.l_ptr[k] + 1; p < h.l_ptr[k + 1]; ++p)
      if (h.l_idx[p] <= (uint64_t)k || (p > h.l_ptr[k] + 1 && h.l_idx[p] <= h.l_idx[p - 1]))
        fail(YBUS_HIP_INVALID_ARGUMENT, "refactor needs L's columns in increasing row order");
    longest_l = std::max<int32_t>(longest_l, (int32_t)(h.l_ptr[k + 1] - h.l_ptr[k] - 1));
    x_cap = std::max<int64_t>(x_cap, (int64_t)(h.u_ptr[k + 1] - h.u_ptr[k] + h.l_ptr[k + 1] - h.l_ptr[k] - 1));
  }
  x_cap = round_up(x_cap, 8);
  const int32_t stage = refactor_stage_for(longest_l);
  if (longest_l > REFACTOR_STAGE ||
      refactor_lds_bytes((int32_t)std::min<int64_t>(x_cap, INDEX_LIMIT), stage, f->complex) > 65536 - 256)
    fail(YBUS_HIP_TOO_LARGE, "a column of the factor is too long for the device refactor (L below its diagonal: " +
                                 std::to_string(longest_l) + ", L and U: " + std::to_string(x_cap) + ")");
  std::vector<uint64_t> step_of(n);
  for (int64_t s = 0; s < n; ++s) step_of[h.row_perm[s]] = (uint64_t)s;
  // slot of pivot step `row` in column k's x, or -1
  auto slot_of = [&](int64_t k, uint64_t row) -> int64_t {
    const uint64_t u0 = h.u_ptr[k], nu = h.u_ptr[k + 1] - u0;
    if (row == (uint64_t)k) return (int64_t)nu - 1;
    if (row < (uint64_t)k) {
      const int64_t p = find_row(h.u_idx, u0, h.u_ptr[k + 1] - 1, row);
      return p < 0 ? -1 : p - (int64_t)u0;
    }
    const int64_t p = find_row(h.l_idx, h.l_ptr[k] + 1, h.l_ptr[k + 1], row);
    return p < 0 ? -1 : (int64_t)nu + (p - (int64_t)h.l_ptr[k] - 1);
  };
  std::vector<int32_t> a_slot(nnz_a), level(n, 0);
  std::vector<int64_t> upd_start(n + 1, 0);
  std::vector<int32_t> target;
  std::vector<int4> step(h.u_idx.size(), make_int4(0, 0, 0, 0));
  for (int64_t k = 0; k < n; ++k) {
    const uint64_t col = h.col_perm[k];
    for (uint64_t p = a_ptr[col]; p < a_ptr[col + 1]; ++p) {
      if (p > a_ptr[col] && a_idx[p] <= a_idx[p - 1])
        fail(YBUS_HIP_INVALID_ARGUMENT, "the matrix's columns must list their rows in increasing order");
      const int64_t slot = slot_of(k, step_of[a_idx[p]]);
      if (slot < 0) fail(YBUS_HIP_PATTERN_MISMATCH, "the matrix has an entry outside the factor's pattern");
      a_slot[p] = (int32_t)slot;
    }
    upd_start[k] = (int64_t)target.size();
    for (uint64_t p = h.u_ptr[k]; p + 1 < h.u_ptr[k + 1]; ++p) {
      const uint64_t j = h.u_idx[p];
      level[k] = std::max(level[k], level[j] + 1);
      step[p] = make_int4((int32_t)j, (int32_t)(h.l_ptr[j] + 1), (int32_t)(h.l_ptr[j + 1] - h.l_ptr[j] - 1), 0);
      for (uint64_t q = h.l_ptr[j] + 1; q < h.l_ptr[j + 1]; ++q) {
        const int64_t slot = slot_of(k, h.l_idx[q]);
        if (slot < 0) fail(YBUS_HIP_INVALID_ARGUMENT, "the factor's pattern is not closed under elimination");
        target.push_back((int32_t)slot);
      }
    }
    if ((int64_t)target.size() > INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "too many multiply-subtracts for the device refactor");
  }
  upd_start[n] = (int64_t)target.size();
  // columns by (level, column), each position's column and ranges
  const int32_t levels = n ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  std::vector<int32_t> start(levels + 1, 0);
  for (int64_t k = 0; k < n; ++k) ++start[level[k] + 1];
  for (int32_t v = 0; v < levels; ++v) start[v + 1] += start[v];
  std::vector<int4> head(n), span(n);
  for (int64_t k = 0; k < n; ++k) {
    const int32_t c = start[level[k]]++;
    const uint64_t col = h.col_perm[k];
    head[c] = make_int4((int32_t)k, (int32_t)h.u_ptr[k], (int32_t)(h.u_ptr[k + 1] - h.u_ptr[k]), (int32_t)h.l_ptr[k]);
    span[c] = make_int4((int32_t)(h.l_ptr[k + 1] - h.l_ptr[k] - 1), (int32_t)a_ptr[col], (int32_t)a_ptr[col + 1],
                        (int32_t)upd_start[k]);
  }
  hipStream_t stream = f->stream();
  plan.head.upload(head, stream);
  plan.span.upload(span, stream);
  plan.step.upload(step, stream);
  plan.a_slot.upload(a_slot, stream);
  plan.target.upload(target, stream);
  plan.a_val.reserve((f->complex ? 2 : 1) * nnz_a);
  plan.control.reserve(4);
  check(hipStreamSynchronize(stream), "refactor plan");
  plan.x_cap = (int32_t)x_cap;
  plan.stage = stage;
  // two wavefronts per SIMD: the time is the column graph's chain of hand-overs, which 512 to
  // 8,192 wavefronts all cover (1.06-1.17 ms on a 25,000-bus Jacobian; fewer are slightly faster,
  // as fewer waiting wavefronts poll)
  plan.grid = (int)std::min<int64_t>(std::max<int64_t>(n, 1), 8 * (int64_t)f->ctx->cus);
  plan.stats.multiply_subtracts = (uint64_t)target.size();
  plan.a_ptr.assign(a_ptr, a_ptr + n + 1);
  plan.a_idx.assign(a_idx, a_idx + nnz_a);
  plan.numeric = true;
}

// The device refactor of a real or complex LU factor (values as (re, im) pairs when complex)
static void refactor_lu(ybus_hip_factor* f, uint64_t n, const uint64_t* a_ptr, const uint64_t* a_idx,
                        const double* a_val, double pivot_threshold, int32_t* outcome, bool complex) {
  require_lu(f, complex);
  if (!a_ptr || !outcome) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  if (a_ptr[n] && (!a_idx || !a_val)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  if (!(pivot_threshold > 0.0 && pivot_threshold <= 1.0))
    fail(YBUS_HIP_INVALID_ARGUMENT, "the pivot threshold must be in (0, 1]");
  if ((int64_t)n != f->n) fail(YBUS_HIP_PATTERN_MISMATCH, "the matrix's shape differs from the factor's");
  std::string why;
  if (!valid_compressed((int64_t)n, (int64_t)n, a_ptr, a_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "A: " + why);
  if (a_ptr[n] > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "the matrix has too many entries");
  check(hipSetDevice(f->ctx->device), "hipSetDevice");
  RefactorPlan& plan = refactor_plan(f);
  if (!plan.numeric) {
    build_numeric(f, plan, a_ptr, a_idx);
  } else if (!std::equal(plan.a_ptr.begin(), plan.a_ptr.end(), a_ptr) ||
             !std::equal(plan.a_idx.begin(), plan.a_idx.end(), a_idx)) {
    fail(YBUS_HIP_PATTERN_MISMATCH, "the matrix's pattern differs from the one this factor was refactored with");
  }
  hipStream_t stream = f->stream();
  const uint64_t parts = complex ? 2 : 1;  // doubles per value
  const uint64_t nnz_a = a_ptr[n];
  f->staging.to_device(stream, plan.a_val.p, a_val, parts * nnz_a * sizeof(double));
  const uint64_t nl = f->lu_host.l_ptr[n];
  check(hipEventRecord(f->ev0, stream), "hipEventRecord");
  check(launch_fill_u64(stream, reinterpret_cast<uint64_t*>(plan.next.p), REFACTOR_PENDING, (int64_t)(parts * nl)),
        "refactor values");
  check(launch_fill(stream, plan.control.p, 0, 4), "refactor control");
  check(launch_fill(stream, plan.control.p + 1, (int32_t)f->n, 1), "refactor control");
  RefactorDev d;
  d.n = (int32_t)f->n;
  d.head = plan.head.p;
  d.span = plan.span.p;
  d.step = plan.step.p;
  d.a_slot = plan.a_slot.p;
  d.a_val = plan.a_val.p;
  d.target = plan.target.p;
  d.l_val = plan.next.p;
  d.u_val = plan.next.p + parts * nl;
  d.control = plan.control.p;
  d.tau = pivot_threshold;
  d.x_cap = plan.x_cap;
  d.spin_limit = refactor_spin_limit();
  check(launch_refactor(stream, d, plan.grid, plan.stage, complex), "refactor");
  check(hipEventRecord(f->ev1, stream), "hipEventRecord");
  int32_t control[4] = {0, 0, 0, 0};
  check(hipMemcpyAsync(control, plan.control.p, sizeof(control), hipMemcpyDeviceToHost, stream), "refactor control");
  check(hipStreamSynchronize(stream), "refactor");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, f->ev0, f->ev1), "hipEventElapsedTime");
  plan.stats.calls += 1;
  plan.stats.refactor_ms += ms;
  if (control[2] != 0)
    fail(YBUS_HIP_INTERNAL, "the device refactor gave up " + std::to_string(control[2]) +
                                " column(s) waiting on another; the factor is unchanged");
  if (control[0] < (int32_t)f->n) fail(YBUS_HIP_INTERNAL, "the device refactor left columns untaken");
  if (control[1] < (int32_t)f->n) {
    *outcome = YBUS_HIP_REFACTOR_NEEDS_PIVOTING;
    return;
  }
  plan.stats.refresh_ms += refresh_sweeps(f, plan);
  plan.stats.reused += 1;
  // a matrix attached for refinement that is this one (CSC, the same pattern) takes its values
  if (!complex && f->a.nnz == nnz_a && !f->a.src_row_major && f->a.src_ptr == plan.a_ptr &&
      f->a.src_idx == plan.a_idx) {
    check(launch_gather(stream, plan.a_val.p, f->a.order_dev.p, (int64_t)nnz_a, f->a.val.p), "matrix values");
    check(launch_gather(stream, plan.a_val.p, f->at.t_order.p, (int64_t)nnz_a, f->at.val.p), "matrix values");
    check(hipStreamSynchronize(stream), "matrix values");
  }
  *outcome = YBUS_HIP_REFACTOR_REUSED;
}

int32_t ybus_hip_factor_refactor_lu(ybus_hip_factor* f, uint64_t n, const uint64_t* a_ptr, const uint64_t* a_idx,
                                    const double* a_val, double pivot_threshold, int32_t* outcome) {
  return guard([&] { refactor_lu(f, n, a_ptr, a_idx, a_val, pivot_threshold, outcome, false); });
}

int32_t ybus_hip_factor_refactor_lu_complex(ybus_hip_factor* f, uint64_t n, const uint64_t* a_ptr,
                                            const uint64_t* a_idx, const double* a_val, double pivot_threshold,
                                            int32_t* outcome) {
  return guard([&] { refactor_lu(f, n, a_ptr, a_idx, a_val, pivot_threshold, outcome, true); });
}

static void update_lu_values(ybus_hip_factor* f, const double* l_val, const double* u_val, bool complex) {
  require_lu(f, complex);
  const uint64_t parts = complex ? 2 : 1;
  const uint64_t nl = parts * f->lu_host.l_ptr[f->n], nu = parts * f->lu_host.u_ptr[f->n];
  if ((nl && !l_val) || (nu && !u_val)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  check(hipSetDevice(f->ctx->device), "hipSetDevice");
  RefactorPlan& plan = refactor_plan(f);
  hipStream_t stream = f->stream();
  f->staging.to_device(stream, plan.next.p, l_val, nl * sizeof(double));
  f->staging.to_device(stream, plan.next.p + nl, u_val, nu * sizeof(double));
  plan.stats.refresh_ms += refresh_sweeps(f, plan);
  plan.stats.updates += 1;
}

int32_t ybus_hip_factor_update_lu_values(ybus_hip_factor* f, const double* l_val, const double* u_val) {
  return guard([&] { update_lu_values(f, l_val, u_val, false); });
}

int32_t ybus_hip_factor_update_lu_values_complex(ybus_hip_factor* f, const double* l_val, const double* u_val) {
  return guard([&] { update_lu_values(f, l_val, u_val, true); });
}

int32_t ybus_hip_factor_refactor_stats(const ybus_hip_factor* f, ybus_hip_refactor_stats* out) {
  return guard([&] {
    require_lu(f, f && f->complex);
    if (!out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    *out = f->refactor ? f->refactor->stats : ybus_hip_refactor_stats{};
  });
}

int32_t ybus_hip_factor_lu_values(ybus_hip_factor* f, double* l_val, double* u_val) {
  return guard([&] {
    require_lu(f, f && f->complex);
    const uint64_t parts = f->complex ? 2 : 1;
    const uint64_t nl = parts * f->lu_host.l_ptr[f->n], nu = parts * f->lu_host.u_ptr[f->n];
    std::vector<double> v;
    if (f->refactor) {
      check(hipSetDevice(f->ctx->device), "hipSetDevice");
      v.resize(nl + nu);
      check(hipMemcpy(v.data(), f->refactor->vals.p, (nl + nu) * sizeof(double), hipMemcpyDeviceToHost), "values");
    }
    const std::vector<double>& src = f->refactor ? v : f->lu_host.values;
    if (l_val && nl) std::memcpy(l_val, src.data(), nl * sizeof(double));
    if (u_val && nu) std::memcpy(u_val, src.data() + nl, nu * sizeof(double));
  });
}

int32_t ybus_hip_factor_upload_cholesky_complex(ybus_hip_ctx* ctx, uint64_t n, const uint64_t* l_ptr,
                                                const uint64_t* l_idx, const double* l_val, const uint64_t* perm,
                                                ybus_hip_factor** out) {
  return guard([&] {
    if (!ctx || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t entries = check_cholesky(n, l_ptr, l_idx, l_val, perm);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto f = std::make_unique<ybus_hip_factor>();
    begin_factor(f.get(), ctx, (int64_t)n);
    f->lu = false;
    f->complex = true;
    f->nnz_factor = 2 * (int64_t)entries;
    std::vector<double> re, im;
    split_complex(l_val, entries, re, im);
    std::vector<int64_t> gather(n), ident(n);
    for (uint64_t i = 0; i < n; ++i) {
      gather[i] = (int64_t)perm[i];
      ident[i] = (int64_t)i;
    }
    // L forward, and L^H (L's columns conjugated below the diagonal) backward
    const RowForm lower =
        complex_form(cholesky_forward(n, l_ptr, l_idx, re.data()), cholesky_forward(n, l_ptr, l_idx, im.data()), false);
    const RowForm upper = complex_form(cholesky_backward(n, l_ptr, l_idx, re.data()),
                                       cholesky_backward(n, l_ptr, l_idx, im.data()), true);
    f->zfwd.upload(schedule(lower, gather, ident), ctx->stream);
    f->zbwd.upload(schedule(upper, ident, gather), ctx->stream);
    check(hipStreamSynchronize(ctx->stream), "upload factor");
    finish_factor(f.get());
    *out = f.release();
  });
}

int32_t ybus_hip_factor_upload_lu_complex(ybus_hip_ctx* ctx, uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx,
                                          const double* l_val, const uint64_t* u_ptr, const uint64_t* u_idx,
                                          const double* u_val, const uint64_t* row_perm, const uint64_t* col_perm,
                                          ybus_hip_factor** out) {
  return guard([&] {
    if (!ctx || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t entries = check_lu(n, l_ptr, l_idx, l_val, u_ptr, u_idx, u_val, row_perm, col_perm);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto f = std::make_unique<ybus_hip_factor>();
    begin_factor(f.get(), ctx, (int64_t)n);
    f->lu = true;
    f->complex = true;
    f->has_transpose = true;
    f->nnz_factor = (int64_t)entries;
    std::vector<double> lre, lim, ure, uim;
    split_complex(l_val, l_ptr[n], lre, lim);
    split_complex(u_val, u_ptr[n], ure, uim);
    std::vector<int64_t> rp(n), cp(n), ident(n);
    for (uint64_t i = 0; i < n; ++i) {
      rp[i] = (int64_t)row_perm[i];
      cp[i] = (int64_t)col_perm[i];
      ident[i] = (int64_t)i;
    }
    // the plain (unconjugated) transpose, as Lu::solve_transpose takes it (lu.rs:502-533)
    const RowForm l = complex_form(lu_forward(n, l_ptr, l_idx, lre.data()), lu_forward(n, l_ptr, l_idx, lim.data()), false);
    const RowForm u = complex_form(lu_backward(n, u_ptr, u_idx, ure.data()), lu_backward(n, u_ptr, u_idx, uim.data()), false);
    const RowForm ut = complex_form(lu_transpose_forward(n, u_ptr, u_idx, ure.data()),
                                    lu_transpose_forward(n, u_ptr, u_idx, uim.data()), false);
    const RowForm lt = complex_form(lu_transpose_backward(n, l_ptr, l_idx, lre.data()),
                                    lu_transpose_backward(n, l_ptr, l_idx, lim.data()), false);
    f->zfwd.upload(schedule(l, rp, ident), ctx->stream);
    f->zbwd.upload(schedule(u, ident, cp), ctx->stream);
    f->ztfwd.upload(schedule(ut, cp, ident), ctx->stream);
    f->ztbwd.upload(schedule(lt, ident, rp), ctx->stream);
    f->lu_host.l_ptr.assign(l_ptr, l_ptr + n + 1);
    f->lu_host.l_idx.assign(l_idx, l_idx + l_ptr[n]);
    f->lu_host.u_ptr.assign(u_ptr, u_ptr + n + 1);
    f->lu_host.u_idx.assign(u_idx, u_idx + u_ptr[n]);
    f->lu_host.row_perm.assign(row_perm, row_perm + n);
    f->lu_host.col_perm.assign(col_perm, col_perm + n);
    f->lu_host.values.assign(l_val, l_val + 2 * l_ptr[n]);
    f->lu_host.values.insert(f->lu_host.values.end(), u_val, u_val + 2 * u_ptr[n]);
    check(hipStreamSynchronize(ctx->stream), "upload factor");
    finish_factor(f.get());
    *out = f.release();
  });
}

int32_t ybus_hip_factor_is_complex(const ybus_hip_factor* f) { return f && f->complex ? 1 : 0; }

void ybus_hip_factor_free(ybus_hip_factor* f) { delete f; }

uint64_t ybus_hip_factor_dim(const ybus_hip_factor* f) { return f ? (uint64_t)f->n : 0; }

uint64_t ybus_hip_factor_block_width(const ybus_hip_factor* f) { return f ? f->width : 0; }

void ybus_hip_factor_set_block_width(ybus_hip_factor* f, uint64_t width) {
  if (!f) return;
  const uint64_t most = f->n == 1 ? SINGLE_ROW_WIDTH : UINT64_MAX;
  f->width = std::clamp<uint64_t>(width, 1, most);
}

void ybus_hip_factor_timings(const ybus_hip_factor* f, ybus_hip_timings* out) {
  if (f && out) *out = f->timings;
}

void ybus_hip_factor_reset_timings(ybus_hip_factor* f) {
  if (f) f->timings = ybus_hip_timings{};
}

int32_t ybus_hip_factor_path(const ybus_hip_factor* f, uint64_t columns, int32_t pairs) {
  return f && f->use_columns((int64_t)std::min<uint64_t>(columns, INT64_MAX), pairs != 0) ? YBUS_HIP_PATH_COLUMNS
                                                                                : YBUS_HIP_PATH_WIDE;
}

void ybus_hip_factor_kernel_stats(const ybus_hip_factor* f, ybus_hip_kernel_stats* out) {
  if (f && out) *out = f->stats;
}

void ybus_hip_factor_reset_kernel_stats(ybus_hip_factor* f) {
  if (f) f->stats = ybus_hip_kernel_stats{};
}

int32_t ybus_hip_factor_set_matrix(ybus_hip_factor* f, int32_t row_major, const uint64_t* ptr, const uint64_t* idx,
                                   const double* val) {
  return guard([&] {
    if (!f || !ptr) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    require_real(f);
    const int64_t n = f->n;
    std::string why;
    if (!valid_compressed(n, n, ptr, idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "A: " + why);
    const uint64_t nnz = ptr[n];
    if (nnz && (!idx || !val)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (nnz > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "the matrix is too large for 32-bit indices");
    std::vector<int32_t> rptr(n + 1), cols(nnz);
    std::vector<double> vals(nnz);
    std::vector<int64_t> order(nnz);
    if (row_major) {
      for (int64_t i = 0; i <= n; ++i) rptr[i] = (int32_t)ptr[i];
      for (uint64_t p = 0; p < nnz; ++p) {
        cols[p] = (int32_t)idx[p];
        vals[p] = val[p];
        order[p] = (int64_t)p;
      }
    } else {
      // CSC to CSR by a counting sort by row: each row's columns come out ascending, the order
      // CsrMatrix::mul_vec sums them in (product.rs:17-27)
      std::vector<int32_t> cnt(n + 1, 0);
      for (uint64_t p = 0; p < nnz; ++p) cnt[idx[p] + 1] += 1;
      for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
      rptr = cnt;
      std::vector<int32_t> cursor(cnt.begin(), cnt.end() - 1);
      for (int64_t j = 0; j < n; ++j)
        for (uint64_t p = ptr[j]; p < ptr[j + 1]; ++p) {
          int32_t q = cursor[idx[p]]++;
          cols[q] = (int32_t)j;
          vals[q] = val[p];
          order[q] = (int64_t)p;
        }
    }
    std::vector<int32_t> rows(nnz);
    for (int64_t i = 0; i < n; ++i)
      for (int32_t q = rptr[i]; q < rptr[i + 1]; ++q) rows[q] = (int32_t)i;
    f->a.rows = f->a.cols = (int32_t)n;
    f->a.nnz = nnz;
    f->a.ptr.upload(rptr, f->stream());
    // the residual's padding entries (column 0, value 0) past the last
    f->a.col.upload(padded(cols, RESIDUAL_CHUNK), f->stream());
    f->a.val.upload(padded(vals, RESIDUAL_CHUNK), f->stream());
    f->a.row.upload(rows, f->stream());
    // A^T by rows: a counting sort of A's entries by column, rows ascending within each
    {
      std::vector<int32_t> tptr(n + 1, 0), tcol(nnz), tord(nnz);
      std::vector<double> tval(nnz);
      for (uint64_t q = 0; q < nnz; ++q) tptr[cols[q] + 1] += 1;
      for (int64_t j = 0; j < n; ++j) tptr[j + 1] += tptr[j];
      std::vector<int32_t> cursor(tptr.begin(), tptr.end() - 1);
      for (int64_t i = 0; i < n; ++i)
        for (int32_t q = rptr[i]; q < rptr[i + 1]; ++q) {
          const int32_t t = cursor[cols[q]]++;
          tcol[t] = (int32_t)i;
          tval[t] = vals[q];
          tord[t] = (int32_t)order[q];
        }
      f->at.rows = f->at.cols = (int32_t)n;
      f->at.nnz = nnz;
      f->at.ptr.upload(tptr, f->stream());
      f->at.col.upload(padded(tcol, RESIDUAL_CHUNK), f->stream());
      f->at.val.upload(padded(tval, RESIDUAL_CHUNK), f->stream());
      f->at.t_order.upload(tord, f->stream());
    }
    f->a.src_row_major = row_major != 0;
    f->a.src_ptr.assign(ptr, ptr + n + 1);
    f->a.src_idx.assign(idx, idx + nnz);
    std::vector<int32_t> order32(order.begin(), order.end());
    f->a.order_dev.upload(order32, f->stream());
    f->a.order = std::move(order);
    check(hipStreamSynchronize(f->stream()), "upload matrix");
  });
}

// Copies count host vectors into a device block of the workspace (the right-hand sides b unless
// another is named), in the workspace's layout
static void upload_vectors(ybus_hip_factor* f, const double* rhs, int64_t count, double* block = nullptr) {
  if (!block) block = f->b.p;
  const size_t len = (size_t)count * (size_t)f->n;
  if (f->layout.columns) {
    f->staging.to_device(f->stream(), block, rhs, len * sizeof(double));
  } else {
    f->staging.to_device(f->stream(), f->vec.p, rhs, len * sizeof(double));
    check(launch_to_block(f->stream(), f->vec.p, block, f->n, count, f->layout.ld), "to_block");
  }
  check(hipStreamSynchronize(f->stream()), "upload");
}

static void download_vectors(ybus_hip_factor* f, const double* block, const Layout& layout, int64_t count,
                             double* out) {
  const size_t len = (size_t)count * (size_t)f->n;
  if (layout.columns) {
    f->staging.to_host(f->stream(), out, block, len * sizeof(double));
  } else {
    check(launch_from_block(f->stream(), block, f->vec.p, f->n, count, layout.ld), "from_block");
    f->staging.to_host(f->stream(), out, f->vec.p, len * sizeof(double));
  }
}

// Copies count device vectors (vector c at src + c * ld) into a block of the workspace, in the
// workspace's layout; the caller's writes to src must be finished
static void copy_in_device(ybus_hip_factor* f, const double* src, int64_t ld, int64_t count, double* block) {
  const size_t n = (size_t)f->n;
  if (f->layout.columns) {
    check(hipMemcpy2DAsync(block, (size_t)f->layout.ld * sizeof(double), src, (size_t)ld * sizeof(double),
                           n * sizeof(double), (size_t)count, hipMemcpyDeviceToDevice, f->stream()),
          "D2D");
  } else {
    const double* vectors = src;
    if (ld != f->n) {
      check(hipMemcpy2DAsync(f->vec.p, n * sizeof(double), src, (size_t)ld * sizeof(double), n * sizeof(double),
                             (size_t)count, hipMemcpyDeviceToDevice, f->stream()),
            "D2D");
      vectors = f->vec.p;
    }
    check(launch_to_block(f->stream(), vectors, block, f->n, count, f->layout.ld), "to_block");
  }
}

// Copies count vectors of a workspace block out to device memory (vector c at dst + c * ld)
static void copy_out_device(ybus_hip_factor* f, const double* block, int64_t count, double* dst, int64_t ld) {
  const size_t n = (size_t)f->n;
  if (f->layout.columns) {
    check(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), block, (size_t)f->layout.ld * sizeof(double),
                           n * sizeof(double), (size_t)count, hipMemcpyDeviceToDevice, f->stream()),
          "D2D");
  } else if (ld == f->n) {
    check(launch_from_block(f->stream(), block, dst, f->n, count, f->layout.ld), "from_block");
  } else {
    check(launch_from_block(f->stream(), block, f->vec.p, f->n, count, f->layout.ld), "from_block");
    check(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), f->vec.p, n * sizeof(double), n * sizeof(double),
                           (size_t)count, hipMemcpyDeviceToDevice, f->stream()),
          "D2D");
  }
}

static void require_unborrowed(const ybus_hip_factor* f) {
  require_real(f);
  if (f->borrowed) fail(YBUS_HIP_INVALID_ARGUMENT, "a block borrows this factor: free it first");
}

static void read_refinement(ybus_hip_factor* f, int64_t count, uint64_t* steps, double* ratio) {
  std::vector<int32_t> st(count);
  std::vector<double> ra(count);
  check(hipMemcpy(st.data(), f->isteps.p, count * sizeof(int32_t), hipMemcpyDeviceToHost), "steps");
  check(hipMemcpy(ra.data(), f->dratio.p, count * sizeof(double), hipMemcpyDeviceToHost), "ratio");
  for (int64_t c = 0; c < count; ++c) {
    if (steps) steps[c] = (uint64_t)st[c];
    if (ratio) ratio[c] = ra[c];
  }
}

static int32_t solve_many_mode(ybus_hip_factor* f, const double* rhs, uint64_t count, double* x, int mode,
                               double tol, uint64_t max_steps, uint64_t* steps, double* ratio) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (count == 0 || f->n == 0) return;
    if (!rhs || !x) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    const bool refine = mode == YBUS_HIP_FP32_REFINED;
    if (refine && !f->a.ptr.p) fail(YBUS_HIP_INVALID_ARGUMENT, "FP32 refinement needs the matrix (ybus_hip_factor_set_matrix)");
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    const int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
    f->ensure_workspace(width, refine);
    for (uint64_t start = 0; start < count; start += (uint64_t)width) {
      const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
      auto t0 = std::chrono::steady_clock::now();
      upload_vectors(f, rhs + start * (uint64_t)f->n, chunk);
      auto t1 = std::chrono::steady_clock::now();
      if (refine) {
        f->refine(chunk, tol, (int32_t)max_steps);
        read_refinement(f, chunk, steps ? steps + start : nullptr, ratio ? ratio + start : nullptr);
      } else {
        f->solve64(mode == YBUS_HIP_TRANSPOSE, false, chunk);
      }
      auto t2 = std::chrono::steady_clock::now();
      download_vectors(f, f->x.p, f->layout, chunk, x + start * (uint64_t)f->n);
      f->timings.upload += std::chrono::duration<double>(t1 - t0).count();
      f->timings.solve += std::chrono::duration<double>(t2 - t1).count();
      f->timings.download += seconds_since(t2);
      f->timings.blocks += 1;
    }
  });
}

// Complex solves of `count` right-hand sides (n interleaved (re, im) values each), in blocks of
// block_width() columns: each block is copied in, turned into a complex block (real and
// imaginary rows), solved by one launch (zsolve), and copied out.
static void require_complex(const ybus_hip_factor* f) {
  if (!f->complex) fail(YBUS_HIP_INVALID_ARGUMENT, "a real factor takes the real solves only");
}

static void require_complex_matrix(const ybus_hip_factor* f, int mode) {
  if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_TRANSPOSE && mode != YBUS_HIP_FP32_REFINED)
    fail(YBUS_HIP_INVALID_ARGUMENT, "unknown mode");
  if (mode == YBUS_HIP_FP32_REFINED && !f->za.ptr.p)
    fail(YBUS_HIP_INVALID_ARGUMENT, "FP32 refinement needs the matrix (ybus_hip_factor_set_matrix_complex)");
}

static int32_t solve_many_complex(ybus_hip_factor* f, const double* rhs, uint64_t count, double* x, int mode,
                                  double tol, uint64_t max_steps, uint64_t* steps, double* ratio) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_complex(f);
    require_complex_matrix(f, mode);
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    if (count == 0 || f->n == 0) return;
    if (!rhs || !x) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    const bool refine = mode == YBUS_HIP_FP32_REFINED;
    const int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
    const int64_t ld = round_up(width, 16);
    const int64_t rows = 2 * f->n;
    f->reserve_complex(width, ld, refine);
    for (uint64_t start = 0; start < count; start += (uint64_t)width) {
      const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
      const size_t len = (size_t)chunk * (size_t)rows;
      auto t0 = std::chrono::steady_clock::now();
      f->staging.to_device(f->stream(), f->zvec.p, rhs + start * (uint64_t)rows, len * sizeof(double));
      check(launch_to_block(f->stream(), f->zvec.p, f->zb.p, rows, chunk, ld), "to_block");
      check(hipStreamSynchronize(f->stream()), "upload");
      auto t1 = std::chrono::steady_clock::now();
      f->zsolve(mode, chunk, ld, tol, (int32_t)max_steps);
      if (refine) read_refinement(f, chunk, steps ? steps + start : nullptr, ratio ? ratio + start : nullptr);
      auto t2 = std::chrono::steady_clock::now();
      check(launch_from_block(f->stream(), f->zx.p, f->zvec.p, rows, chunk, ld), "from_block");
      f->staging.to_host(f->stream(), x + start * (uint64_t)rows, f->zvec.p, len * sizeof(double));
      f->timings.upload += std::chrono::duration<double>(t1 - t0).count();
      f->timings.solve += std::chrono::duration<double>(t2 - t1).count();
      f->timings.download += seconds_since(t2);
      f->timings.blocks += 1;
    }
  });
}

int32_t ybus_hip_factor_solve_many_complex(ybus_hip_factor* f, const double* rhs, uint64_t count, double* x) {
  return solve_many_complex(f, rhs, count, x, YBUS_HIP_FP64, 0.0, 0, nullptr, nullptr);
}

int32_t ybus_hip_factor_solve_transpose_many_complex(ybus_hip_factor* f, const double* rhs, uint64_t count,
                                                     double* x) {
  return solve_many_complex(f, rhs, count, x, YBUS_HIP_TRANSPOSE, 0.0, 0, nullptr, nullptr);
}

int32_t ybus_hip_factor_solve_many_refined_complex(ybus_hip_factor* f, const double* rhs, uint64_t count,
                                                   double tolerance, uint64_t max_steps, double* x, uint64_t* steps,
                                                   double* ratio) {
  return solve_many_complex(f, rhs, count, x, YBUS_HIP_FP32_REFINED, tolerance, max_steps, steps, ratio);
}

int32_t ybus_hip_factor_solve_device_complex(ybus_hip_factor* f, int32_t mode, double tolerance, uint64_t max_steps,
                                             const double* b, uint64_t ld_b, uint64_t count, double* x,
                                             uint64_t ld_x) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_complex(f);
    require_complex_matrix(f, mode);
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    if (count == 0 || f->n == 0) return;
    if (!b || !x) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (std::min(ld_b, ld_x) < (uint64_t)f->n) fail(YBUS_HIP_INVALID_ARGUMENT, "the vectors' stride is below dim()");
    if (count > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "too many vectors");
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    const bool refine = mode == YBUS_HIP_FP32_REFINED;
    const int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
    const int64_t ld = round_up(width, 16);
    const int64_t rows = 2 * f->n;
    const size_t row_bytes = (size_t)rows * sizeof(double);
    f->reserve_complex(width, ld, refine);
    for (uint64_t start = 0; start < count; start += (uint64_t)width) {
      const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
      // vector c of the chunk at b + 2 (start + c) ld_b doubles: packed into the staging block
      // first unless the caller's vectors are already packed
      const double* in = b + 2 * start * ld_b;
      if (ld_b != (uint64_t)f->n) {
        check(hipMemcpy2DAsync(f->zvec.p, row_bytes, in, 2 * ld_b * sizeof(double), row_bytes, (size_t)chunk,
                               hipMemcpyDeviceToDevice, f->stream()),
              "D2D");
        in = f->zvec.p;
      }
      check(launch_to_block(f->stream(), in, f->zb.p, rows, chunk, ld), "to_block");
      f->zsolve(mode, chunk, ld, tolerance, (int32_t)max_steps);
      double* out = x + 2 * start * ld_x;
      if (ld_x == (uint64_t)f->n) {
        check(launch_from_block(f->stream(), f->zx.p, out, rows, chunk, ld), "from_block");
      } else {
        check(launch_from_block(f->stream(), f->zx.p, f->zvec.p, rows, chunk, ld), "from_block");
        check(hipMemcpy2DAsync(out, 2 * ld_x * sizeof(double), f->zvec.p, row_bytes, row_bytes, (size_t)chunk,
                               hipMemcpyDeviceToDevice, f->stream()),
              "D2D");
      }
    }
    check(hipStreamSynchronize(f->stream()), "solve");
  });
}

int32_t ybus_hip_factor_set_refine_stop(ybus_hip_factor* f, int32_t rule) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    if (rule != YBUS_HIP_STOP_NORMWISE && rule != YBUS_HIP_STOP_AT_FLOOR) fail(YBUS_HIP_INVALID_ARGUMENT, "unknown rule");
    f->refine_stop = rule;
  });
}

int32_t ybus_hip_factor_set_matrix_complex(ybus_hip_factor* f, int32_t row_major, const uint64_t* ptr,
                                           const uint64_t* idx, const double* val) {
  return guard([&] {
    if (!f || !ptr) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    require_complex(f);
    const int64_t n = f->n;
    std::string why;
    if (!valid_compressed(n, n, ptr, idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "A: " + why);
    const uint64_t nnz = ptr[n];
    if (nnz && (!idx || !val)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (nnz > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "the matrix is too large for 32-bit indices");
    std::vector<int32_t> rptr(n + 1), cols(nnz);
    std::vector<double2> vals(nnz);
    if (row_major) {
      for (int64_t i = 0; i <= n; ++i) rptr[i] = (int32_t)ptr[i];
      for (uint64_t p = 0; p < nnz; ++p) {
        cols[p] = (int32_t)idx[p];
        vals[p] = make_double2(val[2 * p], val[2 * p + 1]);
      }
    } else {
      // CSC to CSR by a counting sort by row: each row's columns come out ascending, the order
      // CsrMatrix::mul_vec sums them in (product.rs:17-27)
      std::vector<int32_t> cnt(n + 1, 0);
      for (uint64_t p = 0; p < nnz; ++p) cnt[idx[p] + 1] += 1;
      for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
      rptr = cnt;
      std::vector<int32_t> cursor(cnt.begin(), cnt.end() - 1);
      for (int64_t j = 0; j < n; ++j)
        for (uint64_t p = ptr[j]; p < ptr[j + 1]; ++p) {
          const int32_t q = cursor[idx[p]]++;
          cols[q] = (int32_t)j;
          vals[q] = make_double2(val[2 * p], val[2 * p + 1]);
        }
    }
    f->za.rows = (int32_t)n;
    f->za.nnz = nnz;
    f->za.ptr.upload(rptr, f->stream());
    f->za.col.upload(cols, f->stream());
    f->za.val.upload(vals, f->stream());
    check(hipStreamSynchronize(f->stream()), "upload matrix");
  });
}

int32_t ybus_hip_factor_solve_many(ybus_hip_factor* f, const double* rhs, uint64_t count, double* x) {
  return solve_many_mode(f, rhs, count, x, YBUS_HIP_FP64, 0.0, 0, nullptr, nullptr);
}

int32_t ybus_hip_factor_solve_transpose_many(ybus_hip_factor* f, const double* rhs, uint64_t count, double* x) {
  return solve_many_mode(f, rhs, count, x, YBUS_HIP_TRANSPOSE, 0.0, 0, nullptr, nullptr);
}

int32_t ybus_hip_factor_solve_many_refined(ybus_hip_factor* f, const double* rhs, uint64_t count, double tolerance,
                                           uint64_t max_steps, double* x, uint64_t* steps, double* ratio) {
  return solve_many_mode(f, rhs, count, x, YBUS_HIP_FP32_REFINED, tolerance, max_steps, steps, ratio);
}

int32_t ybus_hip_factor_adjoint(ybus_hip_factor* f, int32_t mode, double tolerance, uint64_t max_steps,
                                const double* x_bar, const double* x, uint64_t count, double* b_bar, double* a_bar) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (!f->a.ptr.p) fail(YBUS_HIP_INVALID_ARGUMENT, "the adjoint needs the matrix (ybus_hip_factor_set_matrix)");
    const int64_t nnz = (int64_t)f->a.nnz;
    if (count && (!x_bar || !x || !b_bar)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (nnz && !a_bar) fail(YBUS_HIP_INVALID_ARGUMENT, "a_bar is null");
    if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_FP32_REFINED) fail(YBUS_HIP_INVALID_ARGUMENT, "mode: FP64 or FP32_REFINED");
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    const bool refined = mode == YBUS_HIP_FP32_REFINED;
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    std::vector<double> acc((size_t)nnz, 0.0);
    if (count && f->n) {
      const int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
      f->ensure_workspace(width, refined);
      f->abar.reserve((size_t)std::max<int64_t>(nnz, 1));
      const uint64_t n = (uint64_t)f->n;
      for (uint64_t start = 0; start < count; start += (uint64_t)width) {
        const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
        // b_bar = A^-T x_bar (the transposed solve, lu.rs:502-533; A^T = A for a Cholesky factor)
        upload_vectors(f, x_bar + start * n, chunk);
        if (refined)
          f->refine(chunk, tolerance, (int32_t)max_steps, true);
        else
          f->solve64(true, false, chunk);
        download_vectors(f, f->x.p, f->layout, chunk, b_bar + start * n);
        // the solutions x beside it, then the running sums on A's pattern
        upload_vectors(f, x + start * n, chunk, f->y2.p);
        f->begin_kernel();
        check(launch_sddmm(f->stream(), f->a.row.p, f->a.col.p, nnz, f->x.p, f->y2.p, f->layout.rs(),
                           f->layout.cs(), (int32_t)chunk, f->abar.p, start == 0),
              "adjoint kernel");
        // both blocks read once and the running sums read and written: 16 n c + 16 nnz(A)
        f->end_kernel(16 * (uint64_t)f->n * (uint64_t)chunk + 16 * (uint64_t)nnz);
      }
      if (nnz) {
        check(hipMemcpyAsync(acc.data(), f->abar.p, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost,
                             f->stream()),
              "D2H");
        check(hipStreamSynchronize(f->stream()), "adjoint");
      }
    }
    // A_bar = -(sum over the vectors of b_bar_s x_s^T), in the caller's storage order
    for (int64_t q = 0; q < nnz; ++q) a_bar[f->a.order[q]] = -acc[q];
  });
}

static void check_device_vectors(const ybus_hip_factor* f, uint64_t count, uint64_t ld,
                                 std::initializer_list<const void*> pointers) {
  if (count == 0) return;
  for (const void* p : pointers)
    if (!p) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  if (ld < (uint64_t)f->n) fail(YBUS_HIP_INVALID_ARGUMENT, "the vectors' stride is below dim()");
  if (count > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "too many vectors");
}

int32_t ybus_hip_factor_solve_device(ybus_hip_factor* f, int32_t mode, double tolerance, uint64_t max_steps,
                                     const double* b, uint64_t ld_b, uint64_t count, double* x, uint64_t ld_x) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_TRANSPOSE && mode != YBUS_HIP_FP32_REFINED)
      fail(YBUS_HIP_INVALID_ARGUMENT, "unknown mode");
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    const bool refined = mode == YBUS_HIP_FP32_REFINED;
    if (refined && !f->a.ptr.p)
      fail(YBUS_HIP_INVALID_ARGUMENT, "FP32 refinement needs the matrix (ybus_hip_factor_set_matrix)");
    check_device_vectors(f, count, std::min(ld_b, ld_x), {b, x});
    if (count == 0 || f->n == 0) return;
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    const int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
    f->ensure_workspace(width, refined);
    for (uint64_t start = 0; start < count; start += (uint64_t)width) {
      const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
      copy_in_device(f, b + start * ld_b, (int64_t)ld_b, chunk, f->b.p);
      if (refined)
        f->refine(chunk, tolerance, (int32_t)max_steps);
      else
        f->solve64(mode == YBUS_HIP_TRANSPOSE, false, chunk);
      copy_out_device(f, f->x.p, chunk, x + start * ld_x, (int64_t)ld_x);
    }
    check(hipStreamSynchronize(f->stream()), "solve");
  });
}

static_assert(SCRATCH_BYTES == YBUS_HIP_SCRATCH_BYTES, "the header states the workspace's size");

// A workspace buffer of the factor lent a slice of the no-allocation workspace for one call, and
// given back its own memory (and length) when the call ends, however it ends
struct Lend {
  DevBuf<double>& buf;
  double* p;
  size_t len;
  bool lent;
  Lend(DevBuf<double>& b, double* slice, size_t elems) : buf(b), p(b.p), len(b.len), lent(b.lent) {
    buf.p = slice;
    buf.len = elems;
    buf.lent = true;
  }
  ~Lend() {
    buf.p = p;
    buf.len = len;
    buf.lent = lent;
  }
};

int32_t ybus_hip_factor_solve_device_noalloc(ybus_hip_factor* f, int32_t mode, const double* b, uint64_t ld_b,
                                             uint64_t count, double* x, uint64_t ld_x) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_TRANSPOSE)
      fail(YBUS_HIP_INVALID_ARGUMENT, "mode: YBUS_HIP_FP64 or YBUS_HIP_TRANSPOSE");
    check_device_vectors(f, count, std::min(ld_b, ld_x), {b, x});
    if (count == 0 || f->n == 0) return;
    if (!f->ctx->scratch) fail(YBUS_HIP_INTERNAL, "the context has no no-allocation workspace");
    const size_t n = (size_t)f->n;
    const size_t align = 256 / sizeof(double);
    auto up = [&](size_t v) { return (v + align - 1) / align * align; };
    // the widest chunk (at most block_width() and count columns) whose buffers fit the array: the
    // column path keeps b and x (n per column each); the wide path b (n x ld), t and x (each with
    // the pad row) and the vectors' staging (n per column)
    Layout layout;
    size_t need = 0;
    int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
    for (;; width = width / 2) {
      if (width == 0)
        fail(YBUS_HIP_TOO_LARGE, "one vector's workspace exceeds the no-allocation workspace of " +
                                     std::to_string(SCRATCH_BYTES >> 20) + " MiB");
      layout.columns = f->use_columns(width, false);
      layout.ld = layout.columns ? (int64_t)n : round_up(width, 16);
      const size_t elems = n * (size_t)(layout.columns ? width : layout.ld);
      need = layout.columns ? 2 * up(elems) : up(elems) + 2 * up(elems + (size_t)layout.ld) + up(n * (size_t)width);
      if (need * sizeof(double) <= SCRATCH_BYTES) break;
    }
    const size_t elems = n * (size_t)(layout.columns ? width : layout.ld);
    const size_t padded = layout.columns ? elems : elems + (size_t)layout.ld;
    double* slice_b = static_cast<double*>(f->ctx->scratch);
    double* slice_x = slice_b + up(elems);
    double* slice_t = slice_x + up(padded);
    double* slice_v = slice_t + up(padded);
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    // the factor's own workspace (and what it has staged) is left as it is: its buffers are lent
    // the slices for this call and its layout is put back after it
    Lend lb(f->b, slice_b, elems), lx(f->x, slice_x, padded);
    std::optional<Lend> lt, lv;
    if (!layout.columns) {
      lt.emplace(f->t, slice_t, padded);
      lv.emplace(f->vec, slice_v, n * (size_t)width);
    }
    const Layout own = f->layout;
    struct Restore {
      ybus_hip_factor* f;
      Layout layout;
      ~Restore() {
        f->layout = layout;
        f->no_alloc = false;
      }
    } restore{f, own};
    f->layout = layout;
    f->no_alloc = true;
    for (uint64_t start = 0; start < count; start += (uint64_t)width) {
      const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
      copy_in_device(f, b + start * ld_b, (int64_t)ld_b, chunk, f->b.p);
      f->solve64(mode == YBUS_HIP_TRANSPOSE, false, chunk);
      copy_out_device(f, f->x.p, chunk, x + start * ld_x, (int64_t)ld_x);
    }
    check(hipStreamSynchronize(f->stream()), "solve");
  });
}

int32_t ybus_hip_factor_adjoint_device(ybus_hip_factor* f, int32_t mode, double tolerance, uint64_t max_steps,
                                       const double* x_bar, const double* x, uint64_t ld, uint64_t count,
                                       double* b_bar, double* a_bar) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (!f->a.ptr.p) fail(YBUS_HIP_INVALID_ARGUMENT, "the adjoint needs the matrix (ybus_hip_factor_set_matrix)");
    const int64_t nnz = (int64_t)f->a.nnz;
    check_device_vectors(f, count, ld, {x_bar, x, b_bar});
    if (nnz && !a_bar) fail(YBUS_HIP_INVALID_ARGUMENT, "a_bar is null");
    if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_FP32_REFINED) fail(YBUS_HIP_INVALID_ARGUMENT, "mode: FP64 or FP32_REFINED");
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    const bool refined = mode == YBUS_HIP_FP32_REFINED;
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    f->abar.reserve((size_t)std::max<int64_t>(nnz, 1));
    if (count && f->n) {
      const int64_t width = (int64_t)std::min<uint64_t>(f->width, count);
      f->ensure_workspace(width, refined);
      for (uint64_t start = 0; start < count; start += (uint64_t)width) {
        const int64_t chunk = (int64_t)std::min<uint64_t>((uint64_t)width, count - start);
        copy_in_device(f, x_bar + start * ld, (int64_t)ld, chunk, f->b.p);
        if (refined)
          f->refine(chunk, tolerance, (int32_t)max_steps, true);
        else
          f->solve64(true, false, chunk);
        copy_out_device(f, f->x.p, chunk, b_bar + start * ld, (int64_t)ld);
        copy_in_device(f, x + start * ld, (int64_t)ld, chunk, f->y2.p);
        f->begin_kernel();
        check(launch_sddmm(f->stream(), f->a.row.p, f->a.col.p, nnz, f->x.p, f->y2.p, f->layout.rs(),
                           f->layout.cs(), (int32_t)chunk, f->abar.p, start == 0),
              "adjoint kernel");
        f->end_kernel(16 * (uint64_t)f->n * (uint64_t)chunk + 16 * (uint64_t)nnz);
      }
    }
    check(launch_scatter_negated(f->stream(), f->abar.p, f->a.order_dev.p, nnz, count && f->n, a_bar),
          "adjoint scatter");
    check(hipStreamSynchronize(f->stream()), "adjoint");
  });
}

int32_t ybus_hip_factor_stage(ybus_hip_factor* f, const double* rhs, uint64_t count) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (count > f->width) fail(YBUS_HIP_INVALID_ARGUMENT, "more right-hand sides than block_width()");
    if (count && !rhs) fail(YBUS_HIP_INVALID_ARGUMENT, "rhs is null");
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    f->ensure_workspace((int64_t)count, true);
    if (count && f->n) upload_vectors(f, rhs, (int64_t)count);
    f->staged = (int64_t)count;
    f->staged_pairs = false;
  });
}

// pairs.rs:153-187: checks the pairs and copies them to the device (f->plus, f->minus)
static void stage_pairs(ybus_hip_factor* f, const int64_t* plus, const int64_t* minus, uint64_t k) {
  require_unborrowed(f);
  if (k > f->width)
    fail(YBUS_HIP_INVALID_ARGUMENT,
         "at most block_width() (" + std::to_string(f->width) + ") pairs, not " + std::to_string(k));
  if (k && (!plus || !minus)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
  std::vector<int32_t> p(k), m(k);
  for (uint64_t c = 0; c < k; ++c) {
    if (plus[c] < -1 || plus[c] >= f->n || minus[c] < -1 || minus[c] >= f->n)
      fail(YBUS_HIP_INVALID_ARGUMENT, "unknowns below dim() (" + std::to_string(f->n) + ")");
    if (plus[c] >= 0 && plus[c] == minus[c]) fail(YBUS_HIP_INVALID_ARGUMENT, "a pair's two unknowns must differ");
    p[c] = (int32_t)plus[c];
    m[c] = (int32_t)minus[c];
  }
  check(hipSetDevice(f->ctx->device), "hipSetDevice");
  f->ensure_workspace((int64_t)k, false, true);
  if (k) {
    check(hipMemcpyAsync(f->plus.p, p.data(), k * sizeof(int32_t), hipMemcpyHostToDevice, f->stream()), "H2D");
    check(hipMemcpyAsync(f->minus.p, m.data(), k * sizeof(int32_t), hipMemcpyHostToDevice, f->stream()), "H2D");
    check(hipStreamSynchronize(f->stream()), "stage pairs");
    // the wide path sweeps only each tile's reach forward, where its LDS bitmap fits
    if (!f->layout.columns && f->use_reach && f->n &&
        reach_lds_bytes(f->fwd.view(), f->bwd.view()) <= f->lds_cap)
      f->stage_reach(p, m, f->tile_for((int64_t)k, true), f->use_parts((int64_t)k, true));
  }
}

int32_t ybus_hip_factor_stage_pairs(ybus_hip_factor* f, const int64_t* plus, const int64_t* minus, uint64_t k) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    stage_pairs(f, plus, minus, k);
    f->staged = (int64_t)k;
    f->staged_pairs = true;
  });
}

int32_t ybus_hip_factor_autotune(ybus_hip_factor* f, uint64_t columns, int32_t pairs, const int64_t* tune_plus,
                                 const int64_t* tune_minus, ybus_hip_tuning* out) {
  return guard([&] {
    if (!f) fail(YBUS_HIP_INVALID_ARGUMENT, "factor is null");
    require_unborrowed(f);
    if (columns == 0 || columns > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_INVALID_ARGUMENT, "columns out of range");
    if (f->n < 2) fail(YBUS_HIP_INVALID_ARGUMENT, "a factor of dim() >= 2 (pairs need two unknowns)");
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    const int64_t k = (int64_t)std::min<uint64_t>(columns, f->width);
    // a synthetic block of that width: pairs with +1 at c mod n and -1 at (7c + 1) mod n in
    // column c, or vectors of values in [-1, 1) from a fixed linear congruential sequence
    std::vector<int64_t> plus, minus;
    std::vector<double> vectors;
    if (pairs && tune_plus && tune_minus) {
      // the caller's own pairs (checked by stage_pairs like any other)
      plus.assign(tune_plus, tune_plus + k);
      minus.assign(tune_minus, tune_minus + k);
    } else if (pairs) {
      plus.resize(k);
      minus.resize(k);
      for (int64_t c = 0; c < k; ++c) {
        plus[c] = c % f->n;
        minus[c] = (7 * c + 1) % f->n;
        if (minus[c] == plus[c]) minus[c] = (plus[c] + 1) % f->n;
      }
    } else {
      vectors.resize((size_t)k * (size_t)f->n);
      uint64_t state = 0x9e3779b97f4a7c15ull;
      for (double& v : vectors) {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        v = (double)(state >> 11) * 0x1.0p-52 - 1.0;
      }
    }
    std::vector<ybus_hip_factor::Tuned> candidates;
    if (f->fits && f->force != 1)
      for (int th : {256, 512}) candidates.push_back({true, th, 0});
    if (f->force != 2)
      for (int parts : {0, 1})
        for (int th : {256, 512, 1024})
          for (int tile : {1, 2, 4, 8, 16, 32, 64})
            if ((tile == 1 || tile / 2 < k) && (parts == 0 || f->partitioned))
              candidates.push_back({false, th, tile, parts});
    if (candidates.empty()) fail(YBUS_HIP_INVALID_ARGUMENT, "no kernel configuration to try");
    const int key = 2 * ybus_hip_factor::bucket(k) + (pairs ? 1 : 0);
    f->tuned.erase(key);
    ybus_hip_factor::Tuned best{};
    double best_ms = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    const ybus_hip_kernel_stats saved = f->stats;
    // the best of three timed solves after a warm one
    auto time_config = [&](const ybus_hip_factor::Tuned& c) {
      f->trial = &c;
      if (pairs) {
        stage_pairs(f, plus.data(), minus.data(), (uint64_t)k);
      } else {
        f->ensure_workspace(k, false);
        upload_vectors(f, vectors.data(), k);
      }
      f->solve64(false, pairs != 0, k);  // warm
      double ms = 0.0;
      for (int rep = 0; rep < 3; ++rep) {
        check(hipEventRecord(e0, f->stream()), "hipEventRecord");
        f->solve64(false, pairs != 0, k);
        check(hipEventRecord(e1, f->stream()), "hipEventRecord");
        check(hipEventSynchronize(e1), "autotune");
        float t = 0.0f;
        check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        ms = rep == 0 ? t : std::min(ms, (double)t);
      }
      return ms;
    };
    try {
      for (const ybus_hip_factor::Tuned& c : candidates) {
        const double ms = time_config(c);
        if (best_ms == 0.0 || ms < best_ms) {
          best = c;
          best_ms = ms;
        }
      }
      // a partitioned block of vectors: the top of the tree (few, long levels) on its own
      // configuration, the parts' kept
      if (!pairs && !best.columns && best.parts == 1) {
        const ybus_hip_factor::Tuned parts_best = best;
        for (int th : {64, 128, 256, 512, 1024})
          for (int tile : {1, 2, 4, 8, 16, 32, 64}) {
            if (th < tile || (tile != 1 && tile / 2 >= k) || (tile == parts_best.tile && th == parts_best.threads))
              continue;
            ybus_hip_factor::Tuned c = parts_best;
            c.top_tile = tile;
            c.top_threads = th;
            const double ms = time_config(c);
            if (ms < best_ms) {
              best = c;
              best_ms = ms;
            }
          }
        // then the parts in LDS slots, at the tile widths whose slots fit, the top's kept
        if (f->parts_lds.parts > 0) {
          const ybus_hip_factor::Tuned top_best = best;
          for (int tile : {1, 2, 4, 8, 16})
            for (int th : {256, 512, 1024}) {
              if (th < 64 * tile || (tile != 1 && tile / 2 >= k) ||
                  parts_lds_layout(f->parts_lds.fwd.dev, tile).total > f->lds_cap ||
                  parts_lds_layout(f->parts_lds.bwd.dev, tile).total > f->lds_cap)
                continue;
              ybus_hip_factor::Tuned c = top_best;
              c.parts_lds_tile = tile;
              c.parts_lds_threads = th;
              const double ms = time_config(c);
              if (ms < best_ms) {
                best = c;
                best_ms = ms;
              }
            }
        }
      }
      // a one-schedule wide block: the solve with the last positions in LDS, at each tile width
      // and workgroup size (not for a pair block whose reach is staged: it takes the reach
      // kernel at every tile width, so a windowed candidate would time that kernel again)
      const bool reach_staged =
          pairs && f->use_reach && reach_lds_bytes(f->fwd.view(), f->bwd.view()) <= f->lds_cap;
      if (!best.columns && best.parts != 1 && f->window_env < 0 && !reach_staged) {
        const ybus_hip_factor::Tuned wide_best = best;
        for (int th : {256, 512})
          for (int tile : {1, 2, 4, 8, 16}) {
            if (th < tile || (tile != 1 && tile / 2 >= k)) continue;
            ybus_hip_factor::Tuned c = wide_best;
            c.window = true;
            c.tile = tile;
            c.threads = th;
            c.parts = 0;
            const double ms = time_config(c);
            if (ms < best_ms) {
              best = c;
              best_ms = ms;
            }
          }
      }
    } catch (...) {
      f->trial = nullptr;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    f->trial = nullptr;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f->stats = saved;
    f->tuned[key] = best;
    f->ensure_workspace(k, false);  // drops the synthetic staging
    if (out)
      *out = ybus_hip_tuning{best.columns ? YBUS_HIP_PATH_COLUMNS : YBUS_HIP_PATH_WIDE,
                             best.threads,
                             best.columns ? 1 : best.tile,
                             best_ms,
                             best.columns ? 0 : std::max(best.parts, 0),
                             best.top_tile ? best.top_tile : (best.columns ? 0 : best.tile),
                             best.top_tile ? best.top_threads : (best.columns ? 0 : best.threads),
                             best.parts_lds_tile,
                             best.parts_lds_threads,
                             best.window ? 1 : 0};
  });
}

int32_t ybus_hip_factor_solve_staged(ybus_hip_factor* f, int32_t mode, double tolerance, uint64_t max_steps,
                                     ybus_hip_block** out) {
  return guard([&] {
    if (!f || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    require_unborrowed(f);
    if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_FP32_REFINED && mode != YBUS_HIP_TRANSPOSE)
      fail(YBUS_HIP_INVALID_ARGUMENT, "unknown mode");
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    if (f->staged_pairs && mode != YBUS_HIP_FP64)
      fail(YBUS_HIP_INVALID_ARGUMENT, "staged pairs are solved in YBUS_HIP_FP64 only");
    auto block = new_block(f, f->staged, f->x.p);
    if (f->staged && f->n) {
      check(hipSetDevice(f->ctx->device), "hipSetDevice");
      if (mode == YBUS_HIP_FP32_REFINED) {
        f->refine(f->staged, tolerance, (int32_t)max_steps);
        read_refinement(f, f->staged, block->steps.data(), block->ratio.data());
      } else {
        f->solve64(mode == YBUS_HIP_TRANSPOSE, f->staged_pairs, f->staged);
      }
    }
    f->borrowed = block.get();
    *out = block.release();
  });
}

int32_t ybus_hip_factor_pair_solves(ybus_hip_factor* f, const int64_t* plus, const int64_t* minus, uint64_t k,
                                    ybus_hip_block** out) {
  return guard([&] {
    if (!f || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    stage_pairs(f, plus, minus, k);
    if (k && f->n) f->solve64(false, true, (int64_t)k);
    auto block = new_block(f, (int64_t)k, f->x.p);
    f->borrowed = block.get();
    *out = block.release();
  });
}

int32_t ybus_hip_factor_pair_updates(ybus_hip_factor* f, const double* base, const int64_t* plus,
                                     const int64_t* minus, const double* alpha, const double* inverse_weight,
                                     uint64_t k, ybus_hip_block** out) {
  return guard([&] {
    if (!f || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (!base && f->n) fail(YBUS_HIP_INVALID_ARGUMENT, "one base value per unknown");
    if (k && (!alpha || !inverse_weight)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    stage_pairs(f, plus, minus, k);
    if (k && f->n) f->solve64(false, true, (int64_t)k);
    auto block = new_block(f, (int64_t)k, f->y2.p);
    if (k && f->n) {
      f->upd_base.reserve(f->n);
      f->upd_alpha.reserve(k);
      f->upd_weight.reserve(k);
      f->coef.reserve(k);
      check(hipMemcpyAsync(f->upd_base.p, base, f->n * sizeof(double), hipMemcpyHostToDevice, f->stream()), "H2D");
      check(hipMemcpyAsync(f->upd_alpha.p, alpha, k * sizeof(double), hipMemcpyHostToDevice, f->stream()), "H2D");
      check(hipMemcpyAsync(f->upd_weight.p, inverse_weight, k * sizeof(double), hipMemcpyHostToDevice, f->stream()),
            "H2D");
      check(launch_pair_update(f->stream(), f->x.p, f->y2.p, (int32_t)f->n, (int32_t)k, f->layout.rs(),
                               f->layout.cs(), f->upd_base.p, f->plus.p, f->minus.p, f->upd_alpha.p, f->upd_weight.p,
                               f->singular.p, f->coef.p),
            "pair_update");
      std::vector<int32_t> s(k);
      check(hipMemcpyAsync(s.data(), f->singular.p, k * sizeof(int32_t), hipMemcpyDeviceToHost, f->stream()), "D2H");
      check(hipStreamSynchronize(f->stream()), "pair_update");
      for (uint64_t c = 0; c < k; ++c) block->singular[c] = s[c] != 0;
    }
    f->borrowed = block.get();
    *out = block.release();
  });
}

// Solves the m x m system w z = c (row-major w, both overwritten; z in c) by Gaussian
// elimination with partial pivoting (the row of largest |value| below the diagonal, the first on
// a tie); false when a pivot is zero or the result is not finite
static bool dense_solve(int32_t m, double* w, double* c) {
  for (int32_t j = 0; j < m; ++j) {
    int32_t piv = j;
    for (int32_t i = j + 1; i < m; ++i)
      if (std::fabs(w[i * m + j]) > std::fabs(w[piv * m + j])) piv = i;
    if (w[piv * m + j] == 0.0) return false;
    if (piv != j) {
      for (int32_t q = 0; q < m; ++q) std::swap(w[j * m + q], w[piv * m + q]);
      std::swap(c[j], c[piv]);
    }
    for (int32_t i = j + 1; i < m; ++i) {
      const double l = w[i * m + j] / w[j * m + j];
      for (int32_t q = j + 1; q < m; ++q) w[i * m + q] = w[i * m + q] - l * w[j * m + q];
      c[i] = c[i] - l * c[j];
    }
  }
  for (int32_t j = m; j-- > 0;) {
    double v = c[j];
    for (int32_t q = j + 1; q < m; ++q) v = v - w[j * m + q] * c[q];
    c[j] = v / w[j * m + j];
    if (!std::isfinite(c[j])) return false;
  }
  return true;
}

int32_t ybus_hip_factor_outage_updates(ybus_hip_factor* f, const double* base, const uint64_t* set_ptr, uint64_t sets,
                                       const int64_t* plus, const int64_t* minus, const double* alpha,
                                       const double* inverse_weight, ybus_hip_block** out) {
  return guard([&] {
    if (!f || !out || !set_ptr) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (!base && f->n) fail(YBUS_HIP_INVALID_ARGUMENT, "one base value per unknown");
    if (set_ptr[0] != 0) fail(YBUS_HIP_INVALID_ARGUMENT, "set_ptr[0] must be 0");
    for (uint64_t s = 0; s < sets; ++s) {
      if (set_ptr[s + 1] < set_ptr[s]) fail(YBUS_HIP_INVALID_ARGUMENT, "set_ptr must not decrease");
      if (set_ptr[s + 1] - set_ptr[s] > (uint64_t)MAX_OUTAGE_SET)
        fail(YBUS_HIP_INVALID_ARGUMENT, "at most " + std::to_string(MAX_OUTAGE_SET) + " branches a set");
    }
    const uint64_t k = set_ptr[sets];
    if (k && (!alpha || !inverse_weight)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    stage_pairs(f, plus, minus, k);
    if (k && f->n) f->solve64(false, true, (int64_t)k);
    auto block = new_block(f, (int64_t)sets, f->y2.p);
    if (sets && f->n) {
      // the sets' systems on the device (W and C^T x), solved here, applied on the device
      std::vector<int32_t> sp(sets + 1), w_at(sets);
      size_t w_len = 0;
      for (uint64_t s = 0; s <= sets; ++s) sp[s] = (int32_t)set_ptr[s];
      for (uint64_t s = 0; s < sets; ++s) {
        w_at[s] = (int32_t)w_len;
        w_len += (size_t)(sp[s + 1] - sp[s]) * (size_t)(sp[s + 1] - sp[s]);
      }
      if (w_len > (size_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "too many outage sets");
      f->upd_base.reserve(f->n);
      f->upd_alpha.reserve(std::max<uint64_t>(k, 1));
      f->upd_weight.reserve(std::max<uint64_t>(k, 1));
      f->coef.reserve(std::max<uint64_t>(k, 1));
      DevBuf<int32_t> d_sp, d_at, d_singular;
      DevBuf<double> d_w, d_ctx;
      d_sp.upload(sp, f->stream());
      d_at.upload(w_at, f->stream());
      d_w.reserve(std::max<size_t>(w_len, 1));
      d_ctx.reserve(std::max<uint64_t>(k, 1));
      check(hipMemcpyAsync(f->upd_base.p, base, f->n * sizeof(double), hipMemcpyHostToDevice, f->stream()), "H2D");
      if (k) {
        check(hipMemcpyAsync(f->upd_alpha.p, alpha, k * sizeof(double), hipMemcpyHostToDevice, f->stream()), "H2D");
        check(hipMemcpyAsync(f->upd_weight.p, inverse_weight, k * sizeof(double), hipMemcpyHostToDevice, f->stream()),
              "H2D");
      }
      check(launch_outage_system(f->stream(), f->x.p, f->layout.rs(), f->layout.cs(), f->upd_base.p, d_sp.p,
                                 (int32_t)sets, f->plus.p, f->minus.p, f->upd_alpha.p, f->upd_weight.p, d_at.p, d_w.p,
                                 d_ctx.p),
            "outage system");
      std::vector<double> w(w_len), z(k);
      if (w_len)
        check(hipMemcpyAsync(w.data(), d_w.p, w_len * sizeof(double), hipMemcpyDeviceToHost, f->stream()), "D2H");
      if (k) check(hipMemcpyAsync(z.data(), d_ctx.p, k * sizeof(double), hipMemcpyDeviceToHost, f->stream()), "D2H");
      check(hipStreamSynchronize(f->stream()), "outage system");
      std::vector<int32_t> singular(sets, 0);
      for (uint64_t s = 0; s < sets; ++s) {
        const int32_t m = sp[s + 1] - sp[s];
        if (m && !dense_solve(m, w.data() + w_at[s], z.data() + sp[s])) singular[s] = 1;
      }
      d_singular.upload(singular, f->stream());
      if (k) check(hipMemcpyAsync(f->coef.p, z.data(), k * sizeof(double), hipMemcpyHostToDevice, f->stream()), "H2D");
      check(launch_outage_apply(f->stream(), f->x.p, f->layout.rs(), f->layout.cs(), (int32_t)f->n, f->upd_base.p,
                                d_sp.p, f->upd_alpha.p, f->coef.p, d_singular.p, f->y2.p, f->layout.rs(),
                                f->layout.cs(), (int32_t)sets),
            "outage apply");
      check(hipStreamSynchronize(f->stream()), "outage apply");
      for (uint64_t s = 0; s < sets; ++s) block->singular[s] = singular[s] != 0;
    }
    f->borrowed = block.get();
    *out = block.release();
  });
}

void ybus_hip_block_free(ybus_hip_block* b) {
  if (!b) return;
  if (b->f && b->f->borrowed == b) b->f->borrowed = nullptr;
  delete b;
}

uint64_t ybus_hip_block_columns(const ybus_hip_block* b) { return b ? (uint64_t)b->columns : 0; }

int32_t ybus_hip_block_singular(const ybus_hip_block* b, uint8_t* out) {
  return guard([&] {
    if (!b || (!out && b->columns)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    std::copy(b->singular.begin(), b->singular.end(), out);
  });
}

int32_t ybus_hip_block_refinement(const ybus_hip_block* b, uint64_t* steps, double* ratio) {
  return guard([&] {
    if (!b) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (steps) std::copy(b->steps.begin(), b->steps.end(), steps);
    if (ratio) std::copy(b->ratio.begin(), b->ratio.end(), ratio);
  });
}

int32_t ybus_hip_block_solutions(ybus_hip_block* b, double* out) {
  return guard([&] {
    if (!b || (!out && b->columns)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (!b->columns || !b->f->n) return;
    check(hipSetDevice(b->f->ctx->device), "hipSetDevice");
    download_vectors(b->f, b->data, b->layout, b->columns, out);
  });
}

// The factor's buffers for the maxima of `columns` columns over `rows` rows: the partial results
// of a row-interleaved block's chunks, and the final ones; returns the chunk count
static int32_t reserve_maxima(ybus_hip_factor* f, int32_t rows, int64_t columns) {
  const int32_t chunks = max_chunks(rows, (int32_t)columns, f->ctx->cus);
  const size_t partial = (size_t)chunks * (size_t)columns;
  f->max_pv.reserve(partial);
  f->max_pr.reserve(partial);
  f->max_pb.reserve(partial);
  f->max_best.reserve(columns);
  f->max_row.reserve(columns);
  f->max_bad.reserve(columns);
  return chunks;
}

// Copies the final maxima of `columns` columns to the host
static void read_maxima(ybus_hip_factor* f, int64_t columns, ybus_hip_column_max* out) {
  std::vector<double> hb(columns);
  std::vector<int32_t> hr(columns), hn(columns);
  check(hipMemcpyAsync(hb.data(), f->max_best.p, columns * sizeof(double), hipMemcpyDeviceToHost, f->stream()), "D2H");
  check(hipMemcpyAsync(hr.data(), f->max_row.p, columns * sizeof(int32_t), hipMemcpyDeviceToHost, f->stream()), "D2H");
  check(hipMemcpyAsync(hn.data(), f->max_bad.p, columns * sizeof(int32_t), hipMemcpyDeviceToHost, f->stream()), "D2H");
  check(hipStreamSynchronize(f->stream()), "column_max");
  for (int64_t c = 0; c < columns; ++c) out[c] = ybus_hip_column_max{hr[c], hb[c], hn[c]};
}

static void maxima_of(ybus_hip_factor* f, const double* block, int32_t rows, int64_t rs, int64_t cs,
                      const double* weight, int64_t columns, ybus_hip_column_max* out) {
  const int32_t chunks = reserve_maxima(f, rows, columns);
  check(launch_column_max(f->stream(), block, rows, (int32_t)columns, rs, cs, weight, chunks, f->max_pv.p,
                          f->max_pr.p, f->max_pb.p, f->max_best.p, f->max_row.p, f->max_bad.p),
        "column_max");
  read_maxima(f, columns, out);
}

int32_t ybus_hip_block_maxima(ybus_hip_block* b, ybus_hip_column_max* out) {
  return guard([&] {
    if (!b || (!out && b->columns)) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    if (!b->columns) return;
    ybus_hip_factor* f = b->f;
    if (!f->n) {
      for (int64_t c = 0; c < b->columns; ++c) out[c] = ybus_hip_column_max{-1, 0.0, 0};
      return;
    }
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    if (f->ones.len < (size_t)f->n) f->ones.upload(std::vector<double>(f->n, 1.0), f->stream());
    maxima_of(f, b->data, (int32_t)f->n, b->layout.rs(), b->layout.cs(), f->ones.p, b->columns, out);
  });
}

// Checks each column's row to zero (-1 for none) and copies them to the device
static void stage_zero(ybus_hip_block* b, ybus_hip_operator* op, const int64_t* zero) {
  ybus_hip_factor* f = b->f;
  if (op->f != f) fail(YBUS_HIP_INVALID_ARGUMENT, "an operator of this factorization");
  if (b->columns && !zero) fail(YBUS_HIP_INVALID_ARGUMENT, "one entry to zero (or none) per column");
  std::vector<int32_t> z(b->columns);
  for (int64_t c = 0; c < b->columns; ++c) {
    if (zero[c] < -1 || zero[c] >= op->m.rows)
      fail(YBUS_HIP_INVALID_ARGUMENT, "rows to zero below the operator's rows");
    z[c] = (int32_t)zero[c];
  }
  check(hipSetDevice(f->ctx->device), "hipSetDevice");
  op->zero.upload(z, f->stream());
}

// pairs.rs:312-366: M theta + offset per column, then the zeroed entries, left in op->out in the
// block's layout (column-major blocks: one vector of rows per column; wide: rows x columns)
static void apply(ybus_hip_block* b, ybus_hip_operator* op, const int64_t* zero) {
  ybus_hip_factor* f = b->f;
  stage_zero(b, op, zero);
  op->out.reserve((size_t)std::max<int32_t>(op->m.rows, 1) * (size_t)std::max<int64_t>(b->columns, 1));
  op->out_cols = b->columns;
  if (b->columns && op->m.rows) {
    const int64_t ors = b->layout.columns ? 1 : b->columns, ocs = b->layout.columns ? op->m.rows : 1;
    check(launch_operator(f->stream(), op->m.view(), b->data, b->layout.rs(), b->layout.cs(), op->offset.p,
                          op->zero.p, op->out.p, ors, ocs, (int32_t)b->columns),
          "operator");
  }
}

int32_t ybus_hip_block_product(ybus_hip_block* b, ybus_hip_operator* op, const int64_t* zero, double* out) {
  return guard([&] {
    if (!b || !op) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    apply(b, op, zero);
    const int64_t rows = op->m.rows, cols = b->columns;
    if (!rows || !cols) return;
    if (!out) fail(YBUS_HIP_INVALID_ARGUMENT, "out is null");
    // the caller gets one vector of rows per column
    if (b->layout.columns) {
      b->f->staging.to_host(b->f->stream(), out, op->out.p, (size_t)rows * cols * sizeof(double));
      return;
    }
    // rows x columns on the device: transposed there into one vector per column, then copied out
    op->by_column.reserve((size_t)rows * (size_t)cols);
    check(launch_from_block(b->f->stream(), op->out.p, op->by_column.p, rows, cols, cols), "from_block");
    b->f->staging.to_host(b->f->stream(), out, op->by_column.p, (size_t)rows * cols * sizeof(double));
  });
}

int32_t ybus_hip_block_product_maxima(ybus_hip_block* b, ybus_hip_operator* op, const int64_t* zero,
                                      ybus_hip_column_max* out) {
  return guard([&] {
    if (!b || !op) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    stage_zero(b, op, zero);
    if (!b->columns) return;
    if (!out) fail(YBUS_HIP_INVALID_ARGUMENT, "out is null");
    if (!op->m.rows) {
      for (int64_t c = 0; c < b->columns; ++c) out[c] = ybus_hip_column_max{-1, 0.0, 0};
      return;
    }
    // the product's maxima computed as it is formed, without storing it
    ybus_hip_factor* f = b->f;
    const int32_t chunks = reserve_maxima(f, op->m.rows, b->columns);
    check(launch_product_max(f->stream(), op->m.view(), b->data, b->layout.rs(), b->layout.cs(), op->offset.p,
                             op->zero.p, op->weight.p, (int32_t)b->columns, chunks, f->max_pv.p, f->max_pr.p,
                             f->max_pb.p, f->max_best.p, f->max_row.p, f->max_bad.p),
          "product_max");
    read_maxima(f, b->columns, out);
  });
}

static void copy_levels(const RowForm& form, int32_t* out) {
  if (!out) return;
  std::vector<int32_t> level = levels_of(form);
  std::copy(level.begin(), level.end(), out);
}

int32_t ybus_hip_analyze_cholesky(uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, int32_t* forward_level,
                                  int32_t* backward_level, uint64_t* supernode_start, uint64_t* supernodes) {
  return guard([&] {
    if (!l_ptr || (!l_idx && l_ptr[n])) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    std::string why;
    if (!valid_compressed((int64_t)n, (int64_t)n, l_ptr, l_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "L: " + why);
    for (uint64_t i = 0; i < n; ++i)
      if (l_ptr[i + 1] == l_ptr[i] || l_idx[l_ptr[i + 1] - 1] != i)
        fail(YBUS_HIP_INVALID_ARGUMENT, "each row of L must end with its diagonal");
    std::vector<double> zeros(l_ptr[n], 0.0);
    copy_levels(cholesky_forward((int64_t)n, l_ptr, l_idx, zeros.data()), forward_level);
    copy_levels(cholesky_backward((int64_t)n, l_ptr, l_idx, zeros.data()), backward_level);
    Supernodes s = cholesky_supernodes((int64_t)n, l_ptr, l_idx);
    if (supernode_start) std::copy(s.start.begin(), s.start.end(), supernode_start);
    if (supernodes) *supernodes = s.start.size() - 1;
  });
}

int32_t ybus_hip_analyze_lu(uint64_t n, const uint64_t* l_ptr, const uint64_t* l_idx, const uint64_t* u_ptr,
                            const uint64_t* u_idx, int32_t* forward_level, int32_t* backward_level,
                            int32_t* transpose_forward_level, int32_t* transpose_backward_level) {
  return guard([&] {
    if (!l_ptr || !u_ptr || (!l_idx && l_ptr[n]) || (!u_idx && u_ptr[n]))
      fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    std::string why;
    if (!valid_compressed((int64_t)n, (int64_t)n, l_ptr, l_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "L: " + why);
    if (!valid_compressed((int64_t)n, (int64_t)n, u_ptr, u_idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "U: " + why);
    for (uint64_t j = 0; j < n; ++j)
      if (l_ptr[j + 1] == l_ptr[j] || l_idx[l_ptr[j]] != j || u_ptr[j + 1] == u_ptr[j] || u_idx[u_ptr[j + 1] - 1] != j)
        fail(YBUS_HIP_INVALID_ARGUMENT, "L must start and U end each column with its diagonal");
    std::vector<double> lz(l_ptr[n], 0.0), uz(u_ptr[n], 0.0);
    copy_levels(lu_forward((int64_t)n, l_ptr, l_idx, lz.data()), forward_level);
    copy_levels(lu_backward((int64_t)n, u_ptr, u_idx, uz.data()), backward_level);
    copy_levels(lu_transpose_forward((int64_t)n, u_ptr, u_idx, uz.data()), transpose_forward_level);
    copy_levels(lu_transpose_backward((int64_t)n, l_ptr, l_idx, lz.data()), transpose_backward_level);
  });
}

int32_t ybus_hip_factor_operator(ybus_hip_factor* f, uint64_t rows, const uint64_t* ptr, const uint64_t* idx,
                                 const double* val, ybus_hip_operator** out) {
  return guard([&] {
    if (!f || !ptr || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    require_real(f);
    std::string why;
    if (!valid_compressed((int64_t)rows, f->n, ptr, idx, &why)) fail(YBUS_HIP_INVALID_ARGUMENT, "operator: " + why);
    const uint64_t nnz = ptr[rows];
    if (nnz > (uint64_t)INDEX_LIMIT || rows > (uint64_t)INDEX_LIMIT)
      fail(YBUS_HIP_TOO_LARGE, "an operator with " + std::to_string(nnz) + " entries is too large for 32-bit indices");
    check(hipSetDevice(f->ctx->device), "hipSetDevice");
    auto op = std::make_unique<ybus_hip_operator>();
    op->f = f;
    op->m.rows = (int32_t)rows;
    op->m.cols = (int32_t)f->n;
    op->m.nnz = nnz;
    op->m.ptr.upload(to_i32(ptr, rows + 1, nnz), f->stream());
    op->m.col.upload(to_i32(idx, nnz, nnz), f->stream());
    op->m.val.upload(std::vector<double>(val, val + nnz), f->stream());
    op->offset.upload(std::vector<double>(std::max<uint64_t>(rows, 1), 0.0), f->stream());
    op->weight.upload(std::vector<double>(std::max<uint64_t>(rows, 1), 1.0), f->stream());
    check(hipStreamSynchronize(f->stream()), "operator");
    *out = op.release();
  });
}

void ybus_hip_operator_free(ybus_hip_operator* op) { delete op; }

uint64_t ybus_hip_operator_rows(const ybus_hip_operator* op) { return op ? (uint64_t)op->m.rows : 0; }

int32_t ybus_hip_operator_set_offset(ybus_hip_operator* op, const double* offset) {
  return guard([&] {
    if (!op || (!offset && op->m.rows)) fail(YBUS_HIP_INVALID_ARGUMENT, "one offset per row");
    if (!op->m.rows) return;
    check(hipMemcpy(op->offset.p, offset, op->m.rows * sizeof(double), hipMemcpyHostToDevice), "H2D");
  });
}

int32_t ybus_hip_operator_set_weights(ybus_hip_operator* op, const double* weights) {
  return guard([&] {
    if (!op || (!weights && op->m.rows)) fail(YBUS_HIP_INVALID_ARGUMENT, "one weight per row");
    if (!op->m.rows) return;
    check(hipMemcpy(op->weight.p, weights, op->m.rows * sizeof(double), hipMemcpyHostToDevice), "H2D");
  });
}

}  // extern "C"

// ---- batches: many factors' columns in one persistent launch per LDS size class ----

struct ybus_hip_batch {
  ybus_hip_ctx* ctx = nullptr;
  std::vector<ybus_hip_factor*> factors;
  // what is staged: counts per factor, where each factor's values and columns start
  std::vector<uint64_t> counts;
  std::vector<int64_t> offsets;
  int64_t values = 0, columns = 0;
  bool staged = false;
  DevBuf<double> rhs, x, x1, r;
  DevBuf<int32_t> steps;
  DevBuf<double> ratio;
  DevBuf<BatchEntryDev> entries;
  DevBuf<int32_t> queues;
  // one persistent launch per class of factors that fit `per_cu` workgroups to a compute unit;
  // a plan per mode (FP64, FP32-refined: their LDS needs differ)
  struct Class {
    int per_cu = 1, threads = 1024;
    size_t lds = 0;
    int32_t first = 0, count = 0;
  };
  struct Plan {
    bool ready = false;
    std::vector<Class> classes;
    DevBuf<int2> items;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double tol = 0.0;
    int32_t max_steps = -1;
    void drop_graph() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    }
  };
  Plan plans[2];
  int32_t refine_stop = YBUS_HIP_STOP_NORMWISE;  // the FP32-refined solves' stop rule
  Staging staging;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ybus_hip_kernel_stats stats{};
  // side streams (default priority) and events that fork the classes' launches off the batch's
  // stream inside the captured graph, so the classes run side by side, and join them back
  static constexpr int SIDE = 7;
  hipStream_t side[SIDE] = {};
  hipEvent_t fork = nullptr, join[SIDE] = {};

  hipStream_t stream() const { return ctx->stream; }
  void drop_plans() {
    for (Plan& p : plans) {
      p.drop_graph();
      p.ready = false;
    }
  }
  ~ybus_hip_batch() {
    drop_plans();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (fork) (void)hipEventDestroy(fork);
    for (int i = 0; i < SIDE; ++i) {
      if (join[i]) (void)hipEventDestroy(join[i]);
      if (side[i]) (void)hipStreamDestroy(side[i]);
    }
  }

  // The classes' shapes, most workgroups per compute unit first: (workgroups per CU, threads per
  // workgroup). Small factors run many narrow workgroups to a compute unit (a column's levels are
  // mostly narrow, and a workgroup's barrier waits for all its waves); measured on the 256-grid
  // batch, FP64: 2.31 ms with shapes 4:256, 2:512, 1:1024, 1.96 ms with these; the FP32
  // refinement, whose residual and iterate passes use every thread, keeps wider workgroups for its
  // large factors (9.6 ms against 10.9). YBUS_HIP_BATCH overrides both as "k:threads,k:threads,...".
  static std::vector<std::pair<int, int>> class_shapes(bool refined) {
    std::vector<std::pair<int, int>> shapes =
        refined ? std::vector<std::pair<int, int>>{{16, 64}, {8, 128}, {4, 256}, {2, 512}, {1, 1024}}
                : std::vector<std::pair<int, int>>{{16, 64}, {8, 128}, {4, 256}, {2, 256}, {1, 512}};
    if (const char* e = std::getenv("YBUS_HIP_BATCH")) {
      std::vector<std::pair<int, int>> mine;
      std::string spec(e);
      size_t at = 0;
      while (at < spec.size()) {
        size_t comma = spec.find(',', at);
        if (comma == std::string::npos) comma = spec.size();
        const std::string item = spec.substr(at, comma - at);
        const size_t colon = item.find(':');
        if (colon != std::string::npos) {
          const int k = std::atoi(item.substr(0, colon).c_str()), t = std::atoi(item.substr(colon + 1).c_str());
          if (k >= 1 && k <= 64 && t >= 64 && t <= 1024 && t % 64 == 0 && (t & (t - 1)) == 0) mine.push_back({k, t});
        }
        at = comma + 1;
      }
      if (!mine.empty()) {
        std::sort(mine.begin(), mine.end(), [](auto a, auto b) { return a.first > b.first; });
        shapes = mine;
      }
    }
    return shapes;
  }

  // The class (workgroups per CU) and LDS bytes of a factor's columns whose sweeps have `levels`
  // levels between them (0: none fits)
  static int class_of(const ybus_hip_factor* f, bool refined, int32_t levels, size_t& lds) {
    const int32_t n = (int32_t)std::max<int64_t>(f->n, 1);
    for (const auto& kt : class_shapes(refined)) {
      const size_t bytes = refined ? refine_lds_bytes(n, kt.second, levels) : solve_lds_bytes(n, false, levels);
      if ((size_t)kt.first * bytes <= f->lds_cap) {
        lds = bytes;
        return kt.first;
      }
    }
    return 0;
  }

  // Items in classes for a mode: each factor's columns go to the class of the most workgroups
  // per compute unit whose LDS they fit; longest first within a class (a column's time is set by
  // its levels, then by its rows)
  void make_plan(bool refined) {
    Plan& plan = plans[refined ? 1 : 0];
    struct Item {
      int per_cu;
      int64_t cost;
      int2 w;
      size_t lds;
    };
    std::vector<Item> all;
    all.reserve((size_t)columns);
    for (size_t i = 0; i < factors.size(); ++i) {
      ybus_hip_factor* f = factors[i];
      if (!counts[i]) continue;
      const int32_t levels = f->cfwd.host.levels + f->cbwd.host.levels;
      size_t lds = 0;
      const int per_cu = class_of(f, refined, levels, lds);
      if (!per_cu)
        fail(YBUS_HIP_INVALID_ARGUMENT,
             "factor " + std::to_string(i) + " (dim " + std::to_string(f->n) + ") does not fit in LDS for this mode");
      const int64_t cost = (int64_t)levels * 4096 + f->n;
      for (uint64_t c = 0; c < counts[i]; ++c) all.push_back(Item{per_cu, cost, make_int2((int)i, (int)c), lds});
    }
    std::stable_sort(all.begin(), all.end(), [](const Item& p, const Item& q) {
      return p.per_cu != q.per_cu ? p.per_cu < q.per_cu : p.cost > q.cost;
    });
    std::vector<int2> items(std::max<size_t>(all.size(), 1));
    plan.classes.clear();
    for (size_t k = 0; k < all.size(); ++k) {
      items[k] = all[k].w;
      if (plan.classes.empty() || plan.classes.back().per_cu != all[k].per_cu) {
        Class c;
        c.per_cu = all[k].per_cu;
        for (const auto& kt : class_shapes(refined))
          if (kt.first == c.per_cu) c.threads = kt.second;
        // the FP64 kernel of one or two workgroups a compute unit takes at most 512 threads
        if (!refined && c.per_cu <= 2) c.threads = std::min(c.threads, 512);
        c.first = (int32_t)k;
        plan.classes.push_back(c);
      }
      Class& c = plan.classes.back();
      c.count += 1;
      c.lds = std::max(c.lds, all[k].lds);
    }
    plan.items.upload(items, stream());
    check(hipStreamSynchronize(stream()), "batch plan");
    plan.ready = true;
  }

  // the queue counters to 0, then each class's launch: the first on the batch's stream, the
  // others forked onto side streams and joined back (in a captured graph: parallel branches, so
  // the small classes fill the compute units the large one leaves idle at its tail)
  void enqueue(bool refined, double tol, int32_t max_steps) {
    const Plan& plan = plans[refined ? 1 : 0];
    check(launch_fill(stream(), queues.p, 0, (int32_t)plan.classes.size()), "batch queue reset");
    const int classes = (int)plan.classes.size();
    if (classes > 1) {
      if (!fork) check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
      for (int i = 0; i + 1 < classes && i < SIDE; ++i) {
        if (!side[i]) check(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking), "hipStreamCreate");
        if (!join[i]) check(hipEventCreateWithFlags(&join[i], hipEventDisableTiming), "hipEventCreate");
      }
      check(hipEventRecord(fork, stream()), "hipEventRecord");
    }
    for (int c = 0; c < classes; ++c) {
      const Class& k = plan.classes[c];
      const bool forked = c > 0 && c - 1 < SIDE;
      const hipStream_t on = forked ? side[c - 1] : stream();
      if (forked) check(hipStreamWaitEvent(on, fork, 0), "hipStreamWaitEvent");
      const int grid = std::max(1, std::min<int>(k.count, k.per_cu * ctx->cus));
      if (refined)
        check(launch_batch_refine(on, grid, k.threads, k.lds, entries.p, plan.items.p + k.first, k.count,
                                  queues.p + c, rhs.p, x.p, x1.p, r.p, tol, max_steps, steps.p, ratio.p, refine_stop),
              "batch kernel");
      else
        check(launch_batch(on, grid, k.threads, k.lds, k.per_cu, entries.p, plan.items.p + k.first, k.count,
                           queues.p + c, rhs.p, x.p),
              "batch kernel");
      if (forked) check(hipEventRecord(join[c - 1], on), "hipEventRecord");
    }
    for (int c = 1; c < classes && c - 1 < SIDE; ++c)
      check(hipStreamWaitEvent(stream(), join[c - 1], 0), "hipStreamWaitEvent");
  }
};

extern "C" {

int32_t ybus_hip_batch_create(ybus_hip_ctx* ctx, ybus_hip_batch** out) {
  return guard([&] {
    if (!ctx || !out) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto b = std::make_unique<ybus_hip_batch>();
    b->ctx = ctx;
    check(hipEventCreate(&b->ev0), "hipEventCreate");
    check(hipEventCreate(&b->ev1), "hipEventCreate");
    *out = b.release();
  });
}

void ybus_hip_batch_free(ybus_hip_batch* b) { delete b; }

int32_t ybus_hip_batch_set_refine_stop(ybus_hip_batch* b, int32_t rule) {
  return guard([&] {
    if (!b) fail(YBUS_HIP_INVALID_ARGUMENT, "batch is null");
    if (rule != YBUS_HIP_STOP_NORMWISE && rule != YBUS_HIP_STOP_AT_FLOOR) fail(YBUS_HIP_INVALID_ARGUMENT, "unknown rule");
    if (b->refine_stop != rule) b->plans[1].drop_graph();  // the captured launches carry the rule
    b->refine_stop = rule;
  });
}

uint64_t ybus_hip_batch_factors(const ybus_hip_batch* b) { return b ? (uint64_t)b->factors.size() : 0; }

int32_t ybus_hip_batch_add(ybus_hip_batch* b, ybus_hip_factor* f) {
  return guard([&] {
    if (!b || !f) fail(YBUS_HIP_INVALID_ARGUMENT, "null argument");
    require_real(f);
    if (f->ctx != b->ctx) fail(YBUS_HIP_INVALID_ARGUMENT, "the factor belongs to another context");
    if (!f->fits)
      fail(YBUS_HIP_INVALID_ARGUMENT, "a batch takes factors whose solution fits in LDS (dim " + std::to_string(f->n) +
                                          " does not)");
    b->factors.push_back(f);
    b->staged = false;
    b->drop_plans();
  });
}

int32_t ybus_hip_batch_stage(ybus_hip_batch* b, const double* rhs, const uint64_t* counts) {
  return guard([&] {
    if (!b) fail(YBUS_HIP_INVALID_ARGUMENT, "batch is null");
    const size_t nf = b->factors.size();
    if (nf && !counts) fail(YBUS_HIP_INVALID_ARGUMENT, "one count per factor");
    check(hipSetDevice(b->ctx->device), "hipSetDevice");
    b->staged = false;
    b->drop_plans();
    b->counts.assign(counts, counts + nf);
    b->offsets.assign(nf, 0);
    int64_t values = 0, columns = 0;
    std::vector<BatchEntryDev> entries(std::max<size_t>(nf, 1));
    for (size_t i = 0; i < nf; ++i) {
      ybus_hip_factor* f = b->factors[i];
      if (b->counts[i] > (uint64_t)INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "too many right-hand sides");
      BatchEntryDev& e = entries[i];
      e.fwd = f->cfwd.view();
      e.bwd = f->cbwd.view();
      e.rhs_row = f->rhs_row.p;
      e.sol_of = f->sol_of.p;
      e.n = (int32_t)f->n;
      e.offset = values;
      e.a = f->a.view();
      e.first_column = columns;
      b->offsets[i] = values;
      values += (int64_t)b->counts[i] * f->n;
      columns += (int64_t)b->counts[i];
    }
    if (columns > INDEX_LIMIT) fail(YBUS_HIP_TOO_LARGE, "too many right-hand sides");
    if (values && !rhs) fail(YBUS_HIP_INVALID_ARGUMENT, "rhs is null");
    b->values = values;
    b->columns = columns;
    b->entries.upload(entries, b->stream());
    b->queues.reserve(8);
    const size_t vals = (size_t)std::max<int64_t>(values, 1), cols = (size_t)std::max<int64_t>(columns, 1);
    for (DevBuf<double>* d : {&b->rhs, &b->x, &b->x1, &b->r}) d->reserve(vals);
    b->steps.reserve(cols);
    b->ratio.reserve(cols);
    if (values) b->staging.to_device(b->stream(), b->rhs.p, rhs, (size_t)values * sizeof(double));
    check(hipStreamSynchronize(b->stream()), "batch stage");
    b->staged = true;
  });
}

int32_t ybus_hip_batch_solve_staged(ybus_hip_batch* b, int32_t mode, double tolerance, uint64_t max_steps) {
  return guard([&] {
    if (!b) fail(YBUS_HIP_INVALID_ARGUMENT, "batch is null");
    if (!b->staged) fail(YBUS_HIP_INVALID_ARGUMENT, "nothing staged (ybus_hip_batch_stage)");
    if (mode != YBUS_HIP_FP64 && mode != YBUS_HIP_FP32_REFINED) fail(YBUS_HIP_INVALID_ARGUMENT, "mode: FP64 or FP32_REFINED");
    if (max_steps > 1000000) fail(YBUS_HIP_INVALID_ARGUMENT, "max_steps is out of range");
    const bool refined = mode == YBUS_HIP_FP32_REFINED;
    if (refined)
      for (size_t i = 0; i < b->factors.size(); ++i)
        if (b->counts[i] && !b->factors[i]->a.ptr.p)
          fail(YBUS_HIP_INVALID_ARGUMENT, "FP32 refinement needs each factor's matrix (ybus_hip_factor_set_matrix)");
    if (!b->columns) return;
    check(hipSetDevice(b->ctx->device), "hipSetDevice");
    ybus_hip_batch::Plan& plan = b->plans[refined ? 1 : 0];
    if (!plan.ready) b->make_plan(refined);
    if (plan.exec && refined && (plan.tol != tolerance || plan.max_steps != (int32_t)max_steps)) plan.drop_graph();
    if (!plan.exec) {
      for (const ybus_hip_batch::Class& c : plan.classes)
        check(refined ? prepare_batch_refine(c.lds) : prepare_batch(c.lds, c.per_cu), "batch kernel attribute");
      check(hipStreamBeginCapture(b->stream(), hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        b->enqueue(refined, tolerance, (int32_t)max_steps);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(b->stream(), &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      check(hipStreamEndCapture(b->stream(), &plan.graph), "hipStreamEndCapture");
      check(hipGraphInstantiate(&plan.exec, plan.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      plan.tol = tolerance;
      plan.max_steps = (int32_t)max_steps;
    }
    check(hipEventRecord(b->ev0, b->stream()), "hipEventRecord");
    check(hipGraphLaunch(plan.exec, b->stream()), "hipGraphLaunch");
    check(hipEventRecord(b->ev1, b->stream()), "hipEventRecord");
    check(hipEventSynchronize(b->ev1), "batch solve");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "hipEventElapsedTime");
    b->stats.launches += 1;
    b->stats.milliseconds += ms;
    uint64_t bytes = 0;
    for (size_t i = 0; i < b->factors.size(); ++i)
      if (b->counts[i])
        bytes += (uint64_t)b->factors[i]->nnz_factor * 12 + 8 * (uint64_t)(b->factors[i]->n + 1) +
                 b->counts[i] * (uint64_t)b->factors[i]->n * 16;
    b->stats.bytes += bytes;
  });
}

int32_t ybus_hip_batch_solutions(ybus_hip_batch* b, double* x) {
  return guard([&] {
    if (!b) fail(YBUS_HIP_INVALID_ARGUMENT, "batch is null");
    if (!b->staged) fail(YBUS_HIP_INVALID_ARGUMENT, "nothing staged (ybus_hip_batch_stage)");
    if (!b->values) return;
    if (!x) fail(YBUS_HIP_INVALID_ARGUMENT, "x is null");
    b->staging.to_host(b->stream(), x, b->x.p, (size_t)b->values * sizeof(double));
  });
}

int32_t ybus_hip_batch_refinement(ybus_hip_batch* b, uint64_t* steps, double* ratio) {
  return guard([&] {
    if (!b) fail(YBUS_HIP_INVALID_ARGUMENT, "batch is null");
    if (!b->staged) fail(YBUS_HIP_INVALID_ARGUMENT, "nothing staged (ybus_hip_batch_stage)");
    if (!b->columns) return;
    std::vector<int32_t> st((size_t)b->columns);
    std::vector<double> ra((size_t)b->columns);
    check(hipMemcpy(st.data(), b->steps.p, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost), "steps");
    check(hipMemcpy(ra.data(), b->ratio.p, ra.size() * sizeof(double), hipMemcpyDeviceToHost), "ratio");
    for (int64_t c = 0; c < b->columns; ++c) {
      if (steps) steps[c] = (uint64_t)st[c];
      if (ratio) ratio[c] = ra[c];
    }
  });
}

int32_t ybus_hip_batch_solve(ybus_hip_batch* b, const double* rhs, const uint64_t* counts, double* x) {
  int32_t status = ybus_hip_batch_stage(b, rhs, counts);
  if (status != YBUS_HIP_OK) return status;
  status = ybus_hip_batch_solve_staged(b, YBUS_HIP_FP64, 0.0, 0);
  if (status != YBUS_HIP_OK) return status;
  return ybus_hip_batch_solutions(b, x);
}

void ybus_hip_batch_kernel_stats(const ybus_hip_batch* b, ybus_hip_kernel_stats* out) {
  if (b && out) *out = b->stats;
}

void ybus_hip_batch_reset_kernel_stats(ybus_hip_batch* b) {
  if (b) b->stats = ybus_hip_kernel_stats{};
}

}  // extern "C"
