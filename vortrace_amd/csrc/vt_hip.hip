// libvortrace_hip: HIP kernels for gfx950 and the C-ABI of include/vortrace_hip.h.
//
// Kernels (one ray or one query per lane, persistent grid, rays pulled from a
// device work counter):
//   k_query      PointCloud::queryTreeK           (src/pointcloud.cpp:308-368)
//   k_project    Projection::makeProjection       (src/projection.cpp:26-73)
//                fused with reduce_sum / reduce_max / reduce_min /
//                reduce_volume_render             (src/reduction.cpp:8-83)
//   k_count      segment counts of makeDetailedProjection
//   k_fill       segments + values of makeDetailedProjection
//                                                 (src/projection.cpp:75-144)
//   k_segs       segment lists for the adjoint (adjoint.h)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "../../include/vortrace_hip.h"
#include "adjoint.h"
#include "engine.h"
#include "order.h"
#include "tree_build.h"

#pragma clang fp contract(off)

namespace vt {

constexpr int kWalkCap = 128;  // pending ray points per lane (device memory)
// Intervals walked for another lane of the wave: events a lane can record
// for one interval, and intervals a lane can have given away at once.
constexpr int kHelpEvents = 128;
constexpr int kHelpGifts = 16;

// ---------------------------------------------------------------- sinks ----

template <int NF>
struct SumSink {  // reduce_sum, src/reduction.cpp:8-21
  const double *fields;
  double acc[NF];
  uint64_t n;
  __device__ void init() {
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = 0.0;
    n = 0;
  }
  __device__ __forceinline__ void seg(uint32_t cell, double a, double b) {
    const double ds = b - a;
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = acc[f] + ds * gld(fields + (size_t)cell * NF + f);
    ++n;
  }
  __device__ void store(double *out) const {
#pragma unroll
    for (int f = 0; f < NF; ++f) gst(out + f, acc[f]);
  }
};

struct SumSinkDyn {  // reduce_sum for any number of fields, accumulated in place
  const double *fields;
  int nf;
  double *out;
  uint64_t n;
  __device__ void init() {
    for (int f = 0; f < nf; ++f) gst(out + f, 0.0);
    n = 0;
  }
  __device__ void seg(uint32_t cell, double a, double b) {
    const double ds = b - a;
    for (int f = 0; f < nf; ++f)
      gst(out + f, gld(out + f) + ds * gld(fields + (size_t)cell * nf + f));
    ++n;
  }
  __device__ void store(double *) const {}
};

template <bool kMax>
struct ExtSink {  // reduce_max / reduce_min, src/reduction.cpp:23-53
  const double *fields;
  int nf;
  double *out;
  uint64_t n;
  __device__ void init() {
    for (int f = 0; f < nf; ++f) gst(out + f, kMax ? -(double)INFINITY : (double)INFINITY);
    n = 0;
  }
  __device__ void seg(uint32_t cell, double, double) {
    for (int f = 0; f < nf; ++f) {
      const double v = gld(fields + (size_t)cell * nf + f);
      const double cur = gld(out + f);
      if (kMax ? (v > cur) : (v < cur)) gst(out + f, v);
    }
    ++n;
  }
  __device__ void store(double *) const {}
};

struct VolSink {  // reduce_volume_render, src/reduction.cpp:55-83
  const double *fields;
  double T, c[3];
  uint64_t n;
  __device__ void init() {
    T = 1.0;
    c[0] = c[1] = c[2] = 0.0;
    n = 0;
  }
  __device__ __forceinline__ void seg(uint32_t cell, double a, double b) {
    const double2 *fv = reinterpret_cast<const double2 *>(fields + (size_t)cell * 4);
    const double2 rg = gld(fv), ba = gld(fv + 1);
    const double col[3] = {rg.x, rg.y, ba.x};
    const double alpha = ba.y;
    const double ds = b - a;
    const double opacity = 1.0 - exp(-alpha * ds);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] += T * col[k] * opacity;
    T *= (1.0 - opacity);
    ++n;
  }
  __device__ void store(double *out) const {
    gst(out, c[0]);
    gst(out + 1, c[1]);
    gst(out + 2, c[2]);
  }
};

struct CountSink {
  uint64_t n;
  __device__ void init() { n = 0; }
  __device__ __forceinline__ void seg(uint32_t, double, double) { ++n; }
  __device__ void store(double *) const {}
};

template <class Inner>
struct FillSink {  // segments at the ray's offset, plus the reduced values
  Inner in;
  const uint32_t *slot_id;
  int64_t *ids;
  double *se, *sx;
  uint64_t cap;  // room the count pass reserved for this ray
  uint64_t n;
  __device__ void init() {
    in.init();
    n = 0;
  }
  __device__ __forceinline__ void seg(uint32_t cell, double a, double b) {
    if (n < cap) {
      gst(ids + n, (int64_t)gld(slot_id + cell));
      gst(se + n, a);
      gst(sx + n, b);
    }
    ++n;
    in.seg(cell, a, b);
  }
  __device__ void store(double *out) const { in.store(out); }
};

// ------------------------------------------------------------ ray I/O ----
// Per-ray sink set-up and output for run_engine (engine.h).

// Every projection IO also writes the ray's segment count when `counts` is
// given (the forward pass of a differentiable projection keeps them for the
// adjoint).
__device__ __forceinline__ void put_count(int64_t *counts, int64_t ray, uint64_t n) {
  if (counts) gst(counts + ray, (int64_t)n);
}

template <int NF>
struct SumIO {
  using Sink = SumSink<NF>;
  const double *fields;
  double *out;
  int nout;
  int64_t *counts;
  __device__ void begin(int64_t, Sink &sk) const {
    sk.fields = fields;
    sk.init();
  }
  __device__ int end(int64_t ray, Sink &sk) const {
    sk.store(out + ray * nout);
    put_count(counts, ray, sk.n);
    return kErrNone;
  }
};

struct SumDynIO {
  using Sink = SumSinkDyn;
  const double *fields;
  int nf;
  double *out;
  int nout;
  int64_t *counts;
  __device__ void begin(int64_t ray, Sink &sk) const {
    sk.fields = fields;
    sk.nf = nf;
    sk.out = out + ray * nout;
    sk.init();
  }
  __device__ int end(int64_t ray, Sink &sk) const {
    put_count(counts, ray, sk.n);
    return kErrNone;
  }
};

template <bool kMax>
struct ExtIO {
  using Sink = ExtSink<kMax>;
  const double *fields;
  int nf;
  double *out;
  int nout;
  int64_t *counts;
  __device__ void begin(int64_t ray, Sink &sk) const {
    sk.fields = fields;
    sk.nf = nf;
    sk.out = out + ray * nout;
    sk.init();
  }
  __device__ int end(int64_t ray, Sink &sk) const {
    put_count(counts, ray, sk.n);
    return kErrNone;
  }
};

struct VolIO {
  using Sink = VolSink;
  const double *fields;
  double *out;
  int nout;
  int64_t *counts;
  __device__ void begin(int64_t, Sink &sk) const {
    sk.fields = fields;
    sk.init();
  }
  __device__ int end(int64_t ray, Sink &sk) const {
    sk.store(out + ray * nout);
    put_count(counts, ray, sk.n);
    return kErrNone;
  }
};

struct CountIO {
  using Sink = CountSink;
  int64_t *counts;
  __device__ void begin(int64_t, Sink &sk) const { sk.init(); }
  __device__ int end(int64_t ray, Sink &sk) const {
    gst(counts + ray, (int64_t)sk.n);
    return kErrNone;
  }
};

// Segments at the ray's offset (from the count pass) plus the reduced values.
template <class InnerIO>
struct FillIO {
  using Sink = FillSink<typename InnerIO::Sink>;
  InnerIO inner;
  const int64_t *offsets;
  const uint32_t *slot_id;
  int64_t *ids;
  double *se, *sx;
  __device__ void begin(int64_t ray, Sink &sk) const {
    const int64_t off = gld(offsets + ray);
    sk.slot_id = slot_id;
    sk.ids = ids + off;
    sk.se = se + off;
    sk.sx = sx + off;
    sk.cap = (uint64_t)(gld(offsets + ray + 1) - off);
    inner.begin(ray, sk.in);
    sk.n = 0;
  }
  __device__ int end(int64_t ray, Sink &sk) const {
    inner.end(ray, sk.in);
    return sk.n == sk.cap ? kErrNone : kErrFillOverflow;
  }
};

// -------------------------------------------------------------- kernels ----

__device__ __forceinline__ void flush_stats(Stats *stats, uint64_t nseg, uint64_t nq,
                                            uint64_t nslow, int maxw, int maxk) {
  atomicAdd(&stats->segments, (unsigned long long)nseg);
  atomicAdd(&stats->queries, (unsigned long long)nq);
  if (nslow) atomicAdd(&stats->slow_queries, (unsigned long long)nslow);
  atomicMax(&stats->max_walk_depth, (unsigned int)maxw);
  atomicMax(&stats->max_knn_depth, (unsigned int)maxk);
}

#define VT_LANE_SETUP(PER)                                                      \
  const int tid = threadIdx.x;                                                  \
  const int gtid = blockIdx.x * kBlock + tid;                                   \
  const LaneInfo lane_info{tid, gtid, gtid >> 6, (int)(tid & 63)};              \
  KStack st = make_stack(S, lane_info);

// With HELP (fast kernels only), idle lanes walk intervals of busy lanes'
// rays.  Each projection and count kernel is built with and without it: the
// helping code costs every iteration a little, and it pays only where lanes
// run out of rays long before the launch ends (help_pays).
template <bool PER, bool EXP, bool HELP = false, class IO>
__device__ __forceinline__ void run_rays(const DevCloud &C, const Scratch &S, const double *start,
                                         const double *end, int64_t nrays,
                                         const uint32_t *ray_list, uint32_t *defer, ErrRec *err,
                                         Stats *stats, IO &io) {
  VT_LANE_SETUP(PER)
  (void)lane_info;
  EngineCounters cnt;
  run_engine<PER, EXP, !EXP && HELP>(C, S, start, end, nrays, ray_list, defer, err, stats, io,
                                      st, cnt);
  flush_stats(stats, cnt.nseg, cnt.nq + cnt.nslow, cnt.nslow, cnt.maxw, st.maxsp);
}

// Every kernel below runs the rays [0, nrays) (or ray_list[0, nrays)).  The
// fast variants (EXP = false) append rays that need the K_EXPAND
// classification to `defer`; the EXP = true variants handle every ray.

// With REC every segment is also recorded to `pool` (differentiable forward).
template <bool REC, bool PER, bool EXP, bool HELP = false, class IO>
__device__ __forceinline__ void run_io(const DevCloud &C, const Scratch &S, const double *start,
                                       const double *end, int64_t nrays,
                                       const uint32_t *ray_list, uint32_t *defer, ErrRec *err,
                                       Stats *stats, const IO &io, const RecPool &pool) {
  if constexpr (REC) {
    RecordIO<IO> rio{io, pool};
    run_rays<PER, EXP, HELP>(C, S, start, end, nrays, ray_list, defer, err, stats, rio);
  } else {
    IO cio = io;
    run_rays<PER, EXP, HELP>(C, S, start, end, nrays, ray_list, defer, err, stats, cio);
  }
}

// At most 256 VGPRs (two waves per SIMD): the recording kernels with the
// helping code would otherwise take a few more and run one wave per SIMD.
template <int MODE, int NF, bool PER, bool REC = false, bool HELP = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_project(
    DevCloud C, Scratch S, const double *start, const double *end, int64_t nrays, uint32_t *defer,
    double *out, int nout, int64_t *counts, RecPool pool, ErrRec *err, Stats *stats) {
  if constexpr (MODE == VT_SUM && NF > 0) {
    run_io<REC, PER, false, HELP>(C, S, start, end, nrays, nullptr, defer, err, stats,
                                  SumIO<NF>{C.fields, out, nout, counts}, pool);
  } else if constexpr (MODE == VT_SUM) {
    run_io<REC, PER, false, HELP>(C, S, start, end, nrays, nullptr, defer, err, stats,
                                  SumDynIO{C.fields, (int)C.nf, out, nout, counts}, pool);
  } else if constexpr (MODE == VT_MAX || MODE == VT_MIN) {
    run_io<REC, PER, false, HELP>(
        C, S, start, end, nrays, nullptr, defer, err, stats,
        ExtIO<MODE == VT_MAX>{C.fields, (int)C.nf, out, nout, counts}, pool);
  } else {
    run_io<REC, PER, false, HELP>(C, S, start, end, nrays, nullptr, defer, err, stats,
                                  VolIO{C.fields, out, nout, counts}, pool);
  }
}

// The deferred rays of k_project.  Sums accumulate in place, in the same order
// as SumSink's registers, so the values are the same bits.
template <int MODE, bool PER, bool REC = false>
__global__ __launch_bounds__(kBlock) void k_project_x(DevCloud C, Scratch S,
                                                      const double *start, const double *end,
                                                      int64_t nrays, const uint32_t *ray_list,
                                                      double *out, int nout, int64_t *counts,
                                                      RecPool pool, ErrRec *err, Stats *stats) {
  if constexpr (MODE == VT_SUM) {
    run_io<REC, PER, true>(C, S, start, end, nrays, ray_list, nullptr, err, stats,
                           SumDynIO{C.fields, (int)C.nf, out, nout, counts}, pool);
  } else if constexpr (MODE == VT_MAX || MODE == VT_MIN) {
    run_io<REC, PER, true>(C, S, start, end, nrays, ray_list, nullptr, err, stats,
                           ExtIO<MODE == VT_MAX>{C.fields, (int)C.nf, out, nout, counts}, pool);
  } else {
    run_io<REC, PER, true>(C, S, start, end, nrays, ray_list, nullptr, err, stats,
                           VolIO{C.fields, out, nout, counts}, pool);
  }
}

template <bool PER, bool EXP, bool HELP = false>
__global__ __launch_bounds__(kBlock) void k_count(DevCloud C, Scratch S, const double *start,
                                                  const double *end, int64_t nrays,
                                                  const uint32_t *ray_list, uint32_t *defer,
                                                  int64_t *counts, ErrRec *err, Stats *stats) {
  CountIO io{counts};
  run_rays<PER, EXP, HELP>(C, S, start, end, nrays, ray_list, defer, err, stats, io);
}

template <int MODE, bool PER, bool EXP>
__global__ __launch_bounds__(kBlock) void k_fill(DevCloud C, Scratch S, const double *start,
                                                 const double *end, int64_t nrays,
                                                 const uint32_t *ray_list, uint32_t *defer,
                                                 const int64_t *offsets, double *out, int nout,
                                                 const uint32_t *slot_id, int64_t *ids,
                                                 double *se, double *sx, ErrRec *err,
                                                 Stats *stats) {
  if constexpr (MODE == VT_SUM) {
    FillIO<SumDynIO> io{SumDynIO{C.fields, (int)C.nf, out, nout, nullptr}, offsets, slot_id, ids,
                        se, sx};
    run_rays<PER, EXP>(C, S, start, end, nrays, ray_list, defer, err, stats, io);
  } else if constexpr (MODE == VT_MAX || MODE == VT_MIN) {
    FillIO<ExtIO<MODE == VT_MAX>> io{ExtIO<MODE == VT_MAX>{C.fields, (int)C.nf, out, nout, nullptr},
                                     offsets, slot_id, ids, se, sx};
    run_rays<PER, EXP>(C, S, start, end, nrays, ray_list, defer, err, stats, io);
  } else {
    FillIO<VolIO> io{VolIO{C.fields, out, nout, nullptr}, offsets, slot_id, ids, se, sx};
    run_rays<PER, EXP>(C, S, start, end, nrays, ray_list, defer, err, stats, io);
  }
}

template <bool PER, bool EXP>
__global__ __launch_bounds__(kBlock) void k_segs(DevCloud C, Scratch S, const double *start,
                                                 const double *end, int64_t nrays,
                                                 const uint32_t *ray_list, uint32_t *defer,
                                                 const int64_t *offsets, uint32_t *slot,
                                                 double *ds, ErrRec *err, Stats *stats) {
  SegIO io{offsets, slot, ds};
  run_rays<PER, EXP>(C, S, start, end, nrays, ray_list, defer, err, stats, io);
}

template <int K, bool PER>
__global__ __launch_bounds__(kBlock) void k_query_reg(DevCloud C, Scratch S, const double *q,
                                                      int64_t nq, const uint32_t *slot_id,
                                                      int64_t *ids, double *r2, ErrRec *err,
                                                      Stats *stats) {
  VT_LANE_SETUP(PER)
  for (;;) {
    const int64_t i = (int64_t)atomicAdd(&stats->ray_counter, 1ULL);
    if (i >= nq) break;
    uint32_t id[K];
    double d[K];
    query_reg<K, PER>(C, gld(q + 3 * i), gld(q + 3 * i + 1), gld(q + 3 * i + 2), id, d, st);
#pragma unroll
    for (int j = 0; j < K; ++j) {
      gst(ids + i * K + j, id[j] == kNone ? (int64_t)-1 : (int64_t)gld(slot_id + id[j]));
      gst(r2 + i * K + j, d[j]);
    }
  }
  if (st.overflow) report_error(err, kErrKnnStack, 0, 0, 0, 0, 0, 0);
  flush_stats(stats, 0, 0, 0, 0, st.maxsp);
}

template <bool PER>
__global__ __launch_bounds__(kBlock) void k_query_mem(DevCloud C, Scratch S, const double *q,
                                                      int64_t nq, int k, uint32_t *wid,
                                                      double *wd, const uint32_t *slot_id,
                                                      int64_t *ids, double *r2, ErrRec *err,
                                                      Stats *stats) {
  VT_LANE_SETUP(PER)
  uint32_t *bi = wid + (size_t)gtid * 4 * k;
  double *bd = wd + (size_t)gtid * 4 * k;
  for (;;) {
    const int64_t i = (int64_t)atomicAdd(&stats->ray_counter, 1ULL);
    if (i >= nq) break;
    MemRes res{bd, bi, k, 0};
    query_mem<PER>(C, gld(q + 3 * i), gld(q + 3 * i + 1), gld(q + 3 * i + 2), k, res,
                   bi + 2 * k, bd + 2 * k, st);
    for (int j = 0; j < k; ++j) {
      const uint32_t id = gld(bi + j);
      gst(ids + i * k + j, id == kNone ? (int64_t)-1 : (int64_t)gld(slot_id + id));
      gst(r2 + i * k + j, gld(bd + j));
    }
  }
  if (st.overflow) report_error(err, kErrKnnStack, 0, 0, 0, 0, 0, 0);
  flush_stats(stats, 0, 0, 0, 0, st.maxsp);
}

// Slice::makeSlice (src/slice.cpp:7-60): the nearest cell of every pixel of
// an nx x ny grid at z = depth, pixel (i, j) at (x0 + dx * i, y0 + dy * j);
// out[(i * ny + j) * nf + f] is that cell's field f.
template <bool PER>
__global__ __launch_bounds__(kBlock) void k_slice(DevCloud C, Scratch S, double x0, double dx,
                                                  double y0, double dy, double depth,
                                                  int64_t nx, int64_t ny, double *out,
                                                  ErrRec *err, Stats *stats) {
  VT_LANE_SETUP(PER)
  const int64_t np = nx * ny;
  for (;;) {
    const int64_t p = (int64_t)atomicAdd(&stats->ray_counter, 1ULL);
    if (p >= np) break;
    const int64_t i = p / ny, j = p - i * ny;
    const double qx = x0 + dx * (double)i;
    const double qy = y0 + dy * (double)j;
    uint32_t id[1];
    double d[1];
    query_reg<1, PER>(C, qx, qy, depth, id, d, st);
    if (id[0] == kNone) {
      report_error(err, kErrNonFinite, p, qx, qy, depth, 0, 0);
      continue;
    }
    for (uint32_t f = 0; f < C.nf; ++f)
      gst(out + p * C.nf + f, gld(C.fields + (size_t)id[0] * C.nf + f));
  }
  if (st.overflow) report_error(err, kErrKnnStack, 0, 0, 0, 0, 0, 0);
  flush_stats(stats, 0, 0, 0, 0, st.maxsp);
}

}  // namespace vt

// ================================================================ host ====

using namespace vt;

namespace {

thread_local std::string g_err;
thread_local vt_call_stats g_stats{};
std::mutex g_warn_mu;
vt_warning_fn g_warn_fn = nullptr;
void *g_warn_user = nullptr;
std::atomic<int> g_verbose{0};

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

void warn(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_warn_mu);
  if (g_warn_fn)
    g_warn_fn(msg.c_str(), g_warn_user);
  else
    std::fprintf(stderr, "vortrace warning: %s\n", msg.c_str());
}

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_OK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw HipError(std::string("HIP error in " #expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
T *dev_alloc(size_t count) {
  void *p = nullptr;
  if (count == 0) count = 1;
  HIP_OK(hipMalloc(&p, count * sizeof(T)));
  return static_cast<T *>(p);
}

std::string fmt(const char *f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  std::vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

}  // namespace

struct vt_cloud {
  // host copy (filtered order), PointCloud members of include/pointcloud.hpp:21-32
  size_t n = 0, nf = 1;
  std::vector<double> pts, fields;
  std::vector<int64_t> orig_ids;
  double subbox[6] = {0, 0, 0, 0, 0, 0};
  double box[3] = {0, 0, 0};
  double pad = 0.0;
  int periodic = 0, valid_rgba = 0;
  bool tree_built = false;  // host tree built and uploaded
  bool host_tree = false;   // host tree built
  HostTree tree;
  double build_seconds = 0;

  // device copy
  int device = -1;
  DNode *d_nodes = nullptr;
  DRec *d_recs = nullptr;
  Gen *d_gen = nullptr;
  double *d_fields = nullptr;
  uint32_t *d_slot_id = nullptr;
  ErrRec *d_err = nullptr;
  Stats *d_stats = nullptr;
  Scratch scratch{};
  size_t scratch_lanes = 0;
  uint4 *help_ev = nullptr, *help_steal = nullptr;  // interval buffers (scratch_lanes lanes)
  uint32_t ray_row = 0;         // rays per image row of the coming calls (0 = no image)
  uint32_t *d_order = nullptr;  // a caller's ray order for calls of order_n rays
  size_t order_cap = 0;
  int64_t order_n = 0;
  uint32_t *d_defer = nullptr;  // rays left to the K_EXPAND-capable kernels
  size_t defer_cap = 0;
  uint32_t *d_order_hist = nullptr;  // ray-order estimate: generators per histogram cell
  float *d_order_grid = nullptr;     // and their cube roots (built on first use)
  int kspill = 0;
  DevCloud dc{};
  // Work buffers of the adjoint, recording and field-update calls, kept
  // between calls (slot i of a call's i-th request, grown when too small).
  std::vector<std::pair<void *, size_t>> arena;

  void free_device() {
    for (void *p : {(void *)d_nodes, (void *)d_recs, (void *)d_gen, (void *)d_fields, (void *)d_slot_id,
                    (void *)d_err, (void *)d_stats, (void *)scratch.knn_spill,
                    (void *)scratch.walk_tag, (void *)scratch.walk_s,
                    (void *)scratch.big_id, (void *)scratch.big_d, (void *)help_ev,
                    (void *)help_steal, (void *)d_defer,
                    (void *)d_order, (void *)d_order_hist, (void *)d_order_grid})
      if (p) (void)hipFree(p);
    d_order_hist = nullptr;
    d_order_grid = nullptr;
    d_order = nullptr;
    order_cap = 0;
    order_n = 0;
    for (auto &a : arena)
      if (a.first) (void)hipFree(a.first);
    arena.clear();
    d_defer = nullptr;
    defer_cap = 0;
    d_nodes = nullptr;
    d_recs = nullptr;
    d_gen = nullptr;
    d_fields = nullptr;
    d_slot_id = nullptr;
    d_err = nullptr;
    d_stats = nullptr;
    scratch = Scratch{};
    scratch_lanes = 0;
    help_ev = nullptr;
    help_steal = nullptr;
  }
  ~vt_cloud() { free_device(); }
};

namespace {

int num_cus(int dev) {
  static std::mutex mu;
  static std::vector<int> cache;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (!cache[dev]) {
    hipDeviceProp_t p;
    HIP_OK(hipGetDeviceProperties(&p, dev));
    cache[dev] = p.multiProcessorCount;
  }
  return cache[dev];
}

// Persistent grid size for a kernel: as many workgroups as fit on the device
// at once, but no more than the work needs.  Occupancy is queried once per
// kernel.
template <class K>
int grid_for(K kernel, int64_t items, int dev) {
  static std::mutex mu;
  static std::map<const void *, int> cache;
  int per_cu = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(reinterpret_cast<const void *>(kernel));
    if (it != cache.end()) per_cu = it->second;
  }
  if (!per_cu) {
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0));
    if (per_cu < 1) per_cu = 1;
    std::lock_guard<std::mutex> lk(mu);
    cache[reinterpret_cast<const void *>(kernel)] = per_cu;
  }
  int64_t full = (int64_t)per_cu * num_cus(dev);
  int64_t need = (items + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min(full, need));
}

// Per-lane device scratch for `lanes` persistent lanes.
int env_int(const char *name, int dflt, int lo, int hi) {
  const char *e = std::getenv(name);
  if (!e || !*e) return dflt;
  const int b = std::atoi(e);
  return (b >= lo && b <= hi) ? b : dflt;
}

// Ray-engine launch parameters (engine.h): lanes of a wave that run their
// walk steps / leaf steps together, and ray partitions (1 or kParts).
// Defaults from a sweep on the 2M-point clustered mesh; VT_WALK_BATCH,
// VT_LEAF_BATCH and VT_RAY_PARTS override them at load, vt_set_tuning at run
// time.
std::atomic<int> g_walk_batch{env_int("VT_WALK_BATCH", 32, 1, 64)};
std::atomic<int> g_leaf_batch{env_int("VT_LEAF_BATCH", 32, 1, 64)};
std::atomic<int> g_ray_parts{env_int("VT_RAY_PARTS", kParts, 1, kParts) >= kParts ? kParts : 1};
// Lanes with nothing left to fetch walk intervals of other lanes' rays,
// recording up to this many boundaries of one interval (0: they do not;
// VT_HELP_EVENTS overrides it at load).
std::atomic<int> g_help_events{env_int("VT_HELP_EVENTS", kHelpEvents, 0, kHelpEvents)};
int help_events() { return g_help_events.load(); }
int walk_batch() { return g_walk_batch.load(); }
int leaf_batch() { return g_leaf_batch.load(); }
int ray_parts() { return g_ray_parts.load(); }

void ensure_scratch(vt_cloud *c, size_t lanes) {
  if (lanes > c->scratch_lanes) {
    for (void *p : {(void *)c->scratch.knn_spill, (void *)c->scratch.walk_tag,
                    (void *)c->scratch.walk_s, (void *)c->scratch.big_id,
                    (void *)c->scratch.big_d, (void *)c->help_ev, (void *)c->help_steal})
      if (p) (void)hipFree(p);
    c->scratch = Scratch{};
    c->scratch_lanes = 0;
    c->help_ev = nullptr;
    c->help_steal = nullptr;
    Scratch s{};
    s.kspill = c->kspill;
    s.wcap = kWalkCap;
    s.lanes = (int)lanes;
    // Every pointer is stored in c->scratch as soon as it is allocated, so a
    // failed allocation leaves nothing unowned.
    c->scratch.knn_spill = s.knn_spill = dev_alloc<uint4>((size_t)s.kspill * lanes);
    c->scratch.walk_tag = s.walk_tag = dev_alloc<uint32_t>((size_t)s.wcap * lanes);
    c->scratch.walk_s = s.walk_s = dev_alloc<double>((size_t)s.wcap * lanes);
    const size_t waves = (lanes + 63) / 64;
    c->scratch.big_id = s.big_id = dev_alloc<uint32_t>(waves * 4 * kMaxQueryK);
    c->scratch.big_d = s.big_d = dev_alloc<double>(waves * 4 * kMaxQueryK);
    c->scratch = s;
    c->scratch_lanes = lanes;
  }
  // Interval buffers: allocated on the first call that walks intervals for
  // other lanes, for every lane of the scratch.
  const int hev = help_events();
  if (hev > 0 && !c->help_steal) {
    if (!c->help_ev) c->help_ev = dev_alloc<uint4>((size_t)kHelpEvents * c->scratch_lanes);
    c->help_steal = dev_alloc<uint4>((size_t)2 * kHelpGifts * c->scratch_lanes);
  }
  c->scratch.help_ev = hev > 0 ? c->help_ev : nullptr;
  c->scratch.help_steal = hev > 0 ? c->help_steal : nullptr;
  c->scratch.help_cap = hev;
  c->scratch.help_steal_cap = kHelpGifts;
  c->scratch.help_depth = 3;
  c->scratch.walk_batch = walk_batch();
  c->scratch.leaf_batch = leaf_batch();
  c->scratch.parts = ray_parts();
}

void ensure_defer(vt_cloud *c, size_t nrays) {
  if (nrays <= c->defer_cap) return;
  if (c->d_defer) (void)hipFree(c->d_defer);
  c->d_defer = nullptr;
  c->defer_cap = 0;
  c->d_defer = dev_alloc<uint32_t>(nrays);
  c->defer_cap = nrays;
}

void reset_call(vt_cloud *c, hipStream_t st) {
  HIP_OK(hipMemsetAsync(c->d_err, 0, sizeof(ErrRec), st));
  HIP_OK(hipMemsetAsync(c->d_stats, 0, sizeof(Stats), st));
}

// Waits for the call, copies the error record and statistics back and turns
// a kernel error into the reference's message.
// With `more`, the statistics are added to those of the call's first launch.
int finish_call(vt_cloud *c, hipStream_t st, hipEvent_t e0, hipEvent_t e1, bool more = false) {
  ErrRec er;
  Stats ss;
  HIP_OK(hipMemcpyAsync(&er, c->d_err, sizeof er, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(&ss, c->d_stats, sizeof ss, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  float ms = 0;
  if (e0 && e1) HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  if (!more) g_stats = vt_call_stats{};
  g_stats.segments += (int64_t)ss.segments;
  g_stats.queries += (int64_t)ss.queries;
  g_stats.max_walk_depth = std::max<int64_t>(g_stats.max_walk_depth, ss.max_walk_depth);
  g_stats.max_knn_depth = std::max<int64_t>(g_stats.max_knn_depth, ss.max_knn_depth);
  g_stats.slow_queries += (int64_t)ss.slow_queries;
  g_stats.deferred_rays += (int64_t)ss.deferred;
  g_stats.kernel_ms += ms;
  g_stats.intervals += (int64_t)ss.intervals;
  g_stats.intervals_back += (int64_t)ss.intervals_back;
  if (!er.flag) return VT_OK;
  auto cell = [&](uint32_t slot) -> size_t {
    return slot < c->tree.vind.size() ? (size_t)c->tree.vind[slot] : (size_t)slot;
  };
  switch (er.code) {
    case kErrMaxIter:
      return fail(VT_ERUNTIME, "Ray walk exceeded maximum iteration count");
    case kErrSplitOutside:
      return fail(VT_ERUNTIME, fmt("Split point outside segment bounds: s=%f not in (%f, %f) "
                                   "for cells %zu and %zu",
                                   er.v[0], er.v[1], er.v[2], cell(er.id[0]), cell(er.id[1])));
    case kErrClassify:
      return fail(VT_ERUNTIME,
                  "Could not resolve split point classification after expanding to 100-NN. "
                  "This should be very unlikely; consider increasing TOLERANCE or K_EXPAND.");
    case kErrBisect:
      return fail(VT_ERUNTIME, fmt("Binary search did not converge for split point between "
                                   "cells %zu and %zu",
                                   cell(er.id[0]), cell(er.id[1])));
    case kErrPerturb:
      return fail(VT_ERUNTIME,
                  "Perturbation did not resolve degenerate split point after 1000 cycles");
    case kErrPeriodicStart:
      return fail(VT_ERUNTIME, "Could not find intermediate cell by bisection along periodic ray");
    case kErrWalkStack:
      return fail(VT_ERUNTIME, fmt("Ray %lld: more than %d pending split points on one ray "
                                   "(walk stack capacity)",
                                   (long long)er.ray, kWalkCap));
    case kErrNonFinite:
      return fail(VT_EINVAL, fmt("Ray %lld: a query point has non-finite coordinates",
                                 (long long)er.ray));
    case kErrFillOverflow:
      return fail(VT_ERUNTIME, fmt("Ray %lld: segment count changed between passes",
                                   (long long)er.ray));
    case kErrHelpWait:
      return fail(VT_ERUNTIME, fmt("Ray %lld: internal error: a lane waited too long for "
                                   "another lane of its wave",
                                   (long long)er.ray));
    case kErrTies:
      return fail(VT_ERUNTIME, fmt("Ray %lld: 100 or more generators equidistant from one "
                                   "split point across periodic images; classification is "
                                   "ill-defined",
                                   (long long)er.ray));
    default:
      return fail(VT_ERUNTIME, fmt("Ray %lld: kNN traversal stack overflow", (long long)er.ray));
  }
}

struct Timer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Timer() {
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
  }
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

int check_ready(const vt_cloud *c) {
  if (!c) return fail(VT_EINVAL, "null cloud");
  if (!c->tree_built)
    return fail(VT_ERUNTIME, "There is currently no valid tree for this point cloud");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != c->device)
    return fail(VT_ERUNTIME, "the cloud's tree lives on another HIP device than the current one");
  return VT_OK;
}

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const std::invalid_argument &e) {
    return fail(VT_EINVAL, e.what());
  } catch (const std::exception &e) {
    return fail(VT_ERUNTIME, e.what());
  }
}

}  // namespace

namespace {

// One call = the fast kernel over every ray, then, if it deferred any ray
// (K_EXPAND), the K_EXPAND-capable kernel over the deferred rays.
// fast(timer) and slow(timer, n_deferred) launch and record the timer's events.
template <class Fast, class Slow>
int two_pass(vt_cloud *c, hipStream_t st, size_t nrays, Fast &&fast, Slow &&slow) {
  if (nrays > 0xFFFFFFFFull) return fail(VT_EINVAL, "at most 2^32 - 1 rays per call");
  ensure_defer(c, nrays);
  reset_call(c, st);
  {
    Timer tm;
    fast(tm);
    const int rc = finish_call(c, st, tm.e0, tm.e1);
    if (rc) return rc;
  }
  const int64_t nd = g_stats.deferred_rays;
  if (nd == 0) return VT_OK;
  reset_call(c, st);
  Timer tm;
  slow(tm, nd);
  return finish_call(c, st, tm.e0, tm.e1, true);
}

// Whether a launch of n rays on `grid` workgroups uses the helping variant of
// its kernel: at most 4 rays per lane (2 for volume, whose helping variant
// spills registers).  Measured at 2M clustered cells: 2 rays per lane 46.8 ->
// 41.5 ms with helping; 8 per lane (volume) 138.0 -> 139.5 ms; 32 per lane
// +6 %, 112 per lane +6 %; 1/8 ray per lane (100k uniform) 8.4 -> 3.6 ms.
bool help_pays(int mode, int64_t n, int grid) {
  const int64_t lanes = (int64_t)grid * kBlock;
  return n <= (mode == VT_VOLUME ? 2 : 4) * lanes;
}

// The ray engine's launches: a full persistent grid, and when there are
// fewer rays than lanes, only the first fetch_lanes lanes of every wave take
// rays, so the rays spread over all CUs instead of filling a few waves.
template <class K, class... A>
void launch(vt_cloud *c, K kernel, int64_t n, hipStream_t st, Timer &tm, A... args) {
  const int grid = grid_for(kernel, INT64_MAX / 2, c->device);
  ensure_scratch(c, (size_t)grid * kBlock);
  const int64_t lanes = (int64_t)grid * kBlock;
  c->scratch.fetch_lanes =
      (int)std::min<int64_t>(64, std::max<int64_t>(1, (64 * n + lanes - 1) / lanes));
  // A lane gives away a pair of its pending list once it has this many
  // entries below its top: any pair when lanes outnumber rays (100k uniform
  // cells, 1/8 ray per lane: 3.58 -> 2.85 ms), only pairs below the entries
  // held in registers otherwise (2 rays per lane: 42.1 ms, against 43.6 ms
  // when giving any pair).
  c->scratch.help_depth = n < lanes ? 1 : 3;
  c->scratch.tile_row = c->ray_row;
  c->scratch.ray_order = (c->order_n == n && n > 0) ? c->d_order : nullptr;
  HIP_OK(hipEventRecord(tm.e0, st));
  kernel<<<grid, kBlock, 0, st>>>(c->dc, c->scratch, args...);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(tm.e1, st));
}

template <int MODE, int NF, bool PER, bool REC>
int project_p(vt_cloud *c, const double *s, const double *e, int64_t n, double *out, int nout,
              int64_t *counts, const RecPool &pool, hipStream_t st) {
  return two_pass(
      c, st, (size_t)n,
      [&](Timer &tm) {
        if (help_pays(MODE, n, grid_for(k_project<MODE, NF, PER, REC>, INT64_MAX / 2,
                                        c->device)))
          launch(c, k_project<MODE, NF, PER, REC, true>, n, st, tm, s, e, n, c->d_defer, out,
                 nout, counts, pool, c->d_err, c->d_stats);
        else
          launch(c, k_project<MODE, NF, PER, REC>, n, st, tm, s, e, n, c->d_defer, out, nout,
                 counts, pool, c->d_err, c->d_stats);
      },
      [&](Timer &tm, int64_t nd) {
        launch(c, k_project_x<MODE, PER, REC>, nd, st, tm, s, e, nd,
               (const uint32_t *)c->d_defer, out, nout, counts, pool, c->d_err, c->d_stats);
      });
}

template <int MODE, int NF, bool REC = false>
int project(vt_cloud *c, const double *s, const double *e, int64_t n, double *out, int nout,
            int64_t *counts, hipStream_t st, const RecPool &pool = RecPool{}) {
  return c->periodic ? project_p<MODE, NF, true, REC>(c, s, e, n, out, nout, counts, pool, st)
                     : project_p<MODE, NF, false, REC>(c, s, e, n, out, nout, counts, pool, st);
}

// Resident lanes (persistent grid x kBlock) of the forward kernel that a
// projection call of `mode` on cloud c launches (the recording variant for
// REC); every launch of that kernel runs this many lanes.
template <int MODE, int NF, bool REC>
int64_t project_lanes(const vt_cloud *c) {
  const int grid = c->periodic
                       ? grid_for(k_project<MODE, NF, true, REC>, INT64_MAX / 2, c->device)
                       : grid_for(k_project<MODE, NF, false, REC>, INT64_MAX / 2, c->device);
  return (int64_t)grid * kBlock;
}

int check_mode(const vt_cloud *c, int mode) {
  if (mode < VT_SUM || mode > VT_VOLUME) return fail(VT_EINVAL, "Unrecognized reduction mode");
  if (mode == VT_VOLUME) {
    if (c->nf != 4)
      return fail(VT_EINVAL, fmt("VolumeRender requires exactly 4 fields (R, G, B, alpha), "
                                 "got %zu",
                                 c->nf));
    if (!c->valid_rgba)
      return fail(VT_EINVAL, "Volume rendering requires R, G, B values in [0, 1] and alpha >= 0");
  }
  return VT_OK;
}

}  // namespace

namespace vt {

// numpy.linspace(a, b, n)[i] (numpy/_core/function_base.py): step =
// (b - a) / (n - 1), y = i * step + a (i / (n - 1) * (b - a) + a when the
// step underflows to 0), the last element b; n == 1 gives a.
__device__ __forceinline__ double np_linspace(double a, double b, int64_t n, int64_t i) {
  if (n > 1 && i == n - 1) return b;
  const double delta = b - a;
  if (n <= 1) return (double)i * delta + a;
  const double div = (double)(n - 1);
  const double step = delta / div;
  return step == 0.0 ? ((double)i / div) * delta + a : (double)i * step + a;
}

// Pixel centres of grid._generate_base_grid (vortrace/grid.py:35-44, 47-83):
// linspace(lo + d/2, hi - d/2, n) with d = (hi - lo) / n.
__device__ __forceinline__ double pixel_centre(double lo, double hi, int64_t n, int64_t i) {
  const double d = (hi - lo) / (double)n;
  return np_linspace(lo + 0.5 * d, hi - 0.5 * d, n, i);
}

// One ray of grid.generate_projection_grid (vortrace/grid.py:100-192) per
// thread: ray i * ny + j of its view starts at (x_i, y_j, bounds[0]) and ends
// at (x_i, y_j, bounds[1]); with a rotation, p -> R (p - c) + c where
// numpy.dot(p - c, R.T) evaluates each row as fma(p2, R_k2, fma(p1, R_k1,
// p0 * R_k0)) (OpenBLAS' x86-64 kernels behind numpy).
__global__ void k_grid_rays(const vt_view *views, const int64_t *offsets, int nviews,
                            int64_t nrays, double *start, double *end) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  int lo = 0, hi = nviews - 1;  // the view holding ray r: offsets[v] <= r < offsets[v + 1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gld(offsets + mid) <= r) lo = mid; else hi = mid - 1;
  }
  const vt_view *v = views + lo;
  const int64_t local = r - gld(offsets + lo);
  const int64_t ny = v->ny;
  const int64_t i = local / ny, j = local - i * ny;
  const double x = pixel_centre(v->extent[0], v->extent[1], v->nx, i);
  const double y = pixel_centre(v->extent[2], v->extent[3], ny, j);
  double p[2][3] = {{x, y, v->bounds[0]}, {x, y, v->bounds[1]}};
  if (v->rotate) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const double a0 = p[e][0] - v->center[0], a1 = p[e][1] - v->center[1],
                   a2 = p[e][2] - v->center[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double *R = v->rot + 3 * k;
        p[e][k] = __fma_rn(a2, R[2], __fma_rn(a1, R[1], a0 * R[0])) + v->center[k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gst(start + 3 * r + k, p[0][k]);
    gst(end + 3 * r + k, p[1][k]);
  }
}

}  // namespace vt

namespace {

// Device buffers of one call.  With a cloud, the i-th request of a call is
// served from the cloud's arena slot i (grown by a quarter more than asked
// when too small, so repeated calls allocate nothing); without one, each
// request is allocated and freed with the call.
struct DevBufs {
  vt_cloud *c = nullptr;
  size_t next = 0;
  std::vector<void *> ptrs;
  explicit DevBufs(vt_cloud *cloud = nullptr) : c(cloud) {}
  template <class T>
  T *get(size_t count) {
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (!c) {
      T *p = dev_alloc<T>(std::max<size_t>(1, count));
      ptrs.push_back(p);
      return p;
    }
    if (next == c->arena.size()) c->arena.push_back({nullptr, 0});
    auto &slot = c->arena[next++];
    if (slot.second < bytes) {
      if (slot.first) (void)hipFree(slot.first);
      slot.first = nullptr;
      slot.second = 0;
      const size_t grown = bytes + bytes / 4;
      slot.first = dev_alloc<char>(grown);
      slot.second = grown;
    }
    return static_cast<T *>(slot.first);
  }
  ~DevBufs() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

int bits_for(size_t n) {
  int b = 1;
  while (b < 32 && ((size_t)1 << b) < n) ++b;
  return b;
}

// Steps after the segment lists of an adjoint call: contributions per
// segment, a stable sort by slot (of the contributions themselves when there
// is one per segment, else of the segment indices), and the per-cell sums in
// ray-major order.  Returns the GPU time of these steps in ms.
float adjoint_tail(vt_cloud *c, int mode, int64_t n, const int64_t *offsets,
                   const uint32_t *slot, const double *ds, int64_t nseg,
                   const double *d_grad_out, double *d_grad_fields, DevBufs &B, hipStream_t st) {
  const int nf = (int)c->nf;
  Timer tm;
  HIP_OK(hipEventRecord(tm.e0, st));
  const int ncf = mode == VT_VOLUME ? 4 : nf;
  double *contrib = B.get<double>((size_t)nseg * ncf);
  const int threads = 256;
  const int rblocks = (int)((n + threads - 1) / threads);
  if (mode == VT_SUM)
    k_contrib_sum<<<(int)((n * 64 + threads - 1) / threads), threads, 0, st>>>(
        n, offsets, ds, d_grad_out, nf, contrib);
  else
    k_contrib_vol<<<rblocks, threads, 0, st>>>(n, offsets, slot, ds, c->d_fields, d_grad_out,
                                               contrib);
  HIP_OK(hipGetLastError());
  const int sblocks = (int)((nseg + threads - 1) / threads);
  const int cblocks = (int)((c->n + threads - 1) / threads);
  const int nb = bits_for(c->n);
  if (ncf == 1) {
    // One value per segment: sort the contributions themselves by slot.
    uint32_t *keys = B.get<uint32_t>(nseg);
    double *sorted = B.get<double>(nseg);
    size_t tmp = 0;
    HIP_OK(rocprim::radix_sort_pairs(nullptr, tmp, slot, keys, contrib, sorted, (size_t)nseg, 0,
                                     nb, st));
    void *t = B.get<char>(tmp);
    HIP_OK(rocprim::radix_sort_pairs(t, tmp, slot, keys, contrib, sorted, (size_t)nseg, 0, nb,
                                     st));
    uint32_t *rs = B.get<uint32_t>(c->n);
    uint32_t *re = B.get<uint32_t>(c->n);
    HIP_OK(hipMemsetAsync(rs, 0, c->n * sizeof(uint32_t), st));
    HIP_OK(hipMemsetAsync(re, 0, c->n * sizeof(uint32_t), st));
    k_run_bounds<<<sblocks, threads, 0, st>>>(nseg, keys, rs, re);
    HIP_OK(hipGetLastError());
    k_reduce_sorted<<<cblocks, threads, 0, st>>>((uint32_t)c->n, rs, re, sorted, c->d_slot_id,
                                                  d_grad_fields);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(tm.e1, st));
    HIP_OK(hipStreamSynchronize(st));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, tm.e0, tm.e1));
    return ms;
  }
  uint32_t *idx = B.get<uint32_t>(nseg);
  uint32_t *keys = B.get<uint32_t>(nseg);
  uint32_t *perm = B.get<uint32_t>(nseg);
  k_iota<<<sblocks, threads, 0, st>>>(nseg, idx);
  HIP_OK(hipGetLastError());
  {
    size_t tmp = 0;
    HIP_OK(rocprim::radix_sort_pairs(nullptr, tmp, slot, keys, idx, perm, (size_t)nseg, 0, nb,
                                     st));
    void *t = B.get<char>(tmp);
    HIP_OK(rocprim::radix_sort_pairs(t, tmp, slot, keys, idx, perm, (size_t)nseg, 0, nb, st));
  }
  uint32_t *rs = B.get<uint32_t>(c->n);
  uint32_t *re = B.get<uint32_t>(c->n);
  HIP_OK(hipMemsetAsync(rs, 0, c->n * sizeof(uint32_t), st));
  HIP_OK(hipMemsetAsync(re, 0, c->n * sizeof(uint32_t), st));
  k_run_bounds<<<sblocks, threads, 0, st>>>(nseg, keys, rs, re);
  HIP_OK(hipGetLastError());
  k_reduce_cells<<<cblocks, threads, 0, st>>>((uint32_t)c->n, ncf, rs, re, perm, contrib,
                                               c->d_slot_id, d_grad_fields);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(tm.e1, st));
  HIP_OK(hipStreamSynchronize(st));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, tm.e0, tm.e1));
  return ms;
}

// Offsets (exclusive scan with the total at [n]) of per-ray segment counts;
// returns the total.
int64_t scan_counts(const int64_t *cnt, size_t nrays, int64_t *offsets, DevBufs &B,
                    hipStream_t st) {
  HIP_OK(hipMemsetAsync(offsets, 0, sizeof(int64_t), st));
  size_t tmp = 0;
  HIP_OK(rocprim::inclusive_scan(nullptr, tmp, cnt, offsets + 1, nrays,
                                 rocprim::plus<int64_t>(), st));
  void *t = B.get<char>(tmp);
  HIP_OK(rocprim::inclusive_scan(t, tmp, cnt, offsets + 1, nrays, rocprim::plus<int64_t>(), st));
  int64_t nseg = 0;
  HIP_OK(hipMemcpyAsync(&nseg, offsets + nrays, sizeof nseg, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  return nseg;
}

}  // namespace

// ------------------------------------------------ no-allocation projection ----
// vt_project_noalloc runs vt_project's kernels with every per-call buffer --
// the deferred-ray list, the kNN stacks below their LDS windows, the walk's
// pending split points and the K_EXPAND buffers -- carved out of one static
// device array, so that the call allocates nothing.  The array is the
// library's own per device (the loader places it on every GPU the library
// runs on, for the life of the process); calls on one GPU share it, so calls
// that overlap in time on one GPU race.  Nothing in it needs clearing: every
// buffer is written before it is read.

namespace vt {

constexpr size_t kNoAllocBytes = size_t(128) << 20;  // 128 MiB per device
constexpr int64_t kNoAllocMaxRays = int64_t(1) << 22;  // deferred list: 16 MiB at most
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Zeroes a call's error record and statistics (a kernel, so the call issues
// no memset).
__global__ void k_clear_call(ErrRec *err, Stats *stats) {
  uint32_t *e = reinterpret_cast<uint32_t *>(err);
  uint32_t *s = reinterpret_cast<uint32_t *>(stats);
  constexpr int ne = sizeof(ErrRec) / 4, ns = sizeof(Stats) / 4;
  static_assert(sizeof(ErrRec) % 4 == 0 && sizeof(Stats) % 4 == 0, "word-sized records");
  for (int i = threadIdx.x; i < ne + ns; i += blockDim.x) {
    if (i < ne)
      e[i] = 0u;
    else
      s[i - ne] = 0u;
  }
}

}  // namespace vt

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char *g_noalloc_base[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (current), looked up once
// per device.  Host code uses only this address, never the array's name.
unsigned char *noalloc_base(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw HipError("device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_base[dev]) {
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(vt::scratch)));
    if (!p) throw HipError("no device address for the no-allocation scratch array");
    g_noalloc_base[dev] = static_cast<unsigned char *>(p);
  }
  return g_noalloc_base[dev];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Splits the array for a call of nrays rays on cloud c into the deferred
// list and the per-lane buffers of `grid` workgroups: as many workgroups as
// fit, at most full_grid.  Returns 0 when not even one fits.
int noalloc_layout(const vt_cloud *c, unsigned char *base, int64_t nrays, int full_grid,
                   Scratch &s, uint32_t *&defer) {
  const size_t kspill = (size_t)c->kspill;
  const size_t defer_b = align256((size_t)nrays * sizeof(uint32_t));
  const size_t per_wg = (size_t)kBlock * (kspill * sizeof(uint4) +
                                         (size_t)kWalkCap * (sizeof(uint32_t) + sizeof(double))) +
                        (size_t)(kBlock / 64) * 4 * kMaxQueryK * (sizeof(uint32_t) + sizeof(double));
  const size_t slack = 5 * 256;  // alignment of the five per-lane arrays
  if (defer_b + slack >= kNoAllocBytes) return 0;
  const int grid = (int)std::min<size_t>((size_t)full_grid, (kNoAllocBytes - defer_b - slack) / per_wg);
  if (grid < 1) return 0;
  const size_t lanes = (size_t)grid * kBlock, waves = lanes / 64;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char *p = base + off;
    off += align256(bytes);
    return p;
  };
  defer = reinterpret_cast<uint32_t *>(take((size_t)nrays * sizeof(uint32_t)));
  s = Scratch{};
  s.kspill = (int)kspill;
  s.wcap = kWalkCap;
  s.lanes = (int)lanes;
  s.knn_spill = reinterpret_cast<uint4 *>(take(kspill * lanes * sizeof(uint4)));
  s.walk_tag = reinterpret_cast<uint32_t *>(take((size_t)kWalkCap * lanes * sizeof(uint32_t)));
  s.walk_s = reinterpret_cast<double *>(take((size_t)kWalkCap * lanes * sizeof(double)));
  s.big_id = reinterpret_cast<uint32_t *>(take(waves * 4 * kMaxQueryK * sizeof(uint32_t)));
  s.big_d = reinterpret_cast<double *>(take(waves * 4 * kMaxQueryK * sizeof(double)));
  if (off > kNoAllocBytes) return 0;  // cannot happen: per_wg and slack bound it
  s.walk_batch = walk_batch();
  s.leaf_batch = leaf_batch();
  s.parts = ray_parts();
  return grid;
}

// vt_project over the no-allocation array: the fast kernel, then the
// K_EXPAND-capable one over the rays it deferred, each on as many
// workgroups as the array holds buffers for.
template <int MODE, int NF, bool PER>
int project_noalloc(vt_cloud *c, const double *d_s, const double *d_e, int64_t n, double *out,
                    int nout, hipStream_t st) {
  unsigned char *base = noalloc_base(c->device);
  auto kfast = k_project<MODE, NF, PER, false>;
  auto kslow = k_project_x<MODE, PER, false>;
  const int full = std::max(grid_for(kfast, INT64_MAX / 2, c->device),
                            grid_for(kslow, INT64_MAX / 2, c->device));
  // One layout for both launches: the per-lane arrays are strided by the
  // lanes of `fit` workgroups; each launch runs at most that many.
  Scratch s;
  uint32_t *defer = nullptr;
  const int fit = noalloc_layout(c, base, n, full, s, defer);
  if (fit < 1)
    return fail(VT_EINVAL, fmt("vt_project_noalloc: the tree's depth leaves no room for one "
                               "workgroup's buffers in the %zu MiB scratch array",
                               kNoAllocBytes >> 20));
  auto prepare = [&](auto kernel, int64_t nr, bool own_order) {
    const int grid = std::min(fit, grid_for(kernel, INT64_MAX / 2, c->device));
    const int64_t lanes = (int64_t)grid * kBlock;
    s.fetch_lanes =
        (int)std::min<int64_t>(64, std::max<int64_t>(1, (64 * nr + lanes - 1) / lanes));
    s.tile_row = c->ray_row;
    s.ray_order = (own_order && c->order_n == nr && nr > 0) ? c->d_order : nullptr;
    k_clear_call<<<1, 64, 0, st>>>(c->d_err, c->d_stats);
    HIP_OK(hipGetLastError());
    return grid;
  };
  {
    const int grid = prepare(kfast, n, true);
    kfast<<<grid, kBlock, 0, st>>>(c->dc, s, d_s, d_e, n, defer, out, nout, nullptr, RecPool{},
                                   c->d_err, c->d_stats);
    HIP_OK(hipGetLastError());
    const int rc = finish_call(c, st, nullptr, nullptr);
    if (rc) return rc;
  }
  const int64_t nd = g_stats.deferred_rays;
  if (nd == 0) return VT_OK;
  const int grid = prepare(kslow, nd, false);
  kslow<<<grid, kBlock, 0, st>>>(c->dc, s, d_s, d_e, nd, (const uint32_t *)defer, out, nout,
                                 nullptr, RecPool{}, c->d_err, c->d_stats);
  HIP_OK(hipGetLastError());
  return finish_call(c, st, nullptr, nullptr, true);
}

template <int MODE, int NF>
int project_noalloc_any(vt_cloud *c, const double *s, const double *e, int64_t n, double *out,
                        int nout, hipStream_t st) {
  return c->periodic ? project_noalloc<MODE, NF, true>(c, s, e, n, out, nout, st)
                     : project_noalloc<MODE, NF, false>(c, s, e, n, out, nout, st);
}

}  // namespace

extern "C" {

const char *vt_last_error(void) { return g_err.c_str(); }

void vt_set_warning_callback(vt_warning_fn fn, void *user) {
  std::lock_guard<std::mutex> lk(g_warn_mu);
  g_warn_fn = fn;
  g_warn_user = user;
}

void vt_set_verbose(int v) { g_verbose.store(v ? 1 : 0); }
int vt_get_verbose(void) { return g_verbose.load(); }

int vt_cloud_create(const double *pos, size_t npart, const double *fields, size_t nfields,
                    const double subbox[6], const double *vol, size_t nvol, int periodic,
                    vt_cloud **out) {
  return guarded([&]() -> int {
    if (!out || !subbox) return fail(VT_EINVAL, "null argument");
    if (npart && (!pos || !fields)) return fail(VT_EINVAL, "null argument");
    auto c = std::make_unique<vt_cloud>();
    c->nf = nfields;
    std::memcpy(c->subbox, subbox, sizeof c->subbox);
    c->periodic = periodic ? 1 : 0;
    if (c->periodic) {
      c->box[0] = subbox[1] - subbox[0];
      c->box[1] = subbox[3] - subbox[2];
      c->box[2] = subbox[5] - subbox[4];
    }
    // src/pointcloud.cpp:38-53
    if (nvol > 0) {
      double min_vol = std::numeric_limits<double>::max();
      for (size_t i = 0; i < nvol; i++)
        if (vol[i] > 0 && vol[i] < min_vol) min_vol = vol[i];
      if (min_vol < std::numeric_limits<double>::max()) {
        double min_radius = std::cbrt(3.0 * min_vol / (4.0 * M_PI));
        if (min_radius < 1e-6)
          warn("Some cells have radii smaller than 1e-6, which approaches the internal "
               "ray-tracing tolerance (1e-9). Results may be unreliable for very small cells. "
               "Consider rescaling your coordinate system.");
      }
    }
    if (!c->periodic) {
      // Padded filter, src/pointcloud.cpp:55-108.
      const double dx = subbox[1] - subbox[0], dy = subbox[3] - subbox[2],
                   dz = subbox[5] - subbox[4];
      if (nvol > 0) {
        double max_vol = 0.0;
        for (size_t i = 0; i < nvol; i++)
          if (vol[i] > max_vol) max_vol = vol[i];
        c->pad = 10.0 * std::cbrt(3.0 * max_vol / (4.0 * M_PI));
      } else {
        c->pad = 0.15 * std::max({dx, dy, dz});
        warn("No cell volumes provided; using 15% of longest box side as padding. Pass vol= "
             "for adaptive padding.");
      }
      const double lo[3] = {subbox[0] - c->pad, subbox[2] - c->pad, subbox[4] - c->pad};
      const double hi[3] = {subbox[1] + c->pad, subbox[3] + c->pad, subbox[5] + c->pad};
      for (size_t i = 0; i < npart; i++) {
        const double *p = pos + 3 * i;
        if (p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] &&
            p[2] <= hi[2]) {
          c->pts.insert(c->pts.end(), p, p + 3);
          c->fields.insert(c->fields.end(), fields + i * nfields, fields + (i + 1) * nfields);
          c->orig_ids.push_back((int64_t)i);
        }
      }
    } else {
      // All points, validated against the periodic domain, src/pointcloud.cpp:109-163.
      double pmin[3], pmax[3];
      for (int d = 0; d < 3; d++) {
        pmin[d] = std::numeric_limits<double>::max();
        pmax[d] = std::numeric_limits<double>::lowest();
      }
      c->pts.assign(pos, pos + 3 * npart);
      c->fields.assign(fields, fields + npart * nfields);
      c->orig_ids.resize(npart);
      for (size_t i = 0; i < npart; i++) {
        c->orig_ids[i] = (int64_t)i;
        for (int d = 0; d < 3; d++) {
          pmin[d] = std::min(pmin[d], pos[3 * i + d]);
          pmax[d] = std::max(pmax[d], pos[3 * i + d]);
        }
      }
      if (pmin[0] < subbox[0] || pmax[0] > subbox[1] || pmin[1] < subbox[2] ||
          pmax[1] > subbox[3] || pmin[2] < subbox[4] || pmax[2] > subbox[5])
        return fail(VT_EINVAL,
                    "Periodic mode: some particles lie outside the bounding box. All particles "
                    "must be within the periodic domain.");
      for (int d = 0; d < 3; d++)
        if (pmax[d] - pmin[d] < 0.6 * c->box[d]) {
          warn("Periodic mode: particle extent is less than 60% of the bounding box in at "
               "least one dimension. The bounding box may be too large for the particle "
               "distribution.");
          break;
        }
    }
    c->n = c->orig_ids.size();
    // src/pointcloud.cpp:167-179
    c->valid_rgba = 0;
    if (c->nf == 4 && c->n > 0) {
      c->valid_rgba = 1;
      for (size_t i = 0; i < c->n && c->valid_rgba; i++) {
        for (int k = 0; k < 3; k++) {
          const double v = c->fields[i * 4 + k];
          if (v < 0.0 || v > 1.0) {
            c->valid_rgba = 0;
            break;
          }
        }
        if (c->valid_rgba && c->fields[i * 4 + 3] < 0.0) c->valid_rgba = 0;
      }
    }
    *out = c.release();
    return VT_OK;
  });
}

int vt_cloud_build_host_tree(vt_cloud *c) {
  return guarded([&]() -> int {
    if (!c) return fail(VT_EINVAL, "null cloud");
    if (c->n == 0) return fail(VT_ERUNTIME, "There are no points in the cloud; cannot build tree");
    const auto t0 = std::chrono::steady_clock::now();
    const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    HostTree t;
    build_tree(c->pts.data(), c->n, threads, t);
    c->build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->free_device();
    c->tree_built = false;
    c->host_tree = true;
    c->tree = std::move(t);
    if (g_verbose.load())
      std::printf("Tree build took %lld milliseconds.\n", (long long)(c->build_seconds * 1e3));
    return VT_OK;
  });
}

int vt_cloud_upload(vt_cloud *c) {
  return guarded([&]() -> int {
    if (!c) return fail(VT_EINVAL, "null cloud");
    if (!c->host_tree)
      return fail(VT_ERUNTIME, "There is currently no valid tree for this point cloud");
    const HostTree &t = c->tree;
    DeviceTree dt;
    make_device_tree(t, dt);
    // Slot-ordered generators (with their leaf's end slot) and fields.
    std::vector<Gen> gen(c->n);
    std::vector<double> fslot(c->n * c->nf);
    for (size_t s = 0; s < c->n; ++s) {
      const size_t i = t.vind[s];
      gen[s] = Gen{c->pts[3 * i], c->pts[3 * i + 1], c->pts[3 * i + 2], dt.leaf_start[s],
                   dt.leaf_end[s]};
      std::memcpy(&fslot[s * c->nf], &c->fields[i * c->nf], c->nf * sizeof(double));
    }
    c->free_device();
    c->tree_built = false;
    HIP_OK(hipGetDevice(&c->device));
    c->d_nodes = dev_alloc<DNode>(std::max<size_t>(1, dt.nodes.size()));
    c->d_recs = dev_alloc<DRec>(std::max<size_t>(1, dt.recs.size()));
    // kLeafMaxSize zero records of padding: leaf steps load whole chunks.
    c->d_gen = dev_alloc<Gen>(c->n + kLeafMaxSize);
    c->d_fields = dev_alloc<double>(fslot.size());
    c->d_slot_id = dev_alloc<uint32_t>(c->n);
    c->d_err = dev_alloc<ErrRec>(1);
    c->d_stats = dev_alloc<Stats>(1);
    if (!dt.nodes.empty())
      HIP_OK(hipMemcpy(c->d_nodes, dt.nodes.data(), dt.nodes.size() * sizeof(DNode),
                       hipMemcpyHostToDevice));
    if (!dt.recs.empty())
      HIP_OK(hipMemcpy(c->d_recs, dt.recs.data(), dt.recs.size() * sizeof(DRec),
                       hipMemcpyHostToDevice));
    gen.resize(c->n + kLeafMaxSize, Gen{0.0, 0.0, 0.0, 0xFFFFFFFFu, 0xFFFFFFFFu});
    HIP_OK(hipMemcpy(c->d_gen, gen.data(), gen.size() * sizeof(Gen), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_fields, fslot.data(), std::max<size_t>(1, fslot.size()) * sizeof(double),
                     hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_slot_id, t.vind.data(), c->n * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
    // kNN stack: at most one entry per level of the current path.
    c->kspill = (int)t.max_depth + 2;
    DevCloud &d = c->dc;
    d.nodes = c->d_nodes;
    d.recs = c->d_recs;
    d.gen = c->d_gen;
    d.fields = c->d_fields;
    for (int k = 0; k < 3; ++k) {
      d.rb_lo[k] = t.bbox[k][0];
      d.rb_hi[k] = t.bbox[k][1];
      d.box[k] = c->box[k];
    }
    std::memcpy(d.sb, c->subbox, sizeof d.sb);
    d.inv_sqrt2 = 1.0 / std::sqrt(2.0);
    d.n = (uint32_t)c->n;
    d.nf = (uint32_t)c->nf;
    d.periodic = c->periodic;
    d.root = dt.root;
    d.rroot = dt.rroot;
    c->tree_built = true;
    return VT_OK;
  });
}

int vt_cloud_build_tree(vt_cloud *c) {
  int rc = vt_cloud_build_host_tree(c);
  return rc ? rc : vt_cloud_upload(c);
}

int vt_cloud_tree_export(const vt_cloud *c, int64_t *vind, int32_t *feat, int64_t *a,
                         int64_t *b, double *divlow, double *divhigh, double *bbox6) {
  if (!c) return fail(VT_EINVAL, "null cloud");
  if (!c->host_tree)
    return fail(VT_ERUNTIME, "There is currently no valid tree for this point cloud");
  const HostTree &t = c->tree;
  for (size_t i = 0; i < c->n; ++i) vind[i] = (int64_t)t.vind[i];
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    const Node &nd = t.nodes[i];
    if (nd.b & kLeafBit) {
      feat[i] = -1;
      a[i] = nd.a;
      b[i] = nd.b & ~kLeafBit;
      divlow[i] = divhigh[i] = 0.0;
    } else {
      feat[i] = (int32_t)nd.b;
      a[i] = (int64_t)i + 1;
      b[i] = nd.a;
      divlow[i] = nd.lo;
      divhigh[i] = nd.hi;
    }
  }
  for (int d = 0; d < 3; ++d) {
    bbox6[2 * d] = t.bbox[d][0];
    bbox6[2 * d + 1] = t.bbox[d][1];
  }
  return VT_OK;
}

void vt_cloud_destroy(vt_cloud *c) { delete c; }

size_t vt_cloud_npart(const vt_cloud *c) { return c ? c->n : 0; }
size_t vt_cloud_nfields(const vt_cloud *c) { return c ? c->nf : 0; }
int vt_cloud_periodic(const vt_cloud *c) { return c ? c->periodic : 0; }
int vt_cloud_valid_rgba(const vt_cloud *c) { return c ? c->valid_rgba : 0; }
int vt_cloud_tree_built(const vt_cloud *c) { return c ? (int)c->tree_built : 0; }
double vt_cloud_pad(const vt_cloud *c) { return c ? c->pad : 0.0; }
void vt_cloud_subbox(const vt_cloud *c, double out6[6]) {
  if (c && out6) std::memcpy(out6, c->subbox, sizeof c->subbox);
}
void vt_cloud_orig_ids(const vt_cloud *c, int64_t *out) {
  if (c && out) std::memcpy(out, c->orig_ids.data(), c->n * sizeof(int64_t));
}

int vt_cloud_tree_info(const vt_cloud *c, vt_tree_info *out) {
  if (!c || !out) return fail(VT_EINVAL, "null argument");
  if (!c->host_tree)
    return fail(VT_ERUNTIME, "There is currently no valid tree for this point cloud");
  out->nodes = (int64_t)c->tree.nodes.size();
  out->leaves = c->tree.leaves;
  out->max_depth = c->tree.max_depth;
  out->build_seconds = c->build_seconds;
  return VT_OK;
}

int vt_query_knn(const vt_cloud *cc, const double *d_q, size_t nq, int k, int64_t *d_ids,
                 double *d_r2, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (rc) return rc;
    if (k < 1 || k > kMaxQueryK) return fail(VT_EINVAL, "k must be in [1, 128]");
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nq == 0) return VT_OK;
    reset_call(c, st);
    Timer tm;
    if (k <= 3) {
      auto go = [&](auto kernel) {
        const int grid = grid_for(kernel, nq, c->device);
        ensure_scratch(c, (size_t)grid * kBlock);
        HIP_OK(hipEventRecord(tm.e0, st));
        kernel<<<grid, kBlock, 0, st>>>(c->dc, c->scratch, d_q, (int64_t)nq, c->d_slot_id,
                                        d_ids, d_r2, c->d_err, c->d_stats);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(tm.e1, st));
      };
      if (c->periodic) {
        if (k == 1) go(k_query_reg<1, true>);
        if (k == 2) go(k_query_reg<2, true>);
        if (k == 3) go(k_query_reg<3, true>);
      } else {
        if (k == 1) go(k_query_reg<1, false>);
        if (k == 2) go(k_query_reg<2, false>);
        if (k == 3) go(k_query_reg<3, false>);
      }
      return finish_call(c, st, tm.e0, tm.e1);
    }
    auto kmem = c->periodic ? k_query_mem<true> : k_query_mem<false>;
    const int grid = grid_for(kmem, nq, c->device);
    const size_t lanes = (size_t)grid * kBlock;
    ensure_scratch(c, lanes);
    uint32_t *wid = dev_alloc<uint32_t>(lanes * 4 * (size_t)k);
    double *wd = dev_alloc<double>(lanes * 4 * (size_t)k);
    HIP_OK(hipEventRecord(tm.e0, st));
    kmem<<<grid, kBlock, 0, st>>>(c->dc, c->scratch, d_q, (int64_t)nq, k, wid, wd,
                                  c->d_slot_id, d_ids, d_r2, c->d_err, c->d_stats);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(tm.e1, st));
    rc = finish_call(c, st, tm.e0, tm.e1);
    (void)hipFree(wid);
    (void)hipFree(wd);
    return rc;
  });
}


int vt_project(const vt_cloud *cc, const double *d_start, const double *d_end, size_t nrays,
               int mode, double *d_out, void *stream) {
  return vt_project_counted(cc, d_start, d_end, nrays, mode, d_out, nullptr, stream);
}

int vt_project_counted(const vt_cloud *cc, const double *d_start, const double *d_end,
                       size_t nrays, int mode, double *d_out, int64_t *d_counts, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nrays == 0) return VT_OK;
    const int nout = mode == VT_VOLUME ? 3 : (int)c->nf;
    const int64_t n = (int64_t)nrays;
    if (mode == VT_SUM) {
      switch (c->nf) {
        case 1: return project<VT_SUM, 1>(c, d_start, d_end, n, d_out, nout, d_counts, st);
        case 2: return project<VT_SUM, 2>(c, d_start, d_end, n, d_out, nout, d_counts, st);
        case 3: return project<VT_SUM, 3>(c, d_start, d_end, n, d_out, nout, d_counts, st);
        case 4: return project<VT_SUM, 4>(c, d_start, d_end, n, d_out, nout, d_counts, st);
        default: return project<VT_SUM, 0>(c, d_start, d_end, n, d_out, nout, d_counts, st);
      }
    }
    if (mode == VT_MAX) return project<VT_MAX, 0>(c, d_start, d_end, n, d_out, nout, d_counts, st);
    if (mode == VT_MIN) return project<VT_MIN, 0>(c, d_start, d_end, n, d_out, nout, d_counts, st);
    return project<VT_VOLUME, 4>(c, d_start, d_end, n, d_out, nout, d_counts, st);
  });
}

int vt_project_noalloc(const vt_cloud *cc, const double *d_start, const double *d_end,
                       size_t nrays, int mode, double *d_out, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    if (nrays > (size_t)kNoAllocMaxRays)
      return fail(VT_EINVAL, fmt("vt_project_noalloc: %zu rays; at most %lld per call (its "
                                 "buffers live in a fixed %zu MiB device array)",
                                 nrays, (long long)kNoAllocMaxRays, kNoAllocBytes >> 20));
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nrays == 0) return VT_OK;
    const int nout = mode == VT_VOLUME ? 3 : (int)c->nf;
    const int64_t n = (int64_t)nrays;
    if (mode == VT_SUM) {
      switch (c->nf) {
        case 1: return project_noalloc_any<VT_SUM, 1>(c, d_start, d_end, n, d_out, nout, st);
        case 2: return project_noalloc_any<VT_SUM, 2>(c, d_start, d_end, n, d_out, nout, st);
        case 3: return project_noalloc_any<VT_SUM, 3>(c, d_start, d_end, n, d_out, nout, st);
        case 4: return project_noalloc_any<VT_SUM, 4>(c, d_start, d_end, n, d_out, nout, st);
        default: return project_noalloc_any<VT_SUM, 0>(c, d_start, d_end, n, d_out, nout, st);
      }
    }
    if (mode == VT_MAX)
      return project_noalloc_any<VT_MAX, 0>(c, d_start, d_end, n, d_out, nout, st);
    if (mode == VT_MIN)
      return project_noalloc_any<VT_MIN, 0>(c, d_start, d_end, n, d_out, nout, st);
    return project_noalloc_any<VT_VOLUME, 4>(c, d_start, d_end, n, d_out, nout, st);
  });
}

static_assert(sizeof(SegRec) == VT_SEGREC_BYTES, "record size of the C interface");

int vt_project_recorded(const vt_cloud *cc, const double *d_start, const double *d_end,
                        size_t nrays, int mode, double *d_out, int64_t *d_counts,
                        void *d_records, size_t capacity, size_t *n_records, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    if (mode != VT_SUM && mode != VT_VOLUME)
      return fail(VT_EINVAL, "recorded projections are for reductions 'sum' and 'volume'");
    if (!d_counts || !n_records || (!d_records && capacity))
      return fail(VT_EINVAL, "null argument");
    if (reinterpret_cast<uintptr_t>(d_records) % alignof(SegRec))
      return fail(VT_EINVAL, "the record pool must be 16-byte aligned");
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    *n_records = 0;
    if (nrays == 0) return VT_OK;
    const int nout = mode == VT_VOLUME ? 3 : (int)c->nf;
    const int64_t n = (int64_t)nrays;
    DevBufs B(c);
    unsigned long long *head = B.get<unsigned long long>(1);
    int *overflow = B.get<int>(1);
    HIP_OK(hipMemsetAsync(head, 0, sizeof *head, st));
    HIP_OK(hipMemsetAsync(overflow, 0, sizeof *overflow, st));
    const RecPool pool{static_cast<SegRec *>(d_records), (unsigned long long)capacity, head,
                       overflow};
    if (mode == VT_SUM) {
      switch (c->nf) {
        case 1: rc = project<VT_SUM, 1, true>(c, d_start, d_end, n, d_out, nout, d_counts, st, pool); break;
        default: rc = project<VT_SUM, 0, true>(c, d_start, d_end, n, d_out, nout, d_counts, st, pool); break;
      }
    } else {
      rc = project<VT_VOLUME, 4, true>(c, d_start, d_end, n, d_out, nout, d_counts, st, pool);
    }
    if (rc) return rc;
    unsigned long long used = 0;
    int over = 0;
    HIP_OK(hipMemcpyAsync(&used, head, sizeof used, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(&over, overflow, sizeof over, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    // Records fill whole chunks; on overflow the caller falls back to vt_adjoint.
    *n_records = over ? (size_t)-1 : (size_t)std::min<unsigned long long>(used, capacity);
    return VT_OK;
  });
}

int vt_project_count(const vt_cloud *cc, const double *d_start, const double *d_end,
                     size_t nrays, int64_t *d_counts, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (rc) return rc;
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nrays == 0) return VT_OK;
    const int64_t n = (int64_t)nrays;
    auto go = [&](auto kfast, auto kslow) {
      return two_pass(
          c, st, nrays,
          [&](Timer &tm) {
            launch(c, kfast, n, st, tm, d_start, d_end, n, (const uint32_t *)nullptr,
                   c->d_defer, d_counts, c->d_err, c->d_stats);
          },
          [&](Timer &tm, int64_t nd) {
            launch(c, kslow, nd, st, tm, d_start, d_end, nd, (const uint32_t *)c->d_defer,
                   (uint32_t *)nullptr, d_counts, c->d_err, c->d_stats);
          });
    };
    const bool help = help_pays(VT_SUM, n, grid_for(k_count<false, false>, INT64_MAX / 2,
                                                    c->device));
    if (c->periodic)
      return help ? go(k_count<true, false, true>, k_count<true, true>)
                  : go(k_count<true, false>, k_count<true, true>);
    return help ? go(k_count<false, false, true>, k_count<false, true>)
                : go(k_count<false, false>, k_count<false, true>);
  });
}

int vt_project_segments(const vt_cloud *cc, const double *d_start, const double *d_end,
                        size_t nrays, int mode, const int64_t *d_offsets, double *d_out,
                        int64_t *d_cell_ids, double *d_s_enter, double *d_s_exit,
                        void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nrays == 0) return VT_OK;
    const int nout = mode == VT_VOLUME ? 3 : (int)c->nf;
    const int64_t n = (int64_t)nrays;
    auto go = [&](auto kfast, auto kslow) {
      return two_pass(
          c, st, nrays,
          [&](Timer &tm) {
            launch(c, kfast, n, st, tm, d_start, d_end, n, (const uint32_t *)nullptr,
                   c->d_defer, d_offsets, d_out, nout, (const uint32_t *)c->d_slot_id,
                   d_cell_ids, d_s_enter, d_s_exit, c->d_err, c->d_stats);
          },
          [&](Timer &tm, int64_t nd) {
            launch(c, kslow, nd, st, tm, d_start, d_end, nd, (const uint32_t *)c->d_defer,
                   (uint32_t *)nullptr, d_offsets, d_out, nout, (const uint32_t *)c->d_slot_id,
                   d_cell_ids, d_s_enter, d_s_exit, c->d_err, c->d_stats);
          });
    };
    if (c->periodic) {
      if (mode == VT_SUM) return go(k_fill<VT_SUM, true, false>, k_fill<VT_SUM, true, true>);
      if (mode == VT_MAX) return go(k_fill<VT_MAX, true, false>, k_fill<VT_MAX, true, true>);
      if (mode == VT_MIN) return go(k_fill<VT_MIN, true, false>, k_fill<VT_MIN, true, true>);
      return go(k_fill<VT_VOLUME, true, false>, k_fill<VT_VOLUME, true, true>);
    }
    if (mode == VT_SUM) return go(k_fill<VT_SUM, false, false>, k_fill<VT_SUM, false, true>);
    if (mode == VT_MAX) return go(k_fill<VT_MAX, false, false>, k_fill<VT_MAX, false, true>);
    if (mode == VT_MIN) return go(k_fill<VT_MIN, false, false>, k_fill<VT_MIN, false, true>);
    return go(k_fill<VT_VOLUME, false, false>, k_fill<VT_VOLUME, false, true>);
  });
}

int vt_cloud_save_tree(const vt_cloud *c, uint8_t *buf, size_t cap, size_t *size) {
  return guarded([&]() -> int {
    if (!c || !size) return fail(VT_EINVAL, "null argument");
    if (!c->host_tree) return fail(VT_ERUNTIME, "Cannot save tree: tree has not been built");
    std::vector<uint8_t> bytes;
    save_tree_bytes(c->tree, bytes);
    *size = bytes.size();
    if (buf && cap >= bytes.size()) std::memcpy(buf, bytes.data(), bytes.size());
    return VT_OK;
  });
}

int vt_cloud_load_tree(vt_cloud *c, const uint8_t *buf, size_t size, int upload) {
  return guarded([&]() -> int {
    if (!c || (!buf && size)) return fail(VT_EINVAL, "null argument");
    if (c->n == 0) return fail(VT_ERUNTIME, "Cannot load tree: no points loaded");
    HostTree t;
    load_tree_bytes(buf, size, c->n, t);
    c->free_device();
    c->tree_built = false;
    c->host_tree = true;
    c->build_seconds = 0.0;
    c->tree = std::move(t);
    return upload ? vt_cloud_upload(c) : VT_OK;
  });
}

int vt_slice(const vt_cloud *cc, const double extent[4], size_t nx, size_t ny, double depth,
             double *d_out, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (rc) return rc;
    if (!extent || (!d_out && nx * ny)) return fail(VT_EINVAL, "null argument");
    const double *sb = cc->subbox;
    if ((extent[0] < sb[0]) || (extent[1] > sb[1]) || (extent[2] < sb[2]) ||
        (extent[3] > sb[3]) || (depth < sb[4]) || (depth > sb[5]))
      return fail(VT_ERUNTIME, "Slice extent out of bounds of current cloud subbox");
    if (nx == 0 || ny == 0) return VT_OK;
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    // src/slice.cpp:26-31: deltax = (xmax - xmin) / (npix_x - 1)
    const double dx = (extent[1] - extent[0]) / (double)(nx - 1);
    const double dy = (extent[3] - extent[2]) / (double)(ny - 1);
    const int64_t np = (int64_t)(nx * ny);
    reset_call(c, st);
    Timer tm;
    auto go = [&](auto kernel) {
      const int grid = grid_for(kernel, np, c->device);
      ensure_scratch(c, (size_t)grid * kBlock);
      HIP_OK(hipEventRecord(tm.e0, st));
      kernel<<<grid, kBlock, 0, st>>>(c->dc, c->scratch, extent[0], dx, extent[2], dy, depth,
                                      (int64_t)nx, (int64_t)ny, d_out, c->d_err, c->d_stats);
      HIP_OK(hipGetLastError());
      HIP_OK(hipEventRecord(tm.e1, st));
    };
    if (c->periodic)
      go(k_slice<true>);
    else
      go(k_slice<false>);
    return finish_call(c, st, tm.e0, tm.e1);
  });
}

int vt_cloud_set_ray_rows(vt_cloud *c, int64_t row) {
  if (!c) return fail(VT_EINVAL, "null cloud");
  if (row < 0 || row > 0x7FFFFFFF) return fail(VT_EINVAL, "row length out of range [0, 2^31)");
  c->ray_row = (uint32_t)row;
  return VT_OK;
}

int vt_cloud_set_ray_order(vt_cloud *c, const uint32_t *d_order, int64_t n, void *stream) {
  return guarded([&]() -> int {
    if (!c) return fail(VT_EINVAL, "null cloud");
    if (n < 0 || n > 0xFFFFFFFFll) return fail(VT_EINVAL, "ray count out of range [0, 2^32)");
    c->order_n = 0;
    if (n == 0 || !d_order) return VT_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (c->order_cap < (size_t)n) {
      if (c->d_order) (void)hipFree(c->d_order);
      c->d_order = nullptr;
      c->order_cap = 0;
      c->d_order = dev_alloc<uint32_t>((size_t)n);
      c->order_cap = (size_t)n;
    }
    HIP_OK(hipMemcpyAsync(c->d_order, d_order, (size_t)n * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, st));
    DevBufs B(c);
    uint32_t *seen = B.get<uint32_t>((size_t)n);
    int *bad = B.get<int>(1);
    HIP_OK(hipMemsetAsync(seen, 0, (size_t)n * sizeof(uint32_t), st));
    HIP_OK(hipMemsetAsync(bad, 0, sizeof(int), st));
    const int threads = 256;
    k_check_order<<<(unsigned)((n + threads - 1) / threads), threads, 0, st>>>(n, c->d_order,
                                                                               seen, bad);
    HIP_OK(hipGetLastError());
    int invalid = 0;
    HIP_OK(hipMemcpyAsync(&invalid, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (invalid) return fail(VT_EINVAL, "ray order is not a permutation of [0, n)");
    c->order_n = n;
    return VT_OK;
  });
}

int vt_cloud_order_rays(vt_cloud *c, const double *d_start, const double *d_end,
                        const int64_t *d_counts, int64_t n, int64_t row, int64_t block,
                        uint32_t *d_order_out, int64_t *d_run_keys, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(c);
    if (rc) return rc;
    if (n < 0 || n > 0xFFFFFFFFll) return fail(VT_EINVAL, "ray count out of range [0, 2^32)");
    if (row < 0 || row > 0x7FFFFFFF) return fail(VT_EINVAL, "row length out of range [0, 2^31)");
    if (block < 1 || block > 64 * kOrderMaxRuns)
      return fail(VT_EINVAL, fmt("block out of range [1, %d]", 64 * kOrderMaxRuns));
    if (!d_counts && n > 0 && (!d_start || !d_end)) return fail(VT_EINVAL, "null argument");
    c->order_n = 0;
    if (n == 0) return VT_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (c->order_cap < (size_t)n) {
      if (c->d_order) (void)hipFree(c->d_order);
      c->d_order = nullptr;
      c->order_cap = 0;
      c->d_order = dev_alloc<uint32_t>((size_t)n);
      c->order_cap = (size_t)n;
    }
    RunPlan P;
    P.n = n;
    P.parts = ray_parts();
    P.psize = (n + P.parts - 1) / P.parts;
    P.block = block;
    P.per_part = (P.psize + block - 1) / block;
    P.per_block = (block + 63) / 64;
    P.row = (uint32_t)row;
    const uint64_t band = 8ull * (uint64_t)row;
    P.nfull = row ? (uint32_t)((uint64_t)n / band * band) : 0u;
    const int64_t nblocks = (int64_t)P.parts * P.per_part;
    const int64_t nruns = nblocks * P.per_block;
    DevBufs B(c);
    int64_t *keys = d_run_keys ? d_run_keys : B.get<int64_t>((size_t)nruns);
    double lo[3], sz[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = c->subbox[2 * d];
      sz[d] = std::max(c->subbox[2 * d + 1] - c->subbox[2 * d], 1e-300);
    }
    if (!d_counts && !c->d_order_grid) {
      constexpr int cells = kOrderGrid * kOrderGrid * kOrderGrid;
      uint32_t *hist = dev_alloc<uint32_t>(cells);
      float *grid = nullptr;
      try {
        grid = dev_alloc<float>(cells);
      } catch (...) {
        (void)hipFree(hist);
        throw;
      }
      c->d_order_hist = hist;
      c->d_order_grid = grid;
      HIP_OK(hipMemsetAsync(hist, 0, cells * sizeof(uint32_t), st));
      k_order_hist<<<(unsigned)((c->n + 255) / 256), 256, 0, st>>>(
          c->d_gen, (uint32_t)c->n, lo[0], lo[1], lo[2], sz[0], sz[1], sz[2], hist);
      HIP_OK(hipGetLastError());
      k_order_cbrt<<<(cells + 255) / 256, 256, 0, st>>>(hist, grid);
      HIP_OK(hipGetLastError());
    }
    float flo[3], fsz[3];
    for (int d = 0; d < 3; ++d) {
      flo[d] = (float)lo[d];
      fsz[d] = std::max((float)sz[d], 1e-30f);
    }
    const float diag = std::sqrt(fsz[0] * fsz[0] + fsz[1] * fsz[1] + fsz[2] * fsz[2]);
    k_order_keys<<<(unsigned)((nruns + 255) / 256), 256, 0, st>>>(
        P, nruns, d_counts, d_start, d_end, c->d_order_grid, flo[0], flo[1], flo[2], fsz[0],
        fsz[1], fsz[2], 1.0f / diag, c->periodic, keys);
    HIP_OK(hipGetLastError());
    k_order_write<<<(unsigned)nblocks, 256, 0, st>>>(P, keys, c->d_order);
    HIP_OK(hipGetLastError());
    if (d_order_out)
      HIP_OK(hipMemcpyAsync(d_order_out, c->d_order, (size_t)n * sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, st));
    c->order_n = n;
    return VT_OK;
  });
}

int vt_cloud_set_fields(vt_cloud *c, const double *d_fields, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->n == 0) return VT_OK;
    if (!d_fields) return fail(VT_EINVAL, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevBufs B(c);
    int *flag = B.get<int>(1);
    HIP_OK(hipMemsetAsync(flag, 0, sizeof(int), st));
    const int threads = 256;
    const int blocks = (int)((c->n + threads - 1) / threads);
    k_set_fields<<<blocks, threads, 0, st>>>(d_fields, c->d_slot_id, (uint32_t)c->n,
                                             (uint32_t)c->nf, c->d_fields, flag);
    HIP_OK(hipGetLastError());
    int invalid = 0;
    HIP_OK(hipMemcpyAsync(&invalid, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    c->valid_rgba = (c->nf == 4 && !invalid) ? 1 : 0;
    return VT_OK;
  });
}


int vt_adjoint(const vt_cloud *cc, const double *d_start, const double *d_end, size_t nrays,
               int mode, const int64_t *d_counts, const double *d_grad_out,
               double *d_grad_fields, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    if (mode != VT_SUM && mode != VT_VOLUME)
      return fail(VT_EINVAL, "the adjoint is defined for reductions 'sum' and 'volume'");
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nf = (int)c->nf;
    HIP_OK(hipMemsetAsync(d_grad_fields, 0, c->n * nf * sizeof(double), st));
    if (nrays == 0) return VT_OK;
    const int64_t n = (int64_t)nrays;
    DevBufs B(c);
    // 1. segments per ray, and offsets
    int64_t *counts = B.get<int64_t>(nrays);
    int64_t *offsets = B.get<int64_t>(nrays + 1);
    auto count = [&](auto kfast, auto kslow) {
      return two_pass(
          c, st, nrays,
          [&](Timer &tm) {
            launch(c, kfast, n, st, tm, d_start, d_end, n, (const uint32_t *)nullptr,
                   c->d_defer, counts, c->d_err, c->d_stats);
          },
          [&](Timer &tm, int64_t nd) {
            launch(c, kslow, nd, st, tm, d_start, d_end, nd, (const uint32_t *)c->d_defer,
                   (uint32_t *)nullptr, counts, c->d_err, c->d_stats);
          });
    };
    float ms = 0.0f;
    if (!d_counts) {
      rc = c->periodic ? count(k_count<true, false>, k_count<true, true>)
                       : count(k_count<false, false>, k_count<false, true>);
      if (rc) return rc;
      ms = g_stats.kernel_ms;
    }
    const int64_t *cnt = d_counts ? d_counts : counts;
    const int64_t nseg = scan_counts(cnt, nrays, offsets, B, st);
    if (nseg > (int64_t)0xFFFFFFFFll)
      return fail(VT_EINVAL, "more than 2^32 - 1 segments in one adjoint call");
    // 2. segment lists (slot, ds), ray-major
    uint32_t *slot = B.get<uint32_t>(nseg);
    double *ds = B.get<double>(nseg);
    auto segs = [&](auto kfast, auto kslow) {
      return two_pass(
          c, st, nrays,
          [&](Timer &tm) {
            launch(c, kfast, n, st, tm, d_start, d_end, n, (const uint32_t *)nullptr,
                   c->d_defer, (const int64_t *)offsets, slot, ds, c->d_err, c->d_stats);
          },
          [&](Timer &tm, int64_t nd) {
            launch(c, kslow, nd, st, tm, d_start, d_end, nd, (const uint32_t *)c->d_defer,
                   (uint32_t *)nullptr, (const int64_t *)offsets, slot, ds, c->d_err,
                   c->d_stats);
          });
    };
    rc = c->periodic ? segs(k_segs<true, false>, k_segs<true, true>)
                     : segs(k_segs<false, false>, k_segs<false, true>);
    if (rc) return rc;
    ms += g_stats.kernel_ms;
    const int64_t segments = g_stats.segments;
    if (nseg == 0) return VT_OK;
    const float ms2 = adjoint_tail(c, mode, n, offsets, slot, ds, nseg, d_grad_out,
                                   d_grad_fields, B, st);
    g_stats.kernel_ms = ms + ms2;
    g_stats.segments = segments;
    return VT_OK;
  });
}

int vt_adjoint_recorded(const vt_cloud *cc, size_t nrays, int mode, const int64_t *d_counts,
                        const void *d_records, size_t n_records, const double *d_grad_out,
                        double *d_grad_fields, void *stream) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    if (mode != VT_SUM && mode != VT_VOLUME)
      return fail(VT_EINVAL, "the adjoint is defined for reductions 'sum' and 'volume'");
    if (!d_counts || (!d_records && n_records))
      return fail(VT_EINVAL, "null argument");
    if (n_records == (size_t)-1)
      return fail(VT_EINVAL, "the recorded projection overflowed its record pool");
    if (reinterpret_cast<uintptr_t>(d_records) % alignof(SegRec))
      return fail(VT_EINVAL, "the record pool must be 16-byte aligned");
    vt_cloud *c = const_cast<vt_cloud *>(cc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_OK(hipMemsetAsync(d_grad_fields, 0, c->n * c->nf * sizeof(double), st));
    g_stats = vt_call_stats{};
    if (nrays == 0) return VT_OK;
    const int64_t n = (int64_t)nrays;
    DevBufs B(c);
    int64_t *offsets = B.get<int64_t>(nrays + 1);
    const int64_t nseg = scan_counts(d_counts, nrays, offsets, B, st);
    if (nseg > (int64_t)0xFFFFFFFFll)
      return fail(VT_EINVAL, "more than 2^32 - 1 segments in one adjoint call");
    if ((size_t)nseg > n_records)
      return fail(VT_EINVAL, "the records do not hold every segment the counts name");
    if (nseg == 0) return VT_OK;
    uint32_t *slot = B.get<uint32_t>(nseg);
    double *ds = B.get<double>(nseg);
    Timer tm;
    HIP_OK(hipEventRecord(tm.e0, st));
    const int threads = 256;
    k_scatter_records<<<(int)((n_records + threads - 1) / threads), threads, 0, st>>>(
        (uint64_t)n_records, static_cast<const SegRec *>(d_records), n, offsets, d_counts, slot,
        ds);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(tm.e1, st));
    const float ms2 = adjoint_tail(c, mode, n, offsets, slot, ds, nseg, d_grad_out,
                                   d_grad_fields, B, st);
    float ms1 = 0;
    HIP_OK(hipEventElapsedTime(&ms1, tm.e0, tm.e1));
    g_stats.kernel_ms = ms1 + ms2;
    g_stats.segments = nseg;
    return VT_OK;
  });
}

int vt_grid_rays(const vt_view *views, size_t nviews, double *d_start, double *d_end,
                 void *stream) {
  return guarded([&]() -> int {
    if (nviews && !views) return fail(VT_EINVAL, "null argument");
    std::vector<int64_t> off(nviews + 1, 0);
    for (size_t v = 0; v < nviews; ++v) {
      const vt_view &w = views[v];
      if (w.nx < 1 || w.ny < 1) return fail(VT_EINVAL, "every view needs nx, ny >= 1");
      if (w.nx > (int64_t)1 << 31 || w.ny > (int64_t)1 << 31 ||
          w.nx * w.ny > ((int64_t)1 << 40))
        return fail(VT_EINVAL, "view too large");
      off[v + 1] = off[v] + w.nx * w.ny;
    }
    const int64_t n = off[nviews];
    if (n == 0) return VT_OK;
    if (!d_start || !d_end) return fail(VT_EINVAL, "null argument");
    if (nviews > (size_t)INT32_MAX) return fail(VT_EINVAL, "too many views");
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevBufs B;
    vt_view *dv = B.get<vt_view>(nviews);
    int64_t *doff = B.get<int64_t>(nviews + 1);
    HIP_OK(hipMemcpyAsync(dv, views, nviews * sizeof(vt_view), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(doff, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                          st));
    const int threads = 256;
    k_grid_rays<<<(unsigned)((n + threads - 1) / threads), threads, 0, st>>>(
        dv, doff, (int)nviews, n, d_start, d_end);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));  // the view table is freed on return
    return VT_OK;
  });
}

int vt_set_tuning(const vt_tuning *t) {
  if (!t) return fail(VT_EINVAL, "null argument");
  if (t->walk_batch < 1 || t->walk_batch > 64 || t->leaf_batch < 1 || t->leaf_batch > 64 ||
      (t->ray_parts != 1 && t->ray_parts != kParts) || t->help_events < 0 ||
      t->help_events > kHelpEvents)
    return fail(VT_EINVAL, fmt("tuning out of range: walk_batch and leaf_batch in [1, 64], "
                               "ray_parts 1 or %d, help_events in [0, %d]",
                               kParts, kHelpEvents));
  g_walk_batch.store(t->walk_batch);
  g_leaf_batch.store(t->leaf_batch);
  g_ray_parts.store(t->ray_parts);
  g_help_events.store(t->help_events);
  return VT_OK;
}

int vt_get_tuning(vt_tuning *t) {
  if (!t) return fail(VT_EINVAL, "null argument");
  t->walk_batch = walk_batch();
  t->leaf_batch = leaf_batch();
  t->ray_parts = ray_parts();
  t->help_events = help_events();
  return VT_OK;
}

int vt_cloud_lanes(const vt_cloud *cc, int mode, int recorded, int64_t *lanes) {
  return guarded([&]() -> int {
    int rc = check_ready(cc);
    if (!rc) rc = check_mode(cc, mode);
    if (rc) return rc;
    if (!lanes) return fail(VT_EINVAL, "null argument");
    if (recorded && mode != VT_SUM && mode != VT_VOLUME)
      return fail(VT_EINVAL, "recorded projections are for reductions 'sum' and 'volume'");
    const size_t nf = cc->nf;
    if (recorded) {
      *lanes = mode == VT_VOLUME ? project_lanes<VT_VOLUME, 4, true>(cc)
               : nf == 1         ? project_lanes<VT_SUM, 1, true>(cc)
                                 : project_lanes<VT_SUM, 0, true>(cc);
    } else if (mode == VT_SUM) {
      *lanes = nf == 1   ? project_lanes<VT_SUM, 1, false>(cc)
               : nf == 2 ? project_lanes<VT_SUM, 2, false>(cc)
               : nf == 3 ? project_lanes<VT_SUM, 3, false>(cc)
               : nf == 4 ? project_lanes<VT_SUM, 4, false>(cc)
                         : project_lanes<VT_SUM, 0, false>(cc);
    } else if (mode == VT_MAX) {
      *lanes = project_lanes<VT_MAX, 0, false>(cc);
    } else if (mode == VT_MIN) {
      *lanes = project_lanes<VT_MIN, 0, false>(cc);
    } else {
      *lanes = project_lanes<VT_VOLUME, 4, false>(cc);
    }
    return VT_OK;
  });
}

int vt_last_call_stats(vt_call_stats *out) {
  if (!out) return fail(VT_EINVAL, "null argument");
  *out = g_stats;
  return VT_OK;
}

}  // extern "C"
