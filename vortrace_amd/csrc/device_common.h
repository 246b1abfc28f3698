// Device-side definitions shared by the kNN and walk code (gfx950, FP64).
//
// All arithmetic that decides a cell id or an s value is FP64 with every
// multiply and add rounded separately (the translation unit is compiled with
// -ffp-contract=off and `#pragma clang fp contract(off)`), in the operand
// order of the reference, so results are bit-identical to a default x86-64
// build of it.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "layout.h"

#pragma clang fp contract(off)

namespace vt {

// src/ray.cpp:8-12 (TOLERANCE, PARALLEL_BISECTOR_TOL, MAX_ITERATIONS,
// PERTURBATION_EPS = 2 * TOLERANCE, MAX_PERTURBATION_CYCLES).
constexpr double kTol = 1E-9;
constexpr double kParTol = 1E-12;
constexpr int kMaxIter = 100000;
constexpr double kPertEps = 2 * kTol;
constexpr int kMaxCycles = 1000;
constexpr int kExpandK = 100;  // K_EXPAND, src/ray.cpp:321
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kUndo = kRefLeaf | kRefUndo;  // stack tag of an undo entry (| dim)

// Error codes raised inside kernels; the host maps them to the reference's
// exception texts.
enum : int {
  kErrNone = 0,
  kErrMaxIter = 1,         // src/ray.cpp:260-261
  kErrSplitOutside = 2,    // src/ray.cpp:282-289
  kErrClassify = 3,        // src/ray.cpp:331-336
  kErrBisect = 4,          // src/ray.cpp:125-126
  kErrPerturb = 5,         // src/ray.cpp:182-184
  kErrPeriodicStart = 6,   // src/ray.cpp:243-246
  kErrWalkStack = 7,       // walk stack deeper than the allocated capacity
  kErrKnnStack = 8,        // kNN stack deeper than the tree (cannot happen)
  kErrTies = 9,            // >= K_EXPAND equidistant generators at one point,
                           // periodic images merged
  kErrNonFinite = 10,      // a query found no generator (non-finite point)
  kErrFillOverflow = 11,   // fill pass produced more segments than counted
  kErrHelpWait = 12,       // a lane waited too long for another lane of its wave
  // Not an error: the ray needs the K_EXPAND classification, which the fast
  // kernels leave to a second launch over the deferred rays.
  kDeferRay = 64,
};

struct ErrRec {
  int flag;  // 0 until the first error claims the record
  int code;
  int64_t ray;
  double v[3];
  uint32_t id[2];
  int pad[2];
};

struct Stats {
  unsigned long long segments;
  unsigned long long queries;
  unsigned int max_walk_depth;
  unsigned int max_knn_depth;
  unsigned long long slow_queries;
  unsigned long long ray_counter;  // work queue head (query kernels)
  unsigned long long deferred;     // rays appended to the deferred list
  unsigned long long part_counter[8];  // ray engine: work queue head per partition
  unsigned long long intervals;        // ray engine: intervals given to other lanes
  unsigned long long intervals_back;   // ... of which came back unwalked
};

// The ray engine splits its rays into kParts contiguous partitions, one per
// XCD: workgroups are dispatched to the XCDs round-robin, so workgroup b runs
// on XCD b mod 8 and its lanes take rays from partition b mod 8 first.  Rays
// of one partition cover one band of the image, which keeps each XCD's L2
// working set to the cells under that band.
constexpr int kParts = 8;

// Everything a kernel needs to know about a built cloud (passed by value).
struct DevCloud {
  const DNode *nodes;    // inner nodes, pre-order
  const DRec *recs;      // pair records (ray engine searches)
  const Gen *gen;        // slot order, followed by kLeafMaxSize padding records
  const double *fields;  // slot order, nf per slot
  double rb_lo[3], rb_hi[3];  // nanoflann root_bbox
  double sb[6];               // PointCloud::subbox
  double box[3];              // periodic box size
  double inv_sqrt2;           // 1.0 / std::sqrt(2.0), computed on the host
  uint32_t n, nf;
  int periodic;
  uint32_t root;   // reference of the root (a leaf when n <= kLeafMaxSize)
  uint32_t rroot;  // the root among pair records
};

// Per-lane scratch that does not fit in registers/LDS (device memory).
struct Scratch {
  uint4 *knn_spill;    // [kspill][T] kNN stack entries beyond the LDS part
  uint32_t *walk_tag;  // [wcap][T] pending ray points below the top
  double *walk_s;
  uint32_t *big_id;    // [T][3 * kMaxQueryK] k > 3 result / merge buffers
  double *big_d;
  int kspill, wcap;
  int lanes;           // T
  int walk_batch;      // ray engine: lanes of a wave that run their walk steps together
  int leaf_batch;      // ray engine: lanes of a wave that run their leaf steps together
  int parts;           // ray engine: ray partitions (1 or kParts)
  int fetch_lanes;     // ray engine: lanes of each wave that take rays (1..64)
  uint32_t tile_row;   // ray engine: rays per image row for the 8x8-tile ray order (0 = off)
  const uint32_t *ray_order;  // ray engine: the order fetch slots take rays in (null = own order)
  // ray engine, intervals walked for another lane of the wave (null: off):
  uint4 *help_ev;      // [help_cap][T] events of the interval this lane walks
  uint4 *help_steal;   // [2 * help_steal_cap][T] intervals this lane gave away
  int help_cap, help_steal_cap;
  int help_depth;      // pending entries below the top a lane needs to give one pair away
};

constexpr int kMaxQueryK = 128;

// Loads and stores through the global address space.  The data pointers
// travel inside structs, where the compiler cannot see that they are global
// and would otherwise emit flat accesses.
template <class T>
__device__ __forceinline__ T gld(const T *p) {
  return *(const __attribute__((address_space(1))) T *)(p);
}
template <class T>
__device__ __forceinline__ void gst(T *p, T v) {
  *(__attribute__((address_space(1))) T *)(p) = v;
}
typedef double v2f64 __attribute__((ext_vector_type(2)));
// 16-byte global load of two doubles (one dwordx4).
__device__ __forceinline__ double2 gld(const double2 *p) {
  const v2f64 v = *(const __attribute__((address_space(1))) v2f64 *)(p);
  return make_double2(v.x, v.y);
}

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
// 16-byte global load / store of four dwords.
__device__ __forceinline__ uint4 gld(const uint4 *p) {
  const v4u32 v = *(const __attribute__((address_space(1))) v4u32 *)(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(uint4 *p, uint4 e) {
  v4u32 v;
  v.x = e.x;
  v.y = e.y;
  v.z = e.z;
  v.w = e.w;
  *(__attribute__((address_space(1))) v4u32 *)(p) = v;
}

// Non-temporal stores, for data the kernel writes and never reads back: a
// partly written line is not fetched into the L2 first.
__device__ __forceinline__ void gst_nt(uint4 *p, uint4 e) {
  v4u32 v;
  v.x = e.x;
  v.y = e.y;
  v.z = e.z;
  v.w = e.w;
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) v4u32 *)(p));
}
__device__ __forceinline__ void gst_nt(double *p, double v) {
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) double *)(p));
}

// a, b or c for f = 0, 1 or 2 (by value).
__device__ __forceinline__ double sel3(double a, double b, double c, uint32_t f) {
  return f == 0 ? a : (f == 1 ? b : c);
}

// A device node, decoded.
struct DNodeV {
  double lo, hi, pb;
  uint32_t c1, c2;  // child references (dimension bits removed)
  uint32_t dim;
};

__device__ __forceinline__ DNodeV decode_node(const double2 v0, const double2 v1) {
  DNodeV n;
  n.lo = v0.x;
  n.hi = v0.y;
  const uint64_t cc = (uint64_t)__double_as_longlong(v1.x);
  const uint32_t c1 = (uint32_t)cc, c2 = (uint32_t)(cc >> 32);
  n.dim = ((c1 >> 30) & 1u) | ((c2 >> 29) & 2u);
  n.c1 = c1 & ~kDimBit;
  n.c2 = c2 & ~kDimBit;
  n.pb = v1.y;
  return n;
}

__device__ __forceinline__ DNodeV load_node(const DNode *nodes, uint32_t i) {
  const double2 *p = reinterpret_cast<const double2 *>(nodes + i);
  return decode_node(gld(p), gld(p + 1));
}

// A pair record, read at once, with the node's own fields decoded.
struct DRecV {
  DNodeV x;                    // the record's node
  double lo1, hi1, lo2, hi2;   // its children's bounds
  uint32_t c11, c12, c21, c22;  // its children's references, with dim bits
};

__device__ __forceinline__ DRecV load_rec(const DRec *recs, uint32_t r) {
  const double2 *p = reinterpret_cast<const double2 *>(recs + r);
  const double2 v0 = gld(p), v1 = gld(p + 1), v2 = gld(p + 2), v3 = gld(p + 3);
  const uint4 v4 = gld(reinterpret_cast<const uint4 *>(p + 4));
  DRecV R;
  R.x = decode_node(v0, v1);
  R.lo1 = v2.x;
  R.hi1 = v2.y;
  R.lo2 = v3.x;
  R.hi2 = v3.y;
  R.c11 = v4.x;
  R.c12 = v4.y;
  R.c21 = v4.z;
  R.c22 = v4.w;
  return R;
}

// Child b (0: child1, 1: child2) of a record's node, as a node: its bounds
// and children from the record, and the parent's bound facing it as pb.
__device__ __forceinline__ DNodeV rec_child(const DRecV &R, uint32_t b) {
  DNodeV n;
  n.lo = b ? R.lo2 : R.lo1;
  n.hi = b ? R.hi2 : R.hi1;
  const uint32_t a1 = b ? R.c21 : R.c11, a2 = b ? R.c22 : R.c12;
  n.dim = ((a1 >> 30) & 1u) | ((a2 >> 29) & 2u);
  n.c1 = a1 & ~kDimBit;
  n.c2 = a2 & ~kDimBit;
  n.pb = b ? R.x.hi : R.x.lo;
  return n;
}

// Stack tag of a child reference taken as a far child at a node cutting
// along `dim`.
__device__ __forceinline__ uint32_t far_tag(uint32_t ref, uint32_t dim) {
  return (ref & kRefLeaf) ? ref : (ref | (dim << 29));
}

__device__ __forceinline__ void report_error(ErrRec *err, int code, int64_t ray, double a,
                                             double b, double c, uint32_t i0, uint32_t i1) {
  if (atomicCAS(&err->flag, 0, 1) == 0) {
    err->code = code;
    err->ray = ray;
    err->v[0] = a;
    err->v[1] = b;
    err->v[2] = c;
    err->id[0] = i0;
    err->id[1] = i1;
    __threadfence();
  }
}

}  // namespace vt
