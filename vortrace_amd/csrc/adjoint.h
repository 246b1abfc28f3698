// Device side of the adjoint (d image / d cell fields) and of field updates.
//
// The reference has no derivatives; these kernels differentiate its forward
// reductions (src/reduction.cpp:8-21 reduce_sum, 55-83 reduce_volume_render)
// through the segment lists of Projection::makeDetailedProjection
// (src/projection.cpp:75-144), which the forward kernels reproduce exactly.
//
// For an image gradient g (rays x n_out) the field gradient is
//   sum:    dL/drho[c,f] = sum over segments (r, c, ds) of g[r,f] * ds
//   volume: per ray, with op_i = 1 - exp(-alpha_i ds_i), T_i = prod_{j<i}(1 - op_j)
//           and B_i the colour composited from segment i+1 onward,
//             dL/dcol_i = g * T_i * op_i
//             dL/dalpha_i = sum_ch g_ch * T_i * (col_i,ch - B_i,ch) * ds_i * (1 - op_i)
// Each segment's contribution is computed by its ray's lane; the per-cell sums
// then run over the segments sorted stably by cell, so every cell adds its
// contributions in ray-major segment order: the result does not depend on
// scheduling, and no floating-point atomics are used.
#pragma once
#include "engine.h"

namespace vt {

// Segments of a ray at its offset: slot of the cell and ds.
struct SegSink {
  uint32_t *slot;
  double *ds;
  uint64_t cap;
  uint64_t n;
  __device__ void init() { n = 0; }
  __device__ __forceinline__ void seg(uint32_t cell, double a, double b) {
    if (n < cap) {
      gst(slot + n, cell);
      gst(ds + n, b - a);
    }
    ++n;
  }
};

struct SegIO {
  using Sink = SegSink;
  const int64_t *offsets;
  uint32_t *slot;
  double *ds;
  __device__ void begin(int64_t ray, Sink &sk) const {
    const int64_t off = gld(offsets + ray);
    sk.slot = slot + off;
    sk.ds = ds + off;
    sk.cap = (uint64_t)(gld(offsets + ray + 1) - off);
    sk.init();
  }
  __device__ int end(int64_t, Sink &sk) const {
    return sk.n == sk.cap ? kErrNone : kErrFillOverflow;
  }
};

// One segment recorded by a forward pass: the cell's slot, the ray, its
// position k along the ray, and ds.  ray == kNone marks a record of a
// lane's last chunk that the lane did not fill.
struct alignas(16) SegRec {
  uint32_t slot, ray, k, pad;
  double ds, pad2;
};
static_assert(sizeof(SegRec) == 32, "segment record must be 32 bytes");

// Pool the records go to: lanes take chunks of kRecChunk records from `head`
// and fill them in order; `overflow` is set when the pool runs out.  A lane
// marks the records of its last chunk that it did not fill (ray = kNone)
// after its last ray, so every record below the final head is either written
// or marked and the pool needs no clearing beforehand.
struct RecPool {
  SegRec *rec;
  unsigned long long cap;
  unsigned long long *head;
  int *overflow;
};
constexpr int kRecChunk = 32;

// A projection IO that also records every segment.  The sink's chunk state
// survives from ray to ray (begin() leaves it alone).  A ray left to the
// K_EXPAND launch records a prefix of its segments twice, with the same
// values: the backward pass writes both to the same place.
template <class InnerIO>
struct RecordIO {
  // The inner sink with recording added; its segment count `n` is the
  // place of the next segment along the ray.
  struct Sink : InnerIO::Sink {
    using Base = typename InnerIO::Sink;
    RecPool pool;  // set once per lane before the first ray (setup), the same in every lane
    uint32_t ray;
    // Next record of this lane's current chunk; a multiple of kRecChunk
    // means the chunk is used up (chunks start at multiples of kRecChunk).
    unsigned long long next = 0;
    __device__ __forceinline__ void seg(uint32_t cell, double a, double b) {
      const uint32_t k = (uint32_t)this->n;  // the segment's place along the ray
      Base::seg(cell, a, b);
      if ((next & (unsigned long long)(kRecChunk - 1)) == 0ull)
        next = atomicAdd(pool.head, (unsigned long long)kRecChunk);
      if (next < pool.cap) {
        SegRec *r = pool.rec + next;
        // Records are read only by a later kernel.  Measured at 2M clustered
        // cells, 512^2: HBM reads of the recording forward 13.7 -> 8.0 GB.
        gst_nt(reinterpret_cast<uint4 *>(r), make_uint4(cell, ray, k, 0u));
        gst_nt(&r->ds, b - a);
      } else {
        atomicOr(pool.overflow, 1);
      }
      ++next;
    }
  };
  InnerIO inner;
  RecPool pool;
  __device__ void setup(Sink &sk) const { sk.pool = pool; }
  // After the lane's last ray: the rest of its current chunk is unwritten.
  __device__ void finish(Sink &sk) const {
    unsigned long long j = sk.next;
    while ((j & (unsigned long long)(kRecChunk - 1)) != 0ull && j < pool.cap) {
      gst(&pool.rec[j].ray, kNone);
      ++j;
    }
  }
  __device__ void begin(int64_t ray, Sink &sk) const {
    inner.begin(ray, static_cast<typename Sink::Base &>(sk));
    sk.ray = (uint32_t)ray;
  }
  __device__ int end(int64_t ray, Sink &sk) const {
    return inner.end(ray, static_cast<typename Sink::Base &>(sk));
  }
};

// Recorded segments -> the ray-major (slot, ds) lists at offsets[ray] + k.
// Unwritten records (ray == kNone) and records past a ray's count are
// skipped; every (ray, k < counts[ray]) is written by at least one record
// when the pool did not overflow.
__global__ void k_scatter_records(uint64_t nrec, const SegRec *rec, int64_t nrays,
                                  const int64_t *offsets, const int64_t *counts, uint32_t *slot,
                                  double *ds) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nrec) return;
  const uint4 h = gld(reinterpret_cast<const uint4 *>(rec + j));
  if (h.y == kNone || (int64_t)h.y >= nrays) return;
  if ((int64_t)h.z >= gld(counts + h.y)) return;
  const int64_t pos = gld(offsets + h.y) + h.z;
  gst(slot + pos, h.x);
  gst(ds + pos, gld(&rec[j].ds));
}

// Fields given in filtered point order -> slot order, and the RGBA validity
// rule of PointCloud::loadPoints (src/pointcloud.cpp:168-179): a value
// outside [0, 1] in R, G or B, or a negative alpha, makes the set invalid.
__global__ void k_set_fields(const double *src, const uint32_t *slot_id, uint32_t n, uint32_t nf,
                             double *dst, int *invalid_rgba) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = gld(slot_id + s);
  bool bad = false;
  for (uint32_t f = 0; f < nf; ++f) {
    const double v = gld(src + i * nf + f);
    gst(dst + (size_t)s * nf + f, v);
    if (nf == 4) bad |= f < 3 ? (v < 0.0 || v > 1.0) : (v < 0.0);
  }
  if (bad) atomicOr(invalid_rgba, 1);
}

// Sum: contribution of segment j of ray r to field f is g[r, f] * ds_j.
// One wave per ray, its lanes over the ray's segments (coalesced).
__global__ void k_contrib_sum(int64_t nrays, const int64_t *offsets, const double *ds,
                              const double *g, int nf, double *contrib) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = (int)(threadIdx.x & 63);
  if (r >= nrays) return;
  const int64_t a = gld(offsets + r), b = gld(offsets + r + 1);
  for (int f = 0; f < nf; ++f) {
    const double gf = gld(g + r * nf + f);
    for (int64_t j = a + lane; j < b; j += 64) gst(contrib + j * nf + f, gf * gld(ds + j));
  }
}

// Volume: reverse compositing of one ray.  The forward sweep repeats
// reduce_volume_render's arithmetic for op and T and parks T_i in the alpha
// slot; the backward sweep accumulates B and writes the four contributions.
__global__ void k_contrib_vol(int64_t nrays, const int64_t *offsets, const uint32_t *slot,
                              const double *ds, const double *fields, const double *g,
                              double *contrib) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const int64_t a = gld(offsets + r), b = gld(offsets + r + 1);
  double T = 1.0;
  for (int64_t j = a; j < b; ++j) {
    const double alpha = gld(fields + (size_t)gld(slot + j) * 4 + 3);
    const double opacity = 1.0 - exp(-alpha * gld(ds + j));
    gst(contrib + j * 4 + 3, T);
    T *= (1.0 - opacity);
  }
  const double g0 = gld(g + r * 3), g1 = gld(g + r * 3 + 1), g2 = gld(g + r * 3 + 2);
  double B0 = 0.0, B1 = 0.0, B2 = 0.0;
  for (int64_t j = b - 1; j >= a; --j) {
    const double *fv = fields + (size_t)gld(slot + j) * 4;
    const double c0 = gld(fv), c1 = gld(fv + 1), c2 = gld(fv + 2), alpha = gld(fv + 3);
    const double dsj = gld(ds + j);
    const double e = exp(-alpha * dsj);
    const double opacity = 1.0 - e;
    const double Tj = gld(contrib + j * 4 + 3);
    const double w = Tj * opacity;
    gst(contrib + j * 4 + 0, g0 * w);
    gst(contrib + j * 4 + 1, g1 * w);
    gst(contrib + j * 4 + 2, g2 * w);
    const double dalpha =
        (g0 * (c0 - B0) + g1 * (c1 - B1) + g2 * (c2 - B2)) * Tj * dsj * (1.0 - opacity);
    gst(contrib + j * 4 + 3, dalpha);
    B0 = c0 * opacity + (1.0 - opacity) * B0;
    B1 = c1 * opacity + (1.0 - opacity) * B1;
    B2 = c2 * opacity + (1.0 - opacity) * B2;
  }
}

// Run [start, end) of every slot in the slot-sorted segment order.
__global__ void k_run_bounds(int64_t nseg, const uint32_t *key, uint32_t *start, uint32_t *end) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nseg) return;
  const uint32_t k = gld(key + j);
  if (j == 0 || gld(key + j - 1) != k) gst(start + k, (uint32_t)j);
  if (j == nseg - 1 || gld(key + j + 1) != k) gst(end + k, (uint32_t)(j + 1));
}

// Per cell, the sum of its contributions in ray-major segment order (perm is
// the stable slot sort of the segment indices), written in filtered order.
__global__ void k_reduce_cells(uint32_t n, int nf, const uint32_t *start, const uint32_t *end,
                               const uint32_t *perm, const double *contrib,
                               const uint32_t *slot_id, double *grad) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t a = gld(start + s), b = gld(end + s);
  const size_t i = gld(slot_id + s);
  for (int f = 0; f < nf; ++f) {
    double acc = 0.0;
    for (uint32_t j = a; j < b; ++j) acc += gld(contrib + (size_t)gld(perm + j) * nf + f);
    gst(grad + i * nf + f, acc);
  }
}

// The same sums for one field when the contributions themselves were sorted
// by slot (stably, so each cell's run is in ray-major order): contiguous reads.
__global__ void k_reduce_sorted(uint32_t n, const uint32_t *start, const uint32_t *end,
                                const double *sorted, const uint32_t *slot_id, double *grad) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t a = gld(start + s), b = gld(end + s);
  double acc = 0.0;
  for (uint32_t j = a; j < b; ++j) acc += gld(sorted + j);
  gst(grad + gld(slot_id + s), acc);
}

// Marks a ray order that is not a permutation of [0, n): an entry out of
// range, or one that an earlier entry already took.
__global__ void k_check_order(int64_t n, const uint32_t *order, uint32_t *seen, int *bad) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t r = gld(order + j);
  if ((int64_t)r >= n || atomicAdd(seen + r, 1u) != 0u) atomicOr(bad, 1);
}

__global__ void k_iota(int64_t n, uint32_t *v) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) gst(v + j, (uint32_t)j);
}

}  // namespace vt
