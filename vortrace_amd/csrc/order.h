// Ray orders for the ray engine: longest rays first, decided on the device.
//
// The ray kernels' time has a tail: the last rays handed out finish long
// after the rest when they are long ones.  These kernels build an order in
// which each block of fetch slots of a ray partition takes its runs of 64
// slots (what one wave takes at once) longest first:
//   * fetch slot j names ray tile_ray(j) (the 8x8-tile order of row-major
//     images, or j itself);
//   * the slots of partition p are [p * psize, (p + 1) * psize), cut into
//     blocks of `block` slots and each block into runs of 64;
//   * a run's key is the largest segment count of its rays from a previous
//     pass (when the caller has counts) or an estimate: the length of the
//     run's middle ray (in box diagonals) times the sum of
//     (points per cell)^(1/3) of a G^3 histogram of the generators at S
//     points along it (the cells a ray crosses grow with the cube root of
//     the point density);
//   * within a block the runs are laid out by descending key, lower run
//     first among equal keys.
// Only the order in which lanes take rays changes, never a result.
#pragma once
#include "engine.h"

namespace vt {

constexpr int kOrderGrid = 32;     // histogram cells per axis
constexpr int kOrderSamples = 32;  // samples along a run's middle ray
constexpr int kOrderMaxRuns = 512;  // runs per block (block <= 32768)

// The partition / block / run structure of n fetch slots.
struct RunPlan {
  int64_t n, psize, block;
  int64_t per_part;   // blocks per partition
  int64_t per_block;  // runs per block
  int parts;
  uint32_t row, nfull;  // tile order (row = 0: fetch slot = ray)
  // Slots [first, first + len) of run r (len 0 for runs past the slots).
  __device__ __forceinline__ void run_slots(int64_t r, int64_t &first, int64_t &len) const {
    const int64_t g = r / per_block, k = r - g * per_block;
    const int64_t p = g / per_part, b = g - p * per_part;
    const int64_t lo = p * psize + b * block + 64 * k;
    const int64_t hi = min(min(p * psize + (b + 1) * block, (p + 1) * psize), n);
    const int64_t h = min(lo + 64, hi);
    first = lo;
    len = h > lo ? h - lo : 0;
  }
  __device__ __forceinline__ uint32_t ray_of(int64_t slot) const {
    return row ? tile_ray((uint32_t)slot, row, nfull) : (uint32_t)slot;
  }
};

// Generator counts per histogram cell: (x - lo) / size * G truncated and
// clipped to [0, G - 1] per axis, in FP64.
__global__ void k_order_hist(const Gen *gen, uint32_t n, double lo0, double lo1, double lo2,
                             double sz0, double sz1, double sz2, uint32_t *counts) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double2 xy = gld(reinterpret_cast<const double2 *>(gen + s));
  const double z = gld(&gen[s].z);
  const double v[3] = {(xy.x - lo0) / sz0 * kOrderGrid, (xy.y - lo1) / sz1 * kOrderGrid,
                       (z - lo2) / sz2 * kOrderGrid};
  int64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int64_t i = (int64_t)v[d];
    c[d] = i < 0 ? 0 : (i > kOrderGrid - 1 ? kOrderGrid - 1 : i);
  }
  atomicAdd(counts + (c[0] * kOrderGrid + c[1]) * kOrderGrid + c[2], 1u);
}

// (points per cell)^(1/3), float.
__global__ void k_order_cbrt(const uint32_t *counts, float *grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kOrderGrid * kOrderGrid * kOrderGrid) gst(grid + i, cbrtf((float)gld(counts + i)));
}

// Key of every run: the largest count of its rays, or the estimate (x 1024,
// truncated) from the run's middle ray.
__global__ void k_order_keys(RunPlan P, int64_t nruns, const int64_t *counts,
                             const double *start, const double *end, const float *grid,
                             float lo0, float lo1, float lo2, float sz0, float sz1, float sz2,
                             float inv_diag, int periodic, int64_t *keys) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  int64_t first, len;
  P.run_slots(r, first, len);
  int64_t key = 0;
  if (len > 0 && counts) {
    for (int64_t j = 0; j < len; ++j) {
      const int64_t c = gld(counts + P.ray_of(first + j));
      key = c > key ? c : key;
    }
  } else if (len > 0) {
    const uint32_t ray = P.ray_of(first + len / 2);
    const float s[3] = {(float)gld(start + 3 * (size_t)ray), (float)gld(start + 3 * (size_t)ray + 1),
                        (float)gld(start + 3 * (size_t)ray + 2)};
    const float d[3] = {(float)gld(end + 3 * (size_t)ray) - s[0],
                        (float)gld(end + 3 * (size_t)ray + 1) - s[1],
                        (float)gld(end + 3 * (size_t)ray + 2) - s[2]};
    const float lo[3] = {lo0, lo1, lo2}, sz[3] = {sz0, sz1, sz2};
    float sum = 0.0f;
    for (int i = 0; i < kOrderSamples; ++i) {
      const float t = ((float)i + 0.5f) / (float)kOrderSamples;
      int c[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float p = s[k] + t * d[k];
        int v = (int)floorf((p - lo[k]) / sz[k] * (float)kOrderGrid);
        if (periodic) {
          v %= kOrderGrid;
          v = v < 0 ? v + kOrderGrid : v;
        } else {
          v = v < 0 ? 0 : (v > kOrderGrid - 1 ? kOrderGrid - 1 : v);
        }
        c[k] = v;
      }
      sum += gld(grid + (c[0] * kOrderGrid + c[1]) * kOrderGrid + c[2]);
    }
    const float len_diag = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * inv_diag;
    key = (int64_t)(sum * len_diag * 1024.0f);
  }
  gst(keys + r, key);
}

// One workgroup per block: ranks its runs by (key descending, run index),
// lays their lengths out in that order, and writes every run's rays at its
// new place.  blockDim.x = 256, at most kOrderMaxRuns runs per block.
__global__ __launch_bounds__(256) void k_order_write(RunPlan P, const int64_t *keys,
                                                     uint32_t *order) {
  __shared__ int64_t key[kOrderMaxRuns];
  __shared__ uint32_t len_at[kOrderMaxRuns];
  __shared__ uint32_t start_at[kOrderMaxRuns + 1];
  const int64_t g = blockIdx.x;
  const int nr = (int)P.per_block;
  const int64_t r0 = g * P.per_block;
  for (int k = threadIdx.x; k < nr; k += blockDim.x) key[k] = gld(keys + r0 + k);
  __syncthreads();
  int rank[kOrderMaxRuns / 256];
  for (int q = 0; q < kOrderMaxRuns / 256; ++q) {
    const int k = threadIdx.x + 256 * q;
    if (k >= nr) continue;
    const int64_t mine = key[k];
    int rk = 0;
    for (int j = 0; j < nr; ++j) {
      const int64_t o = key[j];
      rk += (o > mine || (o == mine && j < k)) ? 1 : 0;
    }
    rank[q] = rk;
    int64_t first, len;
    P.run_slots(r0 + k, first, len);
    len_at[rk] = (uint32_t)len;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int j = 0; j < nr; ++j) {
      start_at[j] = acc;
      acc += len_at[j];
    }
    start_at[nr] = acc;
  }
  __syncthreads();
  int64_t gfirst, glen;
  P.run_slots(r0, gfirst, glen);  // the block's first slot
  for (int q = 0; q < kOrderMaxRuns / 256; ++q) {
    const int k = threadIdx.x + 256 * q;
    if (k >= nr) continue;
    int64_t first, len;
    P.run_slots(r0 + k, first, len);
    const int64_t out = gfirst + start_at[rank[q]];
    for (int64_t j = 0; j < len; ++j) gst(order + out + j, P.ray_of(first + j));
  }
}

}  // namespace vt
