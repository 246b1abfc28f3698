// Device k-nearest-neighbour search over the flattened kd-tree.
//
// Reproduces nanoflann 1.3.2's KDTreeSingleIndexAdaptor::knnSearch exactly
// (include/nanoflann.hpp:1221-1258 findNeighbors/knnSearch, 1005-1021
// computeInitialDistances, 1346-1408 searchLevel, 175-202 addPoint, 432-440
// evalMetric): the same best-child-first visiting order, the same incremental
// box distance `mindistsq + cut_dist - dists[idx]` and the same pruning and
// insertion rules, so ids, r^2 bits and the order among equal r^2 match.
//
// The recursion becomes a per-lane explicit stack whose first kLdsDepth
// entries live in LDS (the rest spill to device memory).  Two exact
// transformations keep it short:
//  * An entry for a far child is pushed only if its box distance is <= the
//    current worst distance: the worst distance never grows, so any entry
//    that fails now would fail when nanoflann tests it.
//  * On the first root-to-leaf descent the worst distance is still infinite.
//    Its far children go into an LDS ring holding the deepest kLdsDepth
//    levels; once the first leaf is scanned the ring is compacted to the
//    entries that can still be visited.  Only if an entry that fell out of
//    the ring could still be visited is the path walked again.
// Entries are either PENDING(parent node, box distance of its far child) or
// UNDO(dim, saved dists[dim]), which restores the per-dimension distance
// vector exactly where nanoflann's recursion would on return.
#pragma once
#include "device_common.h"

namespace vt {

constexpr uint32_t kSkip = 0xFFFFFFFFu;

// Launch shape shared by every kernel: lanes per workgroup and kNN stack
// entries per lane kept in LDS.
constexpr int kBlock = 128;
constexpr int kLdsDepth = 16;

// The LDS part of every lane's kNN stack: entry i of lane t at [i * kBlock + t],
// so lanes at equal depth touch consecutive banks.
__shared__ uint32_t lds_stack_tag[kLdsDepth * kBlock];
__shared__ double lds_stack_val[kLdsDepth * kBlock];

// Per-lane explicit stack.  Logical entries [0, kLdsDepth) live in LDS at slot
// (base + i) mod kLdsDepth (base is non-zero only after the first descent
// below, which fills the LDS part as a ring); deeper entries spill to device
// memory, interleaved by lane so lanes at equal depth coalesce.
static_assert((kLdsDepth & (kLdsDepth - 1)) == 0, "kLdsDepth must be a power of two");

struct KStack {
  int tid;       // lane within the workgroup
  uint32_t *gt;  // this lane's spill tags: entry i >= kLdsDepth at gt[(i - kLdsDepth) * lanes]
  double *gv;
  int lanes, gcap;
  int sp, maxsp;
  bool overflow;
  int base;

  __device__ __forceinline__ int slot(int i) const {
    return ((base + i) & (kLdsDepth - 1)) * kBlock + tid;
  }
  __device__ __forceinline__ void push(uint32_t tag, double v) {
    if (sp < kLdsDepth) {
      lds_stack_tag[slot(sp)] = tag;
      lds_stack_val[slot(sp)] = v;
    } else if (sp - kLdsDepth < gcap) {
      gst(gt + (size_t)(sp - kLdsDepth) * lanes, tag);
      gst(gv + (size_t)(sp - kLdsDepth) * lanes, v);
    } else {
      overflow = true;
    }
    ++sp;
    maxsp = sp > maxsp ? sp : maxsp;
  }
  __device__ __forceinline__ void pop(uint32_t &tag, double &v) {
    --sp;
    if (sp < kLdsDepth) {
      tag = lds_stack_tag[slot(sp)];
      v = lds_stack_val[slot(sp)];
    } else if (sp - kLdsDepth < gcap) {
      tag = gld(gt + (size_t)(sp - kLdsDepth) * lanes);
      v = gld(gv + (size_t)(sp - kLdsDepth) * lanes);
    } else {
      tag = kSkip;
      v = 0.0;
    }
  }
};

// k <= 3 result set held in registers (KNNResultSet semantics).
template <int K>
struct RegRes {
  double d[K];
  uint32_t id[K];
  int count;

  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      d[j] = DBL_MAX;
      id[j] = kNone;
    }
    count = 0;
  }
  __device__ __forceinline__ double worst() const { return d[K - 1]; }
  // addPoint: entries strictly greater than dist shift down one place and
  // dist lands after any equal entries.
  __device__ __forceinline__ void add(double dist, uint32_t idx) {
    int pos = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) pos += (j < count && d[j] <= dist) ? 1 : 0;
    if (pos < K) {
#pragma unroll
      for (int j = K - 1; j > 0; --j)
        if (j > pos) {
          d[j] = d[j - 1];
          id[j] = id[j - 1];
        }
#pragma unroll
      for (int j = 0; j < K; ++j)
        if (j == pos) {
          d[j] = dist;
          id[j] = idx;
        }
    }
    if (count < K) ++count;
  }
};

// Result set for a runtime k in {1, 2, 3}, in registers.  Entries at index
// >= k are never read (worst() is entry k - 1 and add() only compares the
// first `count` entries), so shifting into them is harmless.
struct RegRes3 {
  double d[3];
  uint32_t id[3];
  int count, k;

  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      d[j] = DBL_MAX;
      id[j] = kNone;
    }
    count = 0;
  }
  __device__ __forceinline__ double worst() const {
    return k == 1 ? d[0] : (k == 2 ? d[1] : d[2]);
  }
  __device__ __forceinline__ void add(double dist, uint32_t idx) {
    int pos = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) pos += (j < count && d[j] <= dist) ? 1 : 0;
    if (pos < k) {
#pragma unroll
      for (int j = 2; j > 0; --j)
        if (j > pos) {
          d[j] = d[j - 1];
          id[j] = id[j - 1];
        }
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (j == pos) {
          d[j] = dist;
          id[j] = idx;
        }
    }
    if (count < k) ++count;
  }
};

// Runtime-k result set in device memory (k up to kMaxQueryK); used for the
// K_EXPAND classification and for the query hook with k > 3.
struct MemRes {
  double *d;
  uint32_t *id;
  int k, count;

  __device__ void init() {
    for (int j = 0; j < k; ++j) {
      gst(d + j, DBL_MAX);
      gst(id + j, kNone);
    }
    count = 0;
  }
  __device__ __forceinline__ double worst() const { return gld(d + (k - 1)); }
  __device__ void add(double dist, uint32_t idx) {
    int i;
    for (i = count; i > 0; --i) {
      const double prev = gld(d + (i - 1));
      if (prev > dist) {
        if (i < k) {
          gst(d + i, prev);
          gst(id + i, gld(id + (i - 1)));
        }
      } else {
        break;
      }
    }
    if (i < k) {
      gst(d + i, dist);
      gst(id + i, idx);
    }
    if (count < k) ++count;
  }
};

// Leaf scan (searchLevel's leaf branch, include/nanoflann.hpp:1349-1366): the
// worst distance is read once, points are offered in slot order.
template <class RS>
__device__ __forceinline__ void scan_leaf(const DevCloud &C, const double q0, const double q1,
                                          const double q2, const Node &nd, RS &rs) {
  const uint32_t end = nd.b & ~kLeafBit;
  const double worst = rs.worst();
  for (uint32_t t = nd.a; t < end; ++t) {
    const double2 *g = reinterpret_cast<const double2 *>(C.gen + t);
    const double2 xy = gld(g);
    const double2 zw = gld(g + 1);
    const double e0 = q0 - xy.x;
    const double e1 = q1 - xy.y;
    const double e2 = q2 - zw.x;
    double dist = e0 * e0;
    dist = dist + e1 * e1;
    dist = dist + e2 * e2;
    if (dist < worst) rs.add(dist, t);
  }
}

// Child choice at an inner node: returns true when child1 (the lower one,
// index node + 1) is visited first, and the squared distance to the far
// child's near bound.
__device__ __forceinline__ bool choose(const Node &nd, double val, double &cut) {
  const double diff1 = val - nd.lo;
  const double diff2 = val - nd.hi;
  const bool low_first = (diff1 + diff2) < 0;
  const double c = low_first ? (val - nd.hi) : (val - nd.lo);
  cut = c * c;
  return low_first;
}

__device__ __forceinline__ double pick(const double q0, const double q1, const double q2,
                                       uint32_t f) {
  return f == 0 ? q0 : (f == 1 ? q1 : q2);
}

template <class RS, class ST>
__device__ void knn_core(const DevCloud &C, const double q0, const double q1, const double q2,
                         RS &rs, ST &st) {
  rs.init();
  // computeInitialDistances against root_bbox
  double dd[3] = {0.0, 0.0, 0.0};
  double mind = 0.0;
  const double qq[3] = {q0, q1, q2};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (qq[i] < C.rb_lo[i]) {
      dd[i] = (qq[i] - C.rb_lo[i]) * (qq[i] - C.rb_lo[i]);
      mind += dd[i];
    }
    if (qq[i] > C.rb_hi[i]) {
      dd[i] = (qq[i] - C.rb_hi[i]) * (qq[i] - C.rb_hi[i]);
      mind += dd[i];
    }
  }
  // The stack is empty between queries.
  st.sp = 0;
  st.base = 0;

  // Phase 1: best-child descent to the first leaf.  The far child of every
  // level (and its box distance) goes into the LDS ring at slot depth mod
  // kLdsDepth; the smallest box distance that falls out of the ring is kept.
  uint32_t node = 0;
  Node nd = load_node(C.nodes, 0);
  int depth = 0;
  double lost_min = DBL_MAX;
  while (!(nd.b & kLeafBit)) {
    double cut;
    const uint32_t f = nd.b;
    const bool low_first = choose(nd, pick(q0, q1, q2, f), cut);
    const double m2 = mind + cut - dd[f];
    const int sl = (depth & (kLdsDepth - 1)) * kBlock + st.tid;
    if (depth >= kLdsDepth) {
      const double old = lds_stack_val[sl];
      lost_min = old < lost_min ? old : lost_min;
    }
    lds_stack_tag[sl] = node;
    lds_stack_val[sl] = m2;
    ++depth;
    node = low_first ? node + 1 : nd.a;
    nd = load_node(C.nodes, node);
  }
  const uint32_t first_leaf = node;
  scan_leaf(C, q0, q1, q2, nd, rs);

  const double worst0 = rs.worst();
  if (depth <= kLdsDepth || !(lost_min <= worst0)) {
    // Every entry that left the ring would be pruned (the worst distance
    // never grows): keep the ring's entries that can still be visited, in
    // order, as the stack.  Compaction writes only to slots already read.
    const int n = depth < kLdsDepth ? depth : kLdsDepth;
    const int base = depth > kLdsDepth ? (depth & (kLdsDepth - 1)) : 0;
    st.base = base;
    int w = 0;
    for (int m = 0; m < n; ++m) {
      const int from = ((base + m) & (kLdsDepth - 1)) * kBlock + st.tid;
      const double v = lds_stack_val[from];
      if (v <= worst0) {
        const uint32_t t = lds_stack_tag[from];
        const int to = ((base + w) & (kLdsDepth - 1)) * kBlock + st.tid;
        lds_stack_tag[to] = t;
        lds_stack_val[to] = v;
        ++w;
      }
    }
    st.sp = w;
    st.maxsp = w > st.maxsp ? w : st.maxsp;
  } else {
    // An entry that fell out of the ring can still be visited: walk the
    // same path again and push every far child that can still be visited.
    node = 0;
    nd = load_node(C.nodes, 0);
    while (node != first_leaf) {
      double cut;
      const uint32_t f = nd.b;
      const bool low_first = choose(nd, pick(q0, q1, q2, f), cut);
      const double m2 = mind + cut - dd[f];
      if (m2 <= worst0) st.push(node, m2);
      node = low_first ? node + 1 : nd.a;
      nd = load_node(C.nodes, node);
    }
  }

  // Remaining traversal in nanoflann's order.
  while (st.sp > 0) {
    uint32_t tag;
    double v;
    st.pop(tag, v);
    if (tag == kSkip) continue;
    if (tag & kUndo) {
      dd[tag & 3u] = v;
      continue;
    }
    if (!(v <= rs.worst())) continue;  // mindistsq * epsError <= worstDist()
    const Node pn = load_node(C.nodes, tag);
    const uint32_t f = pn.b;
    double cut;
    const bool low_first = choose(pn, pick(q0, q1, q2, f), cut);
    node = low_first ? pn.a : tag + 1;  // the far child
    nd = load_node(C.nodes, node);
    if (nd.b & kLeafBit) {
      scan_leaf(C, q0, q1, q2, nd, rs);
      continue;
    }
    st.push(kUndo | f, dd[f]);
    dd[f] = cut;
    mind = v;
    const double worst = rs.worst();
    while (!(nd.b & kLeafBit)) {
      double c2;
      const uint32_t f2 = nd.b;
      const bool lf = choose(nd, pick(q0, q1, q2, f2), c2);
      const double m2 = mind + c2 - dd[f2];
      if (m2 <= worst) st.push(node, m2);
      node = lf ? node + 1 : nd.a;
      nd = load_node(C.nodes, node);
    }
    scan_leaf(C, q0, q1, q2, nd, rs);
  }
}

// PointCloud::minDistSqToBox, src/pointcloud.cpp:283-298.
__device__ __forceinline__ double min_dist_sq_to_box(const DevCloud &C, const double *q) {
  double d2 = 0.0;
#pragma unroll
  for (int dim = 0; dim < 3; ++dim) {
    const double lo = C.sb[2 * dim], hi = C.sb[2 * dim + 1];
    if (q[dim] < lo) {
      const double d = lo - q[dim];
      d2 += d * d;
    } else if (q[dim] > hi) {
      const double d = q[dim] - hi;
      d2 += d * d;
    }
  }
  return d2;
}

// PointCloud::minDistSqToBox (src/pointcloud.cpp:283-298) for the 27 shifted
// queries q + i*L, factored per dimension.  The reference adds, in dimension
// order, (lo - v)^2 or (v - hi)^2 for every coordinate v outside [lo, hi] to
// d2 = 0; adding 0.0 for an inside coordinate changes no bits, so
// d2(ix, iy, iz) = ((c[0][ix] + c[1][iy]) + c[2][iz]) is bit-identical, with
// c[d][s] computed once per query from the same shifted coordinate.
struct ImageBox {
  double lo[3], mid[3], hi[3];  // per dimension: shift -1, 0, +1
  __device__ __forceinline__ static double term(double v, double lo, double hi) {
    if (v < lo) {
      const double t = lo - v;
      return t * t;
    }
    if (v > hi) {
      const double t = v - hi;
      return t * t;
    }
    return 0.0;
  }
  __device__ __forceinline__ void init(const DevCloud &C, const double *q) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double blo = C.sb[2 * d], bhi = C.sb[2 * d + 1];
      lo[d] = term(q[d] + -1 * C.box[d], blo, bhi);
      mid[d] = term(q[d] + 0 * C.box[d], blo, bhi);
      hi[d] = term(q[d] + 1 * C.box[d], blo, bhi);
    }
  }
  __device__ __forceinline__ double pick(int d, int s) const {
    const double a = d == 0 ? lo[0] : (d == 1 ? lo[1] : lo[2]);
    const double b = d == 0 ? mid[0] : (d == 1 ? mid[1] : mid[2]);
    const double c = d == 0 ? hi[0] : (d == 1 ? hi[1] : hi[2]);
    return s < 0 ? a : (s > 0 ? c : b);
  }
  __device__ __forceinline__ double d2(int ix, int iy, int iz) const {
    return (pick(0, ix) + pick(1, iy)) + pick(2, iz);
  }
  // False when every image has a coordinate whose own contribution already
  // exceeds w; sums of non-negative terms round monotonically, so each such
  // image's d2 exceeds w too and the reference prunes all of them.
  __device__ __forceinline__ bool any_within(double w) const {
    bool any = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) any |= (lo[d] <= w) | (hi[d] <= w);
    return any;
  }
};

// PointCloud::queryTreeK for k = K <= 3 (src/pointcloud.cpp:308-368): the
// plain search, then in periodic mode up to 26 image searches merged with
// de-duplication by id (keeping the smaller r^2), a stable sort by r^2 and
// truncation to K after every image.
template <int K, bool kPeriodic, class ST>
__device__ void query_reg(const DevCloud &C, const double q0, const double q1, const double q2,
                          uint32_t (&ids)[K], double (&r2)[K], ST &st) {
  RegRes<K> rs;
  knn_core(C, q0, q1, q2, rs, st);
  if constexpr (!kPeriodic) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      ids[j] = rs.id[j];
      r2[j] = rs.d[j];
    }
    return;
  }
  uint32_t mi[2 * K];
  double md[2 * K];
  int m = K;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    mi[j] = rs.id[j];
    md[j] = rs.d[j];
  }
#pragma unroll
  for (int j = K; j < 2 * K; ++j) {
    mi[j] = kNone;
    md[j] = DBL_MAX;
  }
  const double qv[3] = {q0, q1, q2};
  ImageBox ib;
  ib.init(C, qv);
  if (!ib.any_within(md[K - 1] + kTol)) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      ids[j] = mi[j];
      r2[j] = md[j];
    }
    return;
  }
  for (int ix = -1; ix <= 1; ix++)
    for (int iy = -1; iy <= 1; iy++)
      for (int iz = -1; iz <= 1; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        if (ib.d2(ix, iy, iz) > md[K - 1] + kTol) continue;
        const double sh[3] = {qv[0] + ix * C.box[0], qv[1] + iy * C.box[1],
                              qv[2] + iz * C.box[2]};
        RegRes<K> ri;
        knn_core(C, sh[0], sh[1], sh[2], ri, st);
#pragma unroll
        for (int i = 0; i < K; ++i) {
          bool found = false;
#pragma unroll
          for (int j = 0; j < 2 * K; ++j) {
            if (!found && j < m && mi[j] == ri.id[i]) {
              md[j] = (ri.d[i] < md[j]) ? ri.d[i] : md[j];
              found = true;
            }
          }
          if (!found) {
#pragma unroll
            for (int j = 0; j < 2 * K; ++j)
              if (j == m) {
                mi[j] = ri.id[i];
                md[j] = ri.d[i];
              }
            ++m;
          }
        }
        // Stable insertion sort of the first m entries (std::sort on at
        // most 16 elements is libstdc++'s insertion sort).
#pragma unroll
        for (int i = 1; i < 2 * K; ++i) {
#pragma unroll
          for (int j = i; j > 0; --j) {
            if (i < m && md[j] < md[j - 1]) {
              const double td = md[j];
              md[j] = md[j - 1];
              md[j - 1] = td;
              const uint32_t ti = mi[j];
              mi[j] = mi[j - 1];
              mi[j - 1] = ti;
            }
          }
        }
        if (m > K) m = K;
      }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    ids[j] = mi[j];
    r2[j] = md[j];
  }
}

// queryTreeK for a runtime k in {1, 2, 3} (src/pointcloud.cpp:308-368): one
// code path for every query of the walk.  Same merge as query_reg.
template <bool kPeriodic, class ST>
__device__ __forceinline__ void query_rt(const DevCloud &C, const double q0, const double q1,
                                         const double q2, int k, uint32_t (&ids)[3],
                                         double (&r2)[3], ST &st) {
  RegRes3 rs;
  rs.k = k;
  knn_core(C, q0, q1, q2, rs, st);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    ids[j] = j < k ? rs.id[j] : kNone;
    r2[j] = j < k ? rs.d[j] : DBL_MAX;
  }
  if constexpr (!kPeriodic) return;
  const double qv[3] = {q0, q1, q2};
  {
    ImageBox ib;
    ib.init(C, qv);
    if (!ib.any_within(rs.worst() + kTol)) return;
  }
  uint32_t mi[6];
  double md[6];
  int m = k;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    mi[j] = rs.id[j];
    md[j] = rs.d[j];
  }
#pragma unroll
  for (int j = 3; j < 6; ++j) {
    mi[j] = kNone;
    md[j] = DBL_MAX;
  }
  for (int ix = -1; ix <= 1; ix++)
    for (int iy = -1; iy <= 1; iy++)
      for (int iz = -1; iz <= 1; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        const double sh[3] = {qv[0] + ix * C.box[0], qv[1] + iy * C.box[1],
                              qv[2] + iz * C.box[2]};
        const double wk = k == 1 ? md[0] : (k == 2 ? md[1] : md[2]);
        if (min_dist_sq_to_box(C, sh) > wk + kTol) continue;
        RegRes3 ri;
        ri.k = k;
        knn_core(C, sh[0], sh[1], sh[2], ri, st);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (i < k) {
            bool found = false;
#pragma unroll
            for (int j = 0; j < 6; ++j) {
              if (!found && j < m && mi[j] == ri.id[i]) {
                md[j] = (ri.d[i] < md[j]) ? ri.d[i] : md[j];
                found = true;
              }
            }
            if (!found) {
#pragma unroll
              for (int j = 0; j < 6; ++j)
                if (j == m) {
                  mi[j] = ri.id[i];
                  md[j] = ri.d[i];
                }
              ++m;
            }
          }
        }
#pragma unroll
        for (int i = 1; i < 6; ++i) {
#pragma unroll
          for (int j = i; j > 0; --j) {
            if (i < m && md[j] < md[j - 1]) {
              const double td = md[j];
              md[j] = md[j - 1];
              md[j - 1] = td;
              const uint32_t ti = mi[j];
              mi[j] = mi[j - 1];
              mi[j - 1] = ti;
            }
          }
        }
        if (m > k) m = k;
      }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    ids[j] = j < k ? mi[j] : kNone;
    r2[j] = j < k ? md[j] : DBL_MAX;
  }
}

// queryTreeK for a runtime k (<= kMaxQueryK) with results in `res`
// (k entries) and `merged` workspace (2k entries), both in device memory.
// Entries beyond the number of points are (kNone, DBL_MAX).
template <bool kPeriodic, class ST>
__device__ void query_mem(const DevCloud &C, const double q0, const double q1, const double q2,
                          int k, MemRes &res, uint32_t *mi, double *md, ST &st) {
  knn_core(C, q0, q1, q2, res, st);
  if (!kPeriodic) return;
  int m = k;
  for (int j = 0; j < k; ++j) {
    gst(mi + j, gld(res.id + j));
    gst(md + j, gld(res.d + j));
  }
  MemRes img{res.d + k, res.id + k, k, 0};  // image results after the k slots
  const double qv[3] = {q0, q1, q2};
  ImageBox ib;
  ib.init(C, qv);
  for (int ix = -1; ix <= 1; ix++)
    for (int iy = -1; iy <= 1; iy++)
      for (int iz = -1; iz <= 1; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        if (ib.d2(ix, iy, iz) > gld(md + (k - 1)) + kTol) continue;
        const double sh[3] = {qv[0] + ix * C.box[0], qv[1] + iy * C.box[1],
                              qv[2] + iz * C.box[2]};
        knn_core(C, sh[0], sh[1], sh[2], img, st);
        for (int i = 0; i < k; ++i) {
          const uint32_t iid = gld(img.id + i);
          const double idd = gld(img.d + i);
          bool found = false;
          for (int j = 0; j < m; ++j) {
            if (gld(mi + j) == iid) {
              const double cur = gld(md + j);
              gst(md + j, (idd < cur) ? idd : cur);
              found = true;
              break;
            }
          }
          if (!found) {
            gst(mi + m, iid);
            gst(md + m, idd);
            ++m;
          }
        }
        // Stable insertion sort.  libstdc++'s std::sort is an insertion sort
        // (stable) up to 16 elements; beyond that its order among equal r^2
        // may differ, which only the k > 8 query hook can observe.
        for (int i = 1; i < m; ++i) {
          const double vd = gld(md + i);
          const uint32_t vi = gld(mi + i);
          int j = i;
          while (j > 0 && vd < gld(md + (j - 1))) {
            gst(md + j, gld(md + (j - 1)));
            gst(mi + j, gld(mi + (j - 1)));
            --j;
          }
          gst(md + j, vd);
          gst(mi + j, vi);
        }
        if (m > k) m = k;
      }
  for (int j = 0; j < k; ++j) {
    gst(res.id + j, gld(mi + j));
    gst(res.d + j, gld(md + j));
  }
}

}  // namespace vt
