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
//  * The first root-to-leaf descent pushes nothing; once that leaf is
//    scanned the same path is walked again (cache-resident) and only the far
//    children that survive the now-finite worst distance are pushed.
// Entries are either PENDING(parent node, box distance of its far child) or
// UNDO(dim, saved dists[dim]), which restores the per-dimension distance
// vector exactly where nanoflann's recursion would on return.
#pragma once
#include "device_common.h"

namespace vt {

constexpr uint32_t kSkip = 0xFFFFFFFFu;

// Per-lane explicit stack: LDS for the top kLdsDepth entries, device memory
// beyond (interleaved by lane so lanes at equal depth coalesce).
template <int kLdsDepth, int kBlock>
struct KStack {
  uint32_t *lt;  // LDS tags, entry i at lt[i * kBlock]
  double *lv;
  uint32_t *gt;  // spill tags, entry i >= kLdsDepth at gt[(i - kLdsDepth) * lanes]
  double *gv;
  int lanes, gcap;
  int sp, maxsp;
  bool overflow;

  __device__ __forceinline__ void push(uint32_t tag, double v) {
    if (sp < kLdsDepth) {
      lt[sp * kBlock] = tag;
      lv[sp * kBlock] = v;
    } else if (sp - kLdsDepth < gcap) {
      gt[(size_t)(sp - kLdsDepth) * lanes] = tag;
      gv[(size_t)(sp - kLdsDepth) * lanes] = v;
    } else {
      overflow = true;
    }
    ++sp;
    maxsp = sp > maxsp ? sp : maxsp;
  }
  __device__ __forceinline__ void pop(uint32_t &tag, double &v) {
    --sp;
    if (sp < kLdsDepth) {
      tag = lt[sp * kBlock];
      v = lv[sp * kBlock];
    } else if (sp - kLdsDepth < gcap) {
      tag = gt[(size_t)(sp - kLdsDepth) * lanes];
      v = gv[(size_t)(sp - kLdsDepth) * lanes];
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

// Runtime-k result set in device memory (k up to kMaxQueryK); used for the
// K_EXPAND classification and for the query hook with k > 3.
struct MemRes {
  double *d;
  uint32_t *id;
  int k, count;

  __device__ void init() {
    for (int j = 0; j < k; ++j) {
      d[j] = DBL_MAX;
      id[j] = kNone;
    }
    count = 0;
  }
  __device__ __forceinline__ double worst() const { return d[k - 1]; }
  __device__ void add(double dist, uint32_t idx) {
    int i;
    for (i = count; i > 0; --i) {
      if (d[i - 1] > dist) {
        if (i < k) {
          d[i] = d[i - 1];
          id[i] = id[i - 1];
        }
      } else {
        break;
      }
    }
    if (i < k) {
      d[i] = dist;
      id[i] = idx;
    }
    if (count < k) ++count;
  }
};

template <class RS>
__device__ __forceinline__ void scan_leaf(const DevCloud &C, const double q0, const double q1,
                                          const double q2, const Node &nd, RS &rs) {
  const uint32_t end = nd.b & ~kLeafBit;
  const double worst = rs.worst();  // read once per leaf, as searchLevel does
  for (uint32_t t = nd.a; t < end; ++t) {
    const double2 *g = reinterpret_cast<const double2 *>(C.gen + t);
    const double2 xy = g[0];
    const double2 zw = g[1];
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
  const int sp0 = st.sp;

  // Phase 1: best-child descent to the first leaf, no pushes.
  uint32_t node = 0;
  Node nd = load_node(C.nodes, 0);
  while (!(nd.b & kLeafBit)) {
    double cut;
    node = choose(nd, pick(q0, q1, q2, nd.b), cut) ? node + 1 : nd.a;
    nd = load_node(C.nodes, node);
  }
  const uint32_t first_leaf = node;
  scan_leaf(C, q0, q1, q2, nd, rs);

  // Phase 2: walk the same path again, pushing the far children that can
  // still hold a point closer than the current worst distance.
  {
    const double worst = rs.worst();
    node = 0;
    nd = load_node(C.nodes, 0);
    while (node != first_leaf) {
      double cut;
      const uint32_t f = nd.b;
      const bool low_first = choose(nd, pick(q0, q1, q2, f), cut);
      const double m2 = mind + cut - dd[f];
      if (m2 <= worst) st.push(node, m2);
      node = low_first ? node + 1 : nd.a;
      nd = load_node(C.nodes, node);
    }
  }

  // Remaining traversal in nanoflann's order.
  while (st.sp > sp0) {
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

// PointCloud::minDistSqToBox (src/pointcloud.cpp:283-298) for the 27 shifted
// queries q + i*L, factored per dimension.  The reference adds, in dimension
// order, (lo - v)^2 or (v - hi)^2 for every coordinate v outside [lo, hi] to
// d2 = 0; adding 0.0 for an inside coordinate changes no bits, so
// d2(ix, iy, iz) = ((c[0][ix] + c[1][iy]) + c[2][iz]) is bit-identical, with
// c[d][s] computed once per query from the same shifted coordinate.
struct ImageBox {
  double c[3][3];  // [dim][shift + 1]
  __device__ __forceinline__ void init(const DevCloud &C, const double *q) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double lo = C.sb[2 * d], hi = C.sb[2 * d + 1];
#pragma unroll
      for (int s = -1; s <= 1; ++s) {
        const double v = q[d] + s * C.box[d];
        double e = 0.0;
        if (v < lo) {
          const double t = lo - v;
          e = t * t;
        } else if (v > hi) {
          const double t = v - hi;
          e = t * t;
        }
        c[d][s + 1] = e;
      }
    }
  }
  __device__ __forceinline__ double d2(int ix, int iy, int iz) const {
    return (c[0][ix + 1] + c[1][iy + 1]) + c[2][iz + 1];
  }
  // False when every image has a coordinate whose own contribution already
  // exceeds w; sums of non-negative terms round monotonically, so each such
  // image's d2 exceeds w too and the reference prunes all of them.
  __device__ __forceinline__ bool any_within(double w) const {
    bool any = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) any |= (c[d][0] <= w) | (c[d][2] <= w);
    // an image shifted in one dimension only also needs the others' 0-shift
    // terms small; those are 0 for a query inside the box and checked per
    // image otherwise.
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

// queryTreeK for a runtime k (<= kMaxQueryK) with results in `res`
// (k entries) and `merged` workspace (2k entries), both in device memory.
// Entries beyond the number of points are (kNone, DBL_MAX).
template <class ST>
__device__ void query_mem(const DevCloud &C, const double q0, const double q1, const double q2,
                          int k, MemRes &res, uint32_t *mi, double *md, ST &st) {
  knn_core(C, q0, q1, q2, res, st);
  if (!C.periodic) return;
  int m = k;
  for (int j = 0; j < k; ++j) {
    mi[j] = res.id[j];
    md[j] = res.d[j];
  }
  MemRes img{res.d + k, res.id + k, k, 0};  // image results after the k slots
  const double qv[3] = {q0, q1, q2};
  ImageBox ib;
  ib.init(C, qv);
  for (int ix = -1; ix <= 1; ix++)
    for (int iy = -1; iy <= 1; iy++)
      for (int iz = -1; iz <= 1; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        if (ib.d2(ix, iy, iz) > md[k - 1] + kTol) continue;
        const double sh[3] = {qv[0] + ix * C.box[0], qv[1] + iy * C.box[1],
                              qv[2] + iz * C.box[2]};
        knn_core(C, sh[0], sh[1], sh[2], img, st);
        for (int i = 0; i < k; ++i) {
          bool found = false;
          for (int j = 0; j < m; ++j) {
            if (mi[j] == img.id[i]) {
              md[j] = (img.d[i] < md[j]) ? img.d[i] : md[j];
              found = true;
              break;
            }
          }
          if (!found) {
            mi[m] = img.id[i];
            md[m] = img.d[i];
            ++m;
          }
        }
        // Stable insertion sort.  libstdc++'s std::sort is an insertion sort
        // (stable) up to 16 elements; beyond that its order among equal r^2
        // may differ, which only the k > 8 query hook can observe.
        for (int i = 1; i < m; ++i) {
          const double vd = md[i];
          const uint32_t vi = mi[i];
          int j = i;
          while (j > 0 && vd < md[j - 1]) {
            md[j] = md[j - 1];
            mi[j] = mi[j - 1];
            --j;
          }
          md[j] = vd;
          mi[j] = vi;
        }
        if (m > k) m = k;
      }
  for (int j = 0; j < k; ++j) {
    res.id[j] = mi[j];
    res.d[j] = md[j];
  }
}

}  // namespace vt
