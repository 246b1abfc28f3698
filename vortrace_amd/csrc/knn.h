// Device k-nearest-neighbour search over the device kd-tree (layout.h).
//
// Reproduces nanoflann 1.3.2's KDTreeSingleIndexAdaptor::knnSearch exactly
// (include/nanoflann.hpp:1221-1258 findNeighbors/knnSearch, 1005-1021
// computeInitialDistances, 1346-1408 searchLevel, 175-202 addPoint, 432-440
// evalMetric): the same best-child-first visiting order, the same incremental
// box distance `mindistsq + cut_dist - dists[idx]` and the same pruning and
// insertion rules, so ids, r^2 bits and the order among equal r^2 match.
//
// The recursion becomes a per-lane explicit stack whose first kLdsDepth
// entries live in LDS (the rest spill to device memory).  Three exact
// transformations keep it short:
//  * An entry for a far child is pushed only if its box distance is <= the
//    current worst distance: the worst distance never grows, so any entry
//    that fails now would fail when nanoflann tests it.
//  * On the first root-to-leaf descent the worst distance is still infinite.
//    Its far children go into an LDS ring holding the deepest kLdsDepth
//    levels, which becomes the stack once the first leaf is scanned.  Only if
//    an entry that fell out of the ring could still be visited is the path
//    walked again.
//  * Every entry records a lower bound (a float at or below it) of the box
//    distances of the entries below it.  Once that bound exceeds the worst
//    distance, nanoflann would prune everything still pending, so the search
//    ends there.
// Entries are PENDING(child reference, box distance of that far child, with
// the parent's cut dimension) or UNDO(dim, saved dists[dim]), which restores
// the per-dimension distance vector exactly where nanoflann's recursion would
// on return.
#pragma once
#include "device_common.h"

namespace vt {

constexpr uint32_t kSkip = 0xFFFFFFFFu;

// Launch shape shared by every kernel: lanes per workgroup and kNN stack
// entries per lane kept in LDS.
constexpr int kBlock = 128;
constexpr int kLdsDepth = 16;

// The LDS part of every lane's kNN stack: kLdsDepth 16-byte entries per lane,
// lane t's at [t * kLdsStride, t * kLdsStride + kLdsDepth).  The stride of 17
// entries (68 dwords) starts the 16 lanes of a quarter-wave on 16 different
// 4-bank groups, so 16-byte reads and writes at equal depth do not conflict.
constexpr int kLdsStride = kLdsDepth + 1;
__shared__ uint4 lds_stack[kBlock * kLdsStride];

// A stack entry: tag, the lower bound `pb` of the box distances below it, and
// a value (a box distance, or a saved dists[dim] for undo entries).
__device__ __forceinline__ uint4 pack_entry(uint32_t tag, double v, float pb) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return make_uint4(tag, __float_as_uint(pb), (uint32_t)b, (uint32_t)(b >> 32));
}
__device__ __forceinline__ void unpack_entry(const uint4 e, uint32_t &tag, double &v, float &pb) {
  tag = e.x;
  pb = __uint_as_float(e.y);
  v = __longlong_as_double((long long)(((uint64_t)e.w << 32) | e.z));
}
__device__ __forceinline__ uint4 lds_entry(int i) { return lds_stack[i]; }
__device__ __forceinline__ void set_lds_entry(int i, uint4 e) { lds_stack[i] = e; }

// Per-lane explicit stack.  The top kLdsDepth entries live in LDS: logical
// entry i of the window [lo, sp) at slot (base + i) mod kLdsDepth (base is
// non-zero only after a first descent that filled the LDS part as a ring).
// A push into a full window first moves the window's bottom entry to device
// memory (entry i at gs[i * lanes], interleaved by lane so lanes at equal
// depth coalesce); pops below the window read it back from there.
static_assert((kLdsDepth & (kLdsDepth - 1)) == 0, "kLdsDepth must be a power of two");

struct KStack {
  int tid;    // lane within the workgroup
  uint4 *gs;  // this lane's entries below the window
  int lanes, gcap;
  int sp, maxsp;
  bool overflow;
  int base;
  int lo;  // bottom of the LDS window

  __device__ __forceinline__ int slot(int i) const {
    return tid * kLdsStride + ((base + i) & (kLdsDepth - 1));
  }
  // Ring slot of depth d on a first descent (base 0).
  __device__ __forceinline__ int ring_slot(int d) const {
    return tid * kLdsStride + (d & (kLdsDepth - 1));
  }
  __device__ __forceinline__ void reset() {
    sp = 0;
    base = 0;
    lo = 0;
  }
  // The LDS slot of a new top entry (moving the window's bottom entry to
  // device memory first when the window is full); the caller writes it.
  __device__ __forceinline__ int push_slot() {
    if (sp - lo >= kLdsDepth) {
      if (lo < gcap)
        gst(gs + (size_t)lo * lanes, lds_entry(slot(lo)));
      else
        overflow = true;
      ++lo;
    }
    const int sl = slot(sp);
    ++sp;
    maxsp = sp > maxsp ? sp : maxsp;
    return sl;
  }
  __device__ __forceinline__ void push(uint32_t tag, double v, float pb) {
    set_lds_entry(push_slot(), pack_entry(tag, v, pb));
  }
  // Entry i (i < sp).  An entry lost to overflow reads as (kSkip, 0, 0):
  // skipped, and its bound of 0 stops nothing.
  __device__ __forceinline__ uint4 read(int i) const {
    if (i >= lo) return lds_entry(slot(i));
    if (i < gcap) return gld(gs + (size_t)i * lanes);
    return make_uint4(kSkip, 0u, 0u, 0u);
  }
  // Drops the top n entries.
  __device__ __forceinline__ void drop(int n) {
    sp -= n;
    lo = lo < sp ? lo : sp;
  }
  __device__ __forceinline__ void pop(uint32_t &tag, double &v, float &pb) {
    unpack_entry(read(sp - 1), tag, v, pb);
    drop(1);
  }
};

// Lower bound of a box distance m2 >= 0 as stored in the stack's `pb` bounds:
// m2 rounded to float (relative error <= 2^-24 for m2 in float's normal
// range), scaled by 1 - 2^-22 and rounded again, is below m2; above FLT_MAX
// the bound is FLT_MAX, below FLT_MIN it is 0.  Cheaper than a round-down
// conversion, and at most 2^-22 looser.
__device__ __forceinline__ float bound_below(double m2) {
  const float f = fminf((float)m2 * 0.99999976158142089844f, 3.4028234663852886e+38f);
  return m2 < 1.1754943508222875e-38 ? 0.0f : f;
}

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

// Result set for a runtime k in {1, 2, 3}, in registers (the searches of the
// ray walk).  All three entries are kept sorted: unused entries start at
// DBL_MAX and entries shifted past position k - 1 are never read, so
// addPoint's position (after every entry <= dist) is the number of the first
// two entries <= dist, and `count` is not needed.  w caches worstDist().
struct RegRes3 {
  double d[3];
  uint32_t id[3];
  int k;
  double w;

  __device__ __forceinline__ void init(double bound = DBL_MAX) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      d[j] = bound;
      id[j] = kNone;
    }
    w = bound;
  }
  __device__ __forceinline__ double worst() const { return w; }
  // addPoint for dist < worst() (a smaller distance always lands in the first
  // k entries, at the position after every entry <= dist).
  __device__ __forceinline__ void add(double dist, uint32_t idx) {
    const bool b0 = d[0] <= dist;
    const bool b1 = d[1] <= dist;  // implies b0
    const double n2 = b1 ? dist : d[1];
    const uint32_t j2 = b1 ? idx : id[1];
    const double n1 = b1 ? d[1] : (b0 ? dist : d[0]);
    const uint32_t j1 = b1 ? id[1] : (b0 ? idx : id[0]);
    const double n0 = b0 ? d[0] : dist;
    const uint32_t j0 = b0 ? id[0] : idx;
    d[0] = n0;
    d[1] = n1;
    d[2] = n2;
    id[0] = j0;
    id[1] = j1;
    id[2] = j2;
    w = sel3(n0, n1, n2, (uint32_t)(k - 1));
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

// Leaf scan (searchLevel's leaf branch, include/nanoflann.hpp:1349-1366) of
// the leaf whose reference is `ref`: the worst distance is read once, points
// are offered in slot order.
template <class RS>
__device__ __forceinline__ void scan_leaf(const DevCloud &C, const double q0, const double q1,
                                          const double q2, uint32_t ref, RS &rs) {
  const uint32_t s0 = ref & kSlotMask;
  const uint32_t end = (uint32_t)gld(&C.gen[s0].end);
  const double worst = rs.worst();
  for (uint32_t t = s0; t < end; ++t) {
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

// Child choice at an inner node: returns true when child1 is visited first,
// and the squared distance to the far child's near bound.
__device__ __forceinline__ bool choose(const DNodeV &nd, double val, double &cut) {
  const double diff1 = val - nd.lo;
  const double diff2 = val - nd.hi;
  const bool low_first = (diff1 + diff2) < 0;
  const double c = low_first ? (val - nd.hi) : (val - nd.lo);
  cut = c * c;
  return low_first;
}

__device__ __forceinline__ double pick(const double q0, const double q1, const double q2,
                                       uint32_t f) {
  return sel3(q0, q1, q2, f);
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
  st.reset();
  float ptop = INFINITY;  // bound of everything on the stack

  // Phase 1: best-child descent to the first leaf.  The far child of every
  // level goes into the LDS ring at slot depth mod kLdsDepth; the smallest
  // box distance that falls out of the ring is kept.
  uint32_t ref = C.root;
  int depth = 0;
  double lost_min = DBL_MAX;
  while (!(ref & kRefLeaf)) {
    const DNodeV nd = load_node(C.nodes, ref);
    double cut;
    const bool low_first = choose(nd, pick(q0, q1, q2, nd.dim), cut);
    const double m2 = mind + cut - sel3(dd[0], dd[1], dd[2], nd.dim);
    const int sl = st.ring_slot(depth);
    if (depth >= kLdsDepth) {
      uint32_t ot;
      double old;
      float op;
      unpack_entry(lds_entry(sl), ot, old, op);
      lost_min = old < lost_min ? old : lost_min;
    }
    set_lds_entry(sl, pack_entry(far_tag(low_first ? nd.c2 : nd.c1, nd.dim), m2, ptop));
    ptop = fminf(ptop, bound_below(m2));
    ++depth;
    ref = low_first ? nd.c1 : nd.c2;
  }
  const uint32_t first_leaf = ref;
  scan_leaf(C, q0, q1, q2, ref, rs);

  const double worst0 = rs.worst();
  if (depth <= kLdsDepth || !(lost_min <= worst0)) {
    // Every entry that left the ring would be pruned (the worst distance
    // never grows): the ring is the stack.
    st.sp = depth < kLdsDepth ? depth : kLdsDepth;
    st.base = depth > kLdsDepth ? (depth & (kLdsDepth - 1)) : 0;
    st.maxsp = st.sp > st.maxsp ? st.sp : st.maxsp;
  } else {
    // An entry that fell out of the ring can still be visited: walk the
    // same path again and push every far child that can still be visited.
    st.reset();
    ptop = INFINITY;
    ref = C.root;
    while (ref != first_leaf) {
      const DNodeV nd = load_node(C.nodes, ref);
      double cut;
      const bool low_first = choose(nd, pick(q0, q1, q2, nd.dim), cut);
      const double m2 = mind + cut - sel3(dd[0], dd[1], dd[2], nd.dim);
      if (m2 <= worst0) {
        st.push(far_tag(low_first ? nd.c2 : nd.c1, nd.dim), m2, ptop);
        ptop = fminf(ptop, bound_below(m2));
      }
      ref = low_first ? nd.c1 : nd.c2;
    }
  }

  // Remaining traversal in nanoflann's order.
  while (st.sp > 0) {
    uint32_t tag;
    double v;
    float pb;
    st.pop(tag, v, pb);
    ptop = pb;
    if (tag == kSkip) continue;
    const double worst = rs.worst();
    if ((tag & kUndo) == kUndo) {
      const uint32_t f = tag & 3u;
      dd[0] = f == 0 ? v : dd[0];
      dd[1] = f == 1 ? v : dd[1];
      dd[2] = f == 2 ? v : dd[2];
    } else if (v <= worst) {  // mindistsq * epsError <= worstDist()
      if (tag & kRefLeaf) {
        scan_leaf(C, q0, q1, q2, tag, rs);
        continue;
      }
      // Entering an inner far child: dists[idx] = cut_dist, mindistsq from
      // the entry; the old dists[idx] is restored by the undo entry.
      ref = tag & kIdxMask;
      const uint32_t pdim = (tag >> 29) & 3u;
      DNodeV nd = load_node(C.nodes, ref);
      const double c = pick(q0, q1, q2, pdim) - nd.pb;
      const double dold = sel3(dd[0], dd[1], dd[2], pdim);
      st.push(kUndo | pdim, dold, ptop);
      dd[0] = pdim == 0 ? c * c : dd[0];
      dd[1] = pdim == 1 ? c * c : dd[1];
      dd[2] = pdim == 2 ? c * c : dd[2];
      mind = v;
      for (;;) {
        double cut;
        const bool low_first = choose(nd, pick(q0, q1, q2, nd.dim), cut);
        const double m2 = mind + cut - sel3(dd[0], dd[1], dd[2], nd.dim);
        if (m2 <= worst) {
          st.push(far_tag(low_first ? nd.c2 : nd.c1, nd.dim), m2, ptop);
          ptop = fminf(ptop, bound_below(m2));
        }
        ref = low_first ? nd.c1 : nd.c2;
        if (ref & kRefLeaf) break;
        nd = load_node(C.nodes, ref);
      }
      scan_leaf(C, q0, q1, q2, ref, rs);
      continue;
    }
    // Nothing below can be visited any more: nanoflann prunes all of it.
    if ((double)pb > worst) break;
  }
  st.reset();
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

// queryTreeK for a runtime k (<= kMaxQueryK) with results in `res`
// (k entries) and `merged` workspace (2k entries), both in device memory.
// Entries beyond the number of points are (kNone, DBL_MAX).  Returns true when
// at least one periodic image was searched and merged (the merge then sorted
// k or more entries; see the note at the sort).
template <bool kPeriodic, class ST>
__device__ bool query_mem(const DevCloud &C, const double q0, const double q1, const double q2,
                          int k, MemRes &res, uint32_t *mi, double *md, ST &st) {
  knn_core(C, q0, q1, q2, res, st);
  if (!kPeriodic) return false;
  bool merged = false;
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
        merged = true;
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
        // (stable) up to 16 elements; beyond that (k > 8) it is an introsort
        // whose order among equal r^2 may differ from this one.
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
  return merged;
}

}  // namespace vt
