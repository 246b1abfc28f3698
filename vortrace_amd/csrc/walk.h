// Device ray walk: the reference's split-point traversal (src/ray.cpp:187-430,
// docs/algorithm.rst, docs/edgecases.rst) for one ray per lane.
//
// The reference keeps the ray's points in a singly linked list; new points are
// only ever inserted right after the current one, so the points ahead of the
// current one form a stack.  Here that stack keeps its top (the "next" point)
// in registers and the rest in per-lane device memory, interleaved by lane.
// Segments are handed to a sink as they close, in order, which fuses the
// reductions of src/reduction.cpp into the walk.
#pragma once
#include "knn.h"

namespace vt {

struct RayGeom {
  double a[3];  // start
  double b[3];  // end
  double d[3];  // unit direction
  double len;
};

// Ray::Ray, src/ray.cpp:14-35.
__device__ __forceinline__ void make_ray(const double *s, const double *e, RayGeom &r) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    r.a[j] = s[j];
    r.b[j] = e[j];
    r.d[j] = e[j] - s[j];
  }
  const double L = sqrt(r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2]);
  r.d[0] /= L;
  r.d[1] /= L;
  r.d[2] /= L;
  r.len = L;
}

template <int kLdsDepth, int kBlock, bool kPer>
struct WalkCtx {
  static constexpr bool kPeriodic = kPer;
  const DevCloud &C;
  KStack<kLdsDepth, kBlock> st;
  const Scratch &S;
  int gtid;       // lane index among all persistent lanes
  int lane;       // lane within the wave
  int wave;       // global wave index (for the per-wave K_EXPAND buffer)
  uint64_t nq;    // kNN queries issued
  uint64_t nslow; // K_EXPAND searches
  bool bad;       // a query found no generator (non-finite query point)

  // Ids that index device arrays must be valid: a search that inserted
  // nothing (NaN or overflowing coordinates) flags the ray and yields slot 0.
  __device__ __forceinline__ uint32_t guard(uint32_t id) {
    if (id == kNone) {
      bad = true;
      return 0;
    }
    return id;
  }

  __device__ __forceinline__ uint32_t query1(const double *q) {
    uint32_t id[1];
    double r2[1];
    query_reg<1, kPer>(C, q[0], q[1], q[2], id, r2, st);
    ++nq;
    return guard(id[0]);
  }
};

// findSplitPointDistance(pos1, pos2), src/ray.cpp:38-60.
__device__ __forceinline__ double split_analytic(const RayGeom &r, const double *p1,
                                                 const double *p2) {
  double nrm[3], ppl[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) nrm[j] = p2[j] - p1[j];
#pragma unroll
  for (int j = 0; j < 3; ++j) ppl[j] = 0.5 * (p1[j] + p2[j]) - r.a[j];
  const double ndp = nrm[0] * ppl[0] + nrm[1] * ppl[1] + nrm[2] * ppl[2];
  const double ndd = nrm[0] * r.d[0] + nrm[1] * r.d[1] + nrm[2] * r.d[2];
  if (fabs(ndd) < kParTol) return DBL_MAX;
  return ndp / ndd;
}

// findSplitPointDistance(cloud, id1, id2, s_lo, s_hi), src/ray.cpp:64-127:
// minimum-image generators in periodic mode, the analytic bisector, and a
// 64-step bisection with 1-NN probes when that falls outside the segment.
template <class X>
__device__ int split_point(X &x, const RayGeom &r, uint32_t id1, uint32_t id2, double s_lo,
                           double s_hi, double &out) {
  const DevCloud &C = x.C;
  const Gen g1 = C.gen[id1];
  const Gen g2 = C.gen[id2];
  double p1[3] = {g1.x, g1.y, g1.z};
  double p2[3] = {g2.x, g2.y, g2.z};
  if constexpr (X::kPeriodic) {
    const double s_ref = 0.5 * (s_lo + s_hi);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double ref = r.a[d] + s_ref * r.d[d];
      const double diff1 = p1[d] - ref;
      if (diff1 > C.box[d] / 2)
        p1[d] -= C.box[d];
      else if (diff1 < -C.box[d] / 2)
        p1[d] += C.box[d];
      const double diff2 = p2[d] - ref;
      if (diff2 > C.box[d] / 2)
        p2[d] -= C.box[d];
      else if (diff2 < -C.box[d] / 2)
        p2[d] += C.box[d];
    }
  }
  const double s = split_analytic(r, p1, p2);
  if (s < DBL_MAX * 0.5 && s >= s_lo - kTol && s <= s_hi + kTol) {
    out = s;
    return kErrNone;
  }
  double lo = s_lo, hi = s_hi;
  for (int it = 0; it < 64; ++it) {
    const double mid_s = 0.5 * (lo + hi);
    double mid[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) mid[j] = r.a[j] + mid_s * r.d[j];
    const uint32_t mc = x.query1(mid);
    if (mc != id1 && mc != id2) {
      out = mid_s;
      return kErrNone;
    }
    if (mc == id1)
      lo = mid_s;
    else
      hi = mid_s;
    if (hi - lo < 1e-12) {
      out = 0.5 * (lo + hi);
      return kErrNone;
    }
  }
  return kErrBisect;
}

// Ray::perpPerturbToFindCell, src/ray.cpp:129-185.
template <class X>
__device__ __noinline__ int perp_perturb(X &x, const RayGeom &r, const double *pos,
                                         uint32_t ex1, uint32_t ex2, uint32_t &out) {
  const double s2 = x.C.inv_sqrt2;
  const double w[3][3] = {{s2, s2, 0.0}, {0.0, s2, s2}, {s2, 0.0, s2}};
  int cycle = 0;
  for (int bound = 1; cycle < kMaxCycles; bound++)
    for (int a = -bound; a <= bound && cycle < kMaxCycles; a++)
      for (int b = -bound; b <= bound && cycle < kMaxCycles; b++)
        for (int c = -bound; c <= bound && cycle < kMaxCycles; c++) {
          if (a == 0 && b == 0 && c == 0) continue;
          if (abs(a) < bound && abs(b) < bound && abs(c) < bound) continue;
          double pd[3], cr[3];
          for (int j = 0; j < 3; ++j) pd[j] = a * w[0][j] + b * w[1][j] + c * w[2][j];
          cr[0] = r.d[1] * pd[2] - r.d[2] * pd[1];
          cr[1] = r.d[2] * pd[0] - r.d[0] * pd[2];
          cr[2] = r.d[0] * pd[1] - r.d[1] * pd[0];
          const double mag = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
          if (mag < 1e-15) {
            cycle++;
            continue;
          }
          for (int j = 0; j < 3; ++j) cr[j] /= mag;
          double pp[3];
          for (int j = 0; j < 3; ++j) pp[j] = pos[j] + kPertEps * cr[j];
          const uint32_t cell = x.query1(pp);
          if (cell != ex1 && cell != ex2) {
            out = cell;
            return kErrNone;
          }
          cycle++;
        }
  return kErrPerturb;
}

// The tryAlongRay lambda of src/ray.cpp:374-384.
template <class X>
__device__ uint32_t try_along_ray(X &x, const RayGeom &r, const double *pos, double sign,
                                  uint32_t ct, uint32_t nt) {
  double pp[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) pp[j] = pos[j] + sign * kPertEps * r.d[j];
  uint32_t ids[2];
  double r2[2];
  query_reg<2, X::kPeriodic>(x.C, pp[0], pp[1], pp[2], ids, r2, x.st);
  ++x.nq;
  ids[0] = x.guard(ids[0]);
  if (r2[0] + kTol < r2[1] && ids[0] != ct && ids[0] != nt) return ids[0];
  return kNone;
}

// The K_EXPAND re-classification (src/ray.cpp:320-338) for the lanes of a
// wave that need it, one lane at a time, in the wave's device buffer.
// Sets has_c/has_n; returns an error code.
template <class X>
__device__ __noinline__ int classify_expand(X &x, const double *pos, uint32_t ct, uint32_t nt,
                                            bool &has_c, bool &has_n) {
  const DevCloud &C = x.C;
  const int kb = kExpandK < (int)C.n ? kExpandK : (int)C.n;
  uint32_t *bi = x.S.big_id + (size_t)x.wave * 4 * kMaxQueryK;
  double *bd = x.S.big_d + (size_t)x.wave * 4 * kMaxQueryK;
  MemRes res{bd, bi, kb, 0};
  query_mem(C, pos[0], pos[1], pos[2], kb, res, bi + 2 * kMaxQueryK, bd + 2 * kMaxQueryK,
            x.st);
  ++x.nq;
  ++x.nslow;
  for (int i = 0; i < kb; ++i) {
    if (bd[i] - bd[0] >= kTol) break;
    if (bi[i] == ct) has_c = true;
    if (bi[i] == nt) has_n = true;
    if (i == kb - 1) {
      if (!has_c && !has_n) return kErrClassify;
      // Every entry is equidistant: which ties made the cut depends on the
      // order of std::sort in the reference, which is not reproduced.
      if (kb == kExpandK) return kErrTies;
    }
  }
  return kErrNone;
}

// Walks one ray; emits its segments to `sink` in order.  Returns an error code
// (kErrNone on success) and fills e0..e2 / i0,i1 with the values the host
// formats into the reference's message.
template <class X, class Sink>
__device__ int walk_ray(X &x, const RayGeom &r, Sink &sink, int &max_depth, double *ev,
                        uint32_t *ei) {
  const DevCloud &C = x.C;
  const Scratch &S = x.S;
  x.bad = false;
  const uint32_t c0 = x.query1(r.a);
  const uint32_t c1 = x.query1(r.b);
  if (x.bad) return kErrNonFinite;

  // current point
  uint32_t ct = c0;
  double sc = 0.0;
  // top of the pending stack ("next")
  uint32_t nt = c1;
  double sn = r.len;
  int wsp = 0;  // pending entries below the top, in device memory

  uint32_t open_cell = c0;
  double entry_s = 0.0;

  auto push_top = [&](uint32_t id, double s) -> bool {
    if (wsp >= S.wcap) return false;
    S.walk_tag[(size_t)wsp * S.lanes + x.gtid] = nt;
    S.walk_s[(size_t)wsp * S.lanes + x.gtid] = sn;
    ++wsp;
    max_depth = wsp > max_depth ? wsp : max_depth;
    nt = id;
    sn = s;
    return true;
  };

  if (c0 == c1) {
    if constexpr (!X::kPeriodic) {
      sink.seg(c0, 0.0, r.len);
      return kErrNone;
    }
    // Dyadic probes s = L*i/2^level, src/ray.cpp:220-246.
    bool found = false;
    for (int level = 1; level <= 10 && !found; level++) {
      const int n = 1 << level;
      for (int i = 1; i < n && !found; i += 2) {
        const double mid_s = r.len * i / n;
        double mid[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) mid[j] = r.a[j] + mid_s * r.d[j];
        const uint32_t mc = x.query1(mid);
        if (mc != c0) {
          if (!push_top(mc, mid_s)) return kErrWalkStack;
          found = true;
        }
      }
    }
    if (x.bad) return kErrNonFinite;
    if (!found) return kErrPeriodicStart;
  }

  int iteration = 0;
  for (;;) {
    if (x.bad) return kErrNonFinite;
    if (++iteration > kMaxIter) return kErrMaxIter;
    // Zero-width step (two points inserted at one split point).
    if (sn - sc < 1e-14) {
      ct = nt;
      sc = sn;
      open_cell = ct;
      if (wsp == 0) break;
      --wsp;
      nt = S.walk_tag[(size_t)wsp * S.lanes + x.gtid];
      sn = S.walk_s[(size_t)wsp * S.lanes + x.gtid];
      continue;
    }
    double s;
    int rc = split_point(x, r, ct, nt, sc, sn, s);
    if (rc) {
      ei[0] = ct;
      ei[1] = nt;
      return rc;
    }
    if (s < sc - kTol || s > sn + kTol) {
      ev[0] = s;
      ev[1] = sc;
      ev[2] = sn;
      ei[0] = ct;
      ei[1] = nt;
      return kErrSplitOutside;
    }
    {
      const double lo = sc + 1e-15, hi = sn - 1e-15;
      s = (s < lo) ? lo : ((hi < s) ? hi : s);  // std::clamp
    }
    double pos[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) pos[j] = r.a[j] + s * r.d[j];

    uint32_t ids[3];
    double r2[3];
    query_reg<3, X::kPeriodic>(C, pos[0], pos[1], pos[2], ids, r2, x.st);
    ++x.nq;
    ids[0] = x.guard(ids[0]);
    if (x.bad) return kErrNonFinite;
    bool has_c = false, has_n = false, three_eq = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (three_eq) {
        if (r2[i] - r2[0] >= kTol) {
          three_eq = false;
        } else {
          if (ids[i] == ct) has_c = true;
          if (ids[i] == nt) has_n = true;
        }
      }
    }
    if (three_eq && !has_c && !has_n) {
      // Serialise the lanes of this wave that need the wide search.
      rc = kErrNone;
      uint64_t want = __ballot(1);
      const int me = x.lane;
      while (want) {
        const int l = __ffsll((unsigned long long)want) - 1;
        if (l == me) rc = classify_expand(x, pos, ct, nt, has_c, has_n);
        want &= want - 1;
      }
      if (rc) return rc;
    }

    if (has_c || has_n) {
      // Case 2: the c/n boundary.
      sink.seg(open_cell, entry_s, s);
      open_cell = nt;
      entry_s = s;
      ct = nt;
      sc = sn;
      if (wsp == 0) break;
      --wsp;
      nt = S.walk_tag[(size_t)wsp * S.lanes + x.gtid];
      sn = S.walk_s[(size_t)wsp * S.lanes + x.gtid];
    } else if (r2[0] + kTol < r2[1]) {
      // Case 1: one cell strictly closer; insert it.
      if (!push_top(ids[0], s)) return kErrWalkStack;
    } else {
      // Case 1a: resolve each side by perturbation along, then across, the ray.
      uint32_t scell = try_along_ray(x, r, pos, -1.0, ct, nt);
      if (scell == kNone) {
        double pc[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) pc[j] = pos[j] - kPertEps * r.d[j];
        rc = perp_perturb(x, r, pc, ct, nt, scell);
        if (rc) return rc;
      }
      if (x.bad) return kErrNonFinite;
      uint32_t ncell = try_along_ray(x, r, pos, 1.0, ct, nt);
      if (ncell == kNone) {
        double pn[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) pn[j] = pos[j] + kPertEps * r.d[j];
        rc = perp_perturb(x, r, pn, ct, nt, ncell);
        if (rc) return rc;
      }
      if (x.bad) return kErrNonFinite;
      if (scell == ncell) {
        if (!push_top(scell, s)) return kErrWalkStack;
      } else {
        // current -> scell -> ncell -> next, both at s
        if (!push_top(ncell, s)) return kErrWalkStack;
        if (!push_top(scell, s)) return kErrWalkStack;
      }
    }
  }
  sink.seg(open_cell, entry_s, sc);
  return kErrNone;
}

}  // namespace vt
