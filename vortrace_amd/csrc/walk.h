// Device ray walk: the reference's split-point traversal (src/ray.cpp:187-430,
// docs/algorithm.rst, docs/edgecases.rst), one ray per lane.
//
// The reference keeps the ray's points in a singly linked list; new points are
// only ever inserted right after the current one, so the points ahead of the
// current one form a stack.  Here that stack keeps its top (the "next" point)
// in registers and the rest in per-lane device memory, interleaved by lane.
//
// The walk is a re-entrant state machine around ONE nearest-neighbour query:
// prepare() advances the lane to its next query (endpoint, dyadic probe,
// bisection probe, split-point classification, along-ray or perpendicular
// perturbation) and returns its point and k; consume() takes the query's
// result and moves the state on.  The caller runs the query in between, so a
// wave can interleave the walks and kNN traversals of its lanes freely.  Each
// step makes exactly the queries, in the order and with the arithmetic, of the
// reference.
//
// Segments are handed to a sink as they close, in order, which fuses the
// reductions of src/reduction.cpp into the walk.
#pragma once
#include "knn.h"

namespace vt {

// Ray::Ray, src/ray.cpp:14-35: the unit direction and length.  The end point
// is not kept; the one query at the end reloads it.
struct RayGeom {
  double a[3];  // start
  double d[3];  // unit direction
  double len;
};

__device__ __forceinline__ void make_ray(const double *s, const double *e, RayGeom &r) {
  double b[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    r.a[j] = gld(s + j);
    b[j] = gld(e + j);
    r.d[j] = b[j] - r.a[j];
  }
  const double L = sqrt(r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2]);
  r.d[0] /= L;
  r.d[1] /= L;
  r.d[2] /= L;
  r.len = L;
}

// Lane identity and the per-lane scratch pointers a kNN stack needs.
struct LaneInfo {
  int tid;   // lane within the workgroup
  int gtid;  // lane among all persistent lanes
  int wave;  // global wave index (per-wave K_EXPAND buffer)
  int lane;  // lane within the wave
};

__device__ __forceinline__ KStack make_stack(const Scratch &S, const LaneInfo &L) {
  return KStack{L.tid, S.knn_spill + L.gtid, S.lanes, S.kspill, 0, 0, false, 0, 0};
}

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

// The analytic part of findSplitPointDistance(cloud, id1, id2, s_lo, s_hi)
// (src/ray.cpp:64-94): minimum-image generators in periodic mode, then the
// bisector.  Returns false when the reference falls back to bisection.
// p1 / p2 receive the two generators as used (minimum images).
template <bool kPer>
__device__ __forceinline__ bool split_point(const DevCloud &C, const RayGeom &r, uint32_t id1,
                                            uint32_t id2, double s_lo, double s_hi,
                                            double &out, double (&p1)[3], double (&p2)[3]) {
  const double2 *g1 = reinterpret_cast<const double2 *>(C.gen + id1);
  const double2 *g2 = reinterpret_cast<const double2 *>(C.gen + id2);
  const double2 a1 = gld(g1), b1 = gld(g1 + 1), a2 = gld(g2), b2 = gld(g2 + 1);
  p1[0] = a1.x;
  p1[1] = a1.y;
  p1[2] = b1.x;
  p2[0] = a2.x;
  p2[1] = a2.y;
  p2[2] = b2.x;
  if constexpr (kPer) {
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
  out = s;
  return s < DBL_MAX * 0.5 && s >= s_lo - kTol && s <= s_hi + kTol;
}

// The K_EXPAND re-classification (src/ray.cpp:320-338), in the wave's device
// buffer; the caller runs the lanes of a wave that need it one at a time.
// The result is packed as error | has_c << 8 | has_n << 9.
template <bool kPer>
__device__ __forceinline__ int classify_expand(const DevCloud &C, const Scratch &S, int tid,
                                            int gtid, double x, double y, double z,
                                            uint32_t ct, uint32_t nt) {
  KStack st{tid, S.knn_spill + gtid, S.lanes, S.kspill, 0, 0, false, 0, 0};
  const int wave = gtid >> 6;
  const int kb = kExpandK < (int)C.n ? kExpandK : (int)C.n;
  uint32_t *bi = S.big_id + (size_t)wave * 4 * kMaxQueryK;
  double *bd = S.big_d + (size_t)wave * 4 * kMaxQueryK;
  MemRes res{bd, bi, kb, 0};
  const bool merged =
      query_mem<kPer>(C, x, y, z, kb, res, bi + 2 * kMaxQueryK, bd + 2 * kMaxQueryK, st);
  if (st.overflow) return kErrKnnStack;
  bool has_c = false, has_n = false;
  const double d0 = gld(bd);
  for (int i = 0; i < kb; ++i) {
    if (gld(bd + i) - d0 >= kTol) break;
    const uint32_t id = gld(bi + i);
    if (id == ct) has_c = true;
    if (id == nt) has_n = true;
    if (i == kb - 1) {
      if (!has_c && !has_n) return kErrClassify;
      // Every entry is equidistant.  When no periodic image was merged the
      // 100 entries are nanoflann's own (ids and tie order reproduced), so
      // the result stands.  After an image merge, which ties survive depends
      // on libstdc++'s std::sort (an introsort beyond 16 entries), which is
      // not reproduced: refuse rather than guess.
      if (merged && kb == kExpandK) return kErrTies;
    }
  }
  return (has_c ? 1 << 8 : 0) | (has_n ? 1 << 9 : 0);
}

// Walk states.  A state names the query the lane issues next.
enum : int {
  kStStart = 0,  // 1-NN of the ray start
  kStEnd,        // 1-NN of the ray end
  kStDyadic,     // periodic start: 1-NN at s = L * i / 2^level
  kStSplit,      // (no query) next split point; becomes Classify or Bisect
  kStBisect,     // 1-NN probe of the split-point bisection
  kStClassify,   // 3-NN at the split point
  kStAlongC,     // 2-NN at pos - eps * dir (case 1a, c side)
  kStAlongN,     // 2-NN at pos + eps * dir (case 1a, n side)
  kStPerpC,      // 1-NN of the perpendicular cycling, c side
  kStPerpN,      // 1-NN of the perpendicular cycling, n side
  kStFinal,      // (no query) the last segment [entry_s, sc] is still open
  kStDone,       // every segment has been emitted
};

// Error details for the host's message (the reference's exception texts).
struct WalkErr {
  double v[3];
  uint32_t id[2];
};

// One ray's walk.  Per-lane scratch: the pending stack below its top lives in
// S.walk_tag / S.walk_s at [depth * S.lanes + wbase]; the entry right below
// the top is also kept in registers, so taking it off the stack does not wait
// for memory (its replacement is read then, and needed one pop later).
//
// Any two adjacent points of the list are walked the same way whatever the
// rest of the list holds (src/ray.cpp:250-424 reads only the current point,
// the next one and the ray), so the walk of a ray can be split by intervals:
// another lane may take the bottom pair of the pending stack (the interval
// nearest the ray's end, steal_bottom) and walk it on its own
// (begin_interval); its case-2 boundaries and zero-width switches, replayed
// in order after this lane's own, give the same segments.  Entries below
// `bot` have been given away.
template <bool kPer>
struct Walker {
  RayGeom r;
  uint32_t ct, nt;  // current point, and the top of the pending stack ("next")
  double sc, sn;
  int wsp;  // pending entries below the top (entries bot .. wsp-2 in device memory)
  int bot;  // entries below bot were given to other lanes
  uint32_t nt2;  // entry wsp-1, when wsp > bot
  double sn2;
  uint32_t open_cell;
  double entry_s;
  int iteration;
  int state;
  double s;            // current split point
  uint32_t c0;         // start cell
  uint32_t side_cell;  // case 1a: the c-side cell once resolved
  // per-state registers: dyadic (level, i), bisection (it, lo, hi),
  // perpendicular cycling (bound, a, b, c, cycle)
  int ia, ib, ic, id, ie;
  double lo, hi;

  __device__ __forceinline__ void begin(const double *s0, const double *e0) {
    make_ray(s0, e0, r);
    ct = nt = 0;
    sc = 0.0;
    sn = r.len;
    wsp = 0;
    bot = 0;
    nt2 = 0;
    sn2 = 0.0;
    open_cell = 0;
    entry_s = 0.0;
    iteration = 0;
    state = kStStart;
    s = 0.0;
    c0 = side_cell = 0;
    ia = ib = ic = id = ie = 0;
    lo = hi = 0.0;
  }

  __device__ __forceinline__ bool push_top(const Scratch &S, size_t wbase, int &max_depth,
                                           uint32_t cell, double sv) {
    if (wsp == bot) wsp = bot = 0;  // nothing below the top: start the memory part over
    if (wsp >= S.wcap) return false;
    if (wsp > bot) {
      gst(S.walk_tag + (size_t)(wsp - 1) * S.lanes + wbase, nt2);
      gst(S.walk_s + (size_t)(wsp - 1) * S.lanes + wbase, sn2);
    }
    nt2 = nt;
    sn2 = sn;
    ++wsp;
    max_depth = wsp - bot > max_depth ? wsp - bot : max_depth;
    nt = cell;
    sn = sv;
    return true;
  }
  // Advance the current point to the next one; false when none is left.
  __device__ __forceinline__ bool advance(const Scratch &S, size_t wbase) {
    ct = nt;
    sc = sn;
    if (wsp == bot) return false;
    --wsp;
    nt = nt2;
    sn = sn2;
    if (wsp > bot) {
      nt2 = gld(S.walk_tag + (size_t)(wsp - 1) * S.lanes + wbase);
      sn2 = gld(S.walk_s + (size_t)(wsp - 1) * S.lanes + wbase);
    }
    return true;
  }
  // Whether a pending pair below the top can be given away: at least
  // `min_depth` (>= 1) entries below the top.
  __device__ __forceinline__ bool can_give(int min_depth) const { return wsp - bot >= min_depth; }
  // Gives the bottom pair (p, q) of the pending list away: q leaves the list,
  // p becomes its last point.  q is entry bot (in registers as nt2 when it is
  // entry wsp-1); p is entry bot + 1 (nt2 when that is entry wsp-1, the top
  // itself when bot + 1 == wsp).
  __device__ __forceinline__ void steal_bottom(const Scratch &S, size_t wbase, uint32_t &pc,
                                               double &ps, uint32_t &qc, double &qs) {
    const int d = wsp - bot;
    if (d == 1) {
      qc = nt2;
      qs = sn2;
      pc = nt;
      ps = sn;
    } else {
      qc = gld(S.walk_tag + (size_t)bot * S.lanes + wbase);
      qs = gld(S.walk_s + (size_t)bot * S.lanes + wbase);
      if (d == 2) {
        pc = nt2;
        ps = sn2;
      } else {
        pc = gld(S.walk_tag + (size_t)(bot + 1) * S.lanes + wbase);
        ps = gld(S.walk_s + (size_t)(bot + 1) * S.lanes + wbase);
      }
    }
    ++bot;
  }
  // Walks the interval between points (pc, ps) and (qc, qs) of the ray from
  // s0 to e0, as the ray's own walk would from the moment that pair is the
  // current one.
  __device__ __forceinline__ void begin_interval(const double *s0, const double *e0, uint32_t pc,
                                                 double ps, uint32_t qc, double qs) {
    begin(s0, e0);
    ct = pc;
    sc = ps;
    nt = qc;
    sn = qs;
    c0 = pc;
    open_cell = pc;
    state = kStSplit;
  }
  // A pair given away came back unwalked (its walker ran out of room): it
  // becomes the current pair, with nothing pending below it.
  __device__ __forceinline__ void take_back(uint32_t pc, double ps, uint32_t qc, double qs) {
    ct = pc;
    sc = ps;
    nt = qc;
    sn = qs;
    wsp = bot = 0;
    state = kStSplit;
  }
  // The split point was found (analytically or by bisection): bounds check,
  // clamp (src/ray.cpp:280-291), then classify.
  __device__ __forceinline__ int accept_split(double sv, WalkErr &e) {
    if (sv < sc - kTol || sv > sn + kTol) {
      e.v[0] = sv;
      e.v[1] = sc;
      e.v[2] = sn;
      e.id[0] = ct;
      e.id[1] = nt;
      return kErrSplitOutside;
    }
    const double l2 = sc + 1e-15, h2 = sn - 1e-15;
    s = (sv < l2) ? l2 : ((h2 < sv) ? h2 : sv);  // std::clamp
    state = kStClassify;
    return kErrNone;
  }
  // Case 1a resolved on both sides: insert one or two points at s.
  __device__ __forceinline__ int resolve_1a(const Scratch &S, size_t wbase, int &max_depth,
                                            uint32_t scell, uint32_t ncell) {
    if (scell == ncell) {
      if (!push_top(S, wbase, max_depth, scell, s)) return kErrWalkStack;
    } else {  // current -> scell -> ncell -> next, both at s
      if (!push_top(S, wbase, max_depth, ncell, s)) return kErrWalkStack;
      if (!push_top(S, wbase, max_depth, scell, s)) return kErrWalkStack;
    }
    state = kStSplit;
    return kErrNone;
  }
  // Perpendicular cycling (src/ray.cpp:129-185): move to the next candidate
  // direction that needs a query, counting degenerate ones as cycles.
  // Returns false when MAX_PERTURBATION_CYCLES is reached.
  __device__ __forceinline__ bool perp_next(const DevCloud &C, double (&u)[3]) {
    const double s2 = C.inv_sqrt2;
    for (;;) {
      if (ie >= kMaxCycles) return false;
      // iterator (bound = ia, a = ib, b = ic, c = id) in the reference's loop order
      if (++id > ia) {
        id = -ia;
        if (++ic > ia) {
          ic = -ia;
          if (++ib > ia) {
            ++ia;
            ib = ic = id = -ia;
          }
        }
      }
      const int a = ib, b = ic, c = id;
      if (a == 0 && b == 0 && c == 0) continue;
      if (abs(a) < ia && abs(b) < ia && abs(c) < ia) continue;
      const double w[3][3] = {{s2, s2, 0.0}, {0.0, s2, s2}, {s2, 0.0, s2}};
      double pd[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) pd[j] = a * w[0][j] + b * w[1][j] + c * w[2][j];
      u[0] = r.d[1] * pd[2] - r.d[2] * pd[1];
      u[1] = r.d[2] * pd[0] - r.d[0] * pd[2];
      u[2] = r.d[0] * pd[1] - r.d[1] * pd[0];
      const double mag = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      if (mag < 1e-15) {
        ++ie;
        continue;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) u[j] /= mag;
      return true;
    }
  }
  __device__ __forceinline__ void perp_start(int st_next) {
    ia = 1;   // bound
    ib = -1;  // a
    ic = -1;  // b
    id = -2;  // c (the first perp_next moves it to -1)
    ie = 0;   // cycle
    state = st_next;
  }

  // Advances to the next query: on success with state < kStFinal, q and k are
  // the query; with kStFinal the ray has no query left.  `bound`: the search
  // need only find neighbours closer than it (DBL_MAX: all of them).
  //
  // The classification after an analytic split point reads its 3-NN result
  // only through the entries within kTol of the nearest (ids and r^2) and
  // through the tests r2[i] - r2[0] >= kTol, r2[0] + kTol < r2[1] for the
  // rest.  The nearest is at most as far as the nearer of the two cells the
  // split point lies between, m; so with bound B = m (1 + 1e-12) + 2 kTol
  // every entry within kTol of the nearest is found exactly, and an entry at
  // or beyond B (left as the placeholder (none, B)) passes both tests as
  // the true one does.  When 2 kTol is below the resolution of m, no bound.
  template <class Sink>
  __device__ __forceinline__ int prepare(const DevCloud &C, const Scratch &S, size_t wbase,
                                         const double *end, double (&q)[3], int &k,
                                         double &bound, WalkErr &e, Sink &sink) {
    bound = DBL_MAX;
    double p1[3], p2[3];
    bool have_pair = false;
    if (state == kStSplit) {
      for (;;) {
        if (++iteration > kMaxIter) return kErrMaxIter;
        if (sn - sc < 1e-14) {  // zero-width step
          const bool more = advance(S, wbase);
          open_cell = ct;
          sink.sw(ct);
          if (!more) {
            state = kStFinal;
            return kErrNone;
          }
          continue;
        }
        double sv;
        if (split_point<kPer>(C, r, ct, nt, sc, sn, sv, p1, p2)) {
          const int rc = accept_split(sv, e);
          if (rc) return rc;
          have_pair = true;
        } else {
          lo = sc;
          hi = sn;
          ia = 0;
          state = kStBisect;
        }
        break;
      }
    }
    k = 1;
    if (state == kStStart) {
#pragma unroll
      for (int j = 0; j < 3; ++j) q[j] = r.a[j];
    } else if (state == kStEnd) {
#pragma unroll
      for (int j = 0; j < 3; ++j) q[j] = gld(end + j);
    } else if (state == kStDyadic) {
      const double mid_s = r.len * ib / (1 << ia);
#pragma unroll
      for (int j = 0; j < 3; ++j) q[j] = r.a[j] + mid_s * r.d[j];
    } else if (state == kStBisect) {
      const double mid_s = 0.5 * (lo + hi);
#pragma unroll
      for (int j = 0; j < 3; ++j) q[j] = r.a[j] + mid_s * r.d[j];
    } else {
      double pos[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) pos[j] = r.a[j] + s * r.d[j];
      if (state == kStClassify) {
#pragma unroll
        for (int j = 0; j < 3; ++j) q[j] = pos[j];
        k = 3;
        if (have_pair) {
          double d1 = 0.0, d2 = 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double e1 = pos[j] - p1[j], e2 = pos[j] - p2[j];
            d1 = d1 + e1 * e1;
            d2 = d2 + e2 * e2;
          }
          const double m = (d1 < d2 ? d1 : d2) * (1.0 + 1e-12);
          const double b = m + 2.0 * kTol;
          if (b - m >= 1.5 * kTol) bound = b;
        }
      } else if (state == kStAlongC || state == kStAlongN) {
        const double sign = state == kStAlongC ? -1.0 : 1.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) q[j] = pos[j] + sign * kPertEps * r.d[j];
        k = 2;
      } else {  // kStPerpC / kStPerpN: around pos -/+ eps * dir
        double u[3];
        if (!perp_next(C, u)) return kErrPerturb;
        const bool c_side = state == kStPerpC;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double base = c_side ? pos[j] - kPertEps * r.d[j] : pos[j] + kPertEps * r.d[j];
          q[j] = base + kPertEps * u[j];
        }
      }
    }
    return kErrNone;
  }

  // Takes the result of the query prepare() asked for.  May emit segments; on
  // return the state is the next query's, kStSplit, kStFinal or kStDone.
  // Without kExpand, a ray that needs the K_EXPAND classification returns
  // kDeferRay.  With it, `nslow` counts K_EXPAND searches, and every lane of
  // the wave that calls consume() must do so together (K_EXPAND runs those
  // lanes one at a time).
  template <bool kExpand, class Sink>
  __device__ __forceinline__ int consume(const DevCloud &C, const Scratch &S, size_t wbase,
                                         int tid, int gtid, const uint32_t (&ids)[3],
                                         const double (&r2)[3], Sink &sink, int &max_depth,
                                         uint64_t &nslow, WalkErr &e) {
    // Ids index device arrays: a search that inserted nothing (NaN or
    // overflowing coordinates) fails the ray instead.
    if (ids[0] == kNone) return kErrNonFinite;
    const int wlane = tid & 63;
    if (state == kStStart) {
      c0 = ids[0];
      state = kStEnd;
    } else if (state == kStEnd) {
      const uint32_t c1 = ids[0];
      ct = c0;
      nt = c1;
      open_cell = c0;
      if (c0 != c1) {
        state = kStSplit;
      } else if constexpr (!kPer) {
        sink.seg(c0, 0.0, r.len);
        state = kStDone;
      } else {
        ia = 1;  // level
        ib = 1;  // i
        state = kStDyadic;
      }
    } else if (state == kStDyadic) {
      if (ids[0] != c0) {
        const double mid_s = r.len * ib / (1 << ia);
        if (!push_top(S, wbase, max_depth, ids[0], mid_s)) return kErrWalkStack;
        state = kStSplit;
      } else {
        ib += 2;
        if (ib >= (1 << ia)) {
          ++ia;
          ib = 1;
          if (ia > 10) return kErrPeriodicStart;
        }
      }
    } else if (state == kStBisect) {
      const double mid_s = 0.5 * (lo + hi);
      const uint32_t mc = ids[0];
      int rc = kErrNone;
      if (mc != ct && mc != nt) {
        rc = accept_split(mid_s, e);
      } else {
        if (mc == ct)
          lo = mid_s;
        else
          hi = mid_s;
        if (hi - lo < 1e-12) {
          rc = accept_split(0.5 * (lo + hi), e);
        } else if (++ia >= 64) {
          e.id[0] = ct;
          e.id[1] = nt;
          return kErrBisect;
        }
      }
      if (rc) return rc;
    } else if (state == kStClassify) {
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
        if constexpr (!kExpand) return kDeferRay;
        double pos[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) pos[j] = r.a[j] + s * r.d[j];
        int res = 0;
        uint64_t want = __ballot(1);
        while (want) {
          const int l = __ffsll((unsigned long long)want) - 1;
          if (l == wlane) {
            res = classify_expand<kPer>(C, S, tid, gtid, pos[0], pos[1], pos[2], ct, nt);
            ++nslow;
          }
          want &= want - 1;
        }
        if (res & 0xff) return res & 0xff;
        has_c = (res >> 8) & 1;
        has_n = (res >> 9) & 1;
      }
      if (has_c || has_n) {
        // Case 2: the c/n boundary.
        sink.close(open_cell, entry_s, s, nt);
        open_cell = nt;
        entry_s = s;
        state = advance(S, wbase) ? kStSplit : kStFinal;
      } else if (r2[0] + kTol < r2[1]) {
        // Case 1: one cell strictly closer; insert it.
        if (!push_top(S, wbase, max_depth, ids[0], s)) return kErrWalkStack;
        state = kStSplit;
      } else {
        // Case 1a: resolve each side by perturbation along, then across, the ray.
        state = kStAlongC;
      }
    } else if (state == kStAlongC || state == kStAlongN) {
      const bool ok = r2[0] + kTol < r2[1] && ids[0] != ct && ids[0] != nt;
      if (state == kStAlongC) {
        if (ok) {
          side_cell = ids[0];
          state = kStAlongN;
        } else {
          perp_start(kStPerpC);
        }
      } else if (ok) {
        const int rc = resolve_1a(S, wbase, max_depth, side_cell, ids[0]);
        if (rc) return rc;
      } else {
        perp_start(kStPerpN);
      }
    } else {  // kStPerpC / kStPerpN
      const uint32_t cell = ids[0];
      if (cell != ct && cell != nt) {
        if (state == kStPerpC) {
          side_cell = cell;
          state = kStAlongN;
        } else {
          const int rc = resolve_1a(S, wbase, max_depth, side_cell, cell);
          if (rc) return rc;
        }
      } else {
        ++ie;  // cycle
      }
    }
    return kErrNone;
  }
};

}  // namespace vt
