// The ray engine: every lane of a persistent wave owns one ray at a time and
// advances it one small step per loop iteration.
//
// A ray alternates between its walk (walk.h: consume a query result, emit
// segments, prepare the next query) and one kNN search at a time
// (PointCloud::queryTreeK, src/pointcloud.cpp:308-368, over nanoflann 1.3.2's
// searchLevel, include/nanoflann.hpp:1346-1408).  A search is unrolled into
// node steps, which read one 32-byte inner node, and leaf steps, which read a
// whole leaf's generators at once and then pop the stack.  Lanes of a wave run
// their own searches at their own depths; a lane whose search ends does not
// wait for the others.
//
// Leaf steps and walk steps are long compared with node steps.  The wave runs
// node steps every iteration of a tight inner loop and leaf steps only once
// S.leaf_batch of its lanes are at a leaf (or none is at a node); it leaves
// the loop once S.walk_batch lanes wait for a walk step (or none is
// searching), and then every waiting lane runs its walk step.
//
// Each search visits nodes and leaves, inserts points and prunes entries in
// exactly nanoflann's order with nanoflann's arithmetic (see knn.h for the two
// exact transformations of its recursion used here: prune-at-push and the
// first-descent ring), so ids, r^2 bits and tie order are the reference's.
#pragma once
#include "walk.h"

namespace vt {

// Search modes.
enum : int {
  kTmNode = 0,   // visit inner node `cur`
  kTmEnter = 1,  // visit inner node `cur`, a far child taken off the stack
  kTmLeaf = 2,   // scan the leaf whose reference is `cur`
  kTmPop = 3,    // take the next entry off the stack
  kTmDone = 4,   // search finished
};

// Generators a leaf step loads at once.
constexpr int kLeafChunk = 5;

// One queryTreeK for k in {1, 2, 3}: the plain search, then in periodic mode
// the image searches, each run as a sequence of node and leaf steps.
template <bool kPer>
struct Search {
  double q[3];  // point of the running search (an image in periodic mode)
  double dd[3];
  double mind;
  RegRes3 rs;
  uint32_t cur;  // inner node index, or leaf reference
  int mode;
  int depth;            // >= 0 on the first descent (entries in the LDS ring), else -1
  double lost_min;      // smallest box distance that fell out of the ring
  uint32_t first_leaf;  // first descent's leaf; a re-descent stops there
  double pv;            // popped entry: box distance of the far child
  uint32_t pdim;        // popped entry: the parent's cut dimension
  float ptop;           // lower bound of the box distances on the stack
  // periodic images (src/pointcloud.cpp:327-365)
  double qb[3];  // the query point itself
  uint32_t mi[3];  // merged result so far (m <= k entries between images)
  double md[3];
  int m;
  int img;  // -1: the plain search; else the image (ix+1)*9 + (iy+1)*3 + (iz+1)

  // computeInitialDistances against root_bbox, then the first descent.  With
  // a bound, only points closer than it are looked for (a result set
  // pre-filled with `bound` and no ids).
  __device__ __forceinline__ void begin_search(const DevCloud &C, KStack &st, double x,
                                               double y, double z,
                                               double bound = DBL_MAX) {
    q[0] = x;
    q[1] = y;
    q[2] = z;
    rs.init(bound);
    mind = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      dd[i] = 0.0;
      if (q[i] < C.rb_lo[i]) {
        dd[i] = (q[i] - C.rb_lo[i]) * (q[i] - C.rb_lo[i]);
        mind += dd[i];
      }
      if (q[i] > C.rb_hi[i]) {
        dd[i] = (q[i] - C.rb_hi[i]) * (q[i] - C.rb_hi[i]);
        mind += dd[i];
      }
    }
    st.reset();
    ptop = INFINITY;
    cur = C.rroot;
    mode = (cur & kRefLeaf) ? kTmLeaf : kTmNode;
    depth = 0;
    lost_min = DBL_MAX;
    first_leaf = cur;
  }

  __device__ __forceinline__ void start(const DevCloud &C, KStack &st, const double (&x)[3],
                                        int k, double bound = DBL_MAX) {
    qb[0] = x[0];
    qb[1] = x[1];
    qb[2] = x[2];
    rs.k = k;
    img = -1;
    begin_search(C, st, x[0], x[1], x[2], bound);
  }

  // Component f of q / dd.  Values are selected, never addresses: a select
  // of element addresses would keep the struct out of registers.
  __device__ __forceinline__ double qd(uint32_t f) const { return sel3(q[0], q[1], q[2], f); }
  __device__ __forceinline__ double ddd(uint32_t f) const {
    return sel3(dd[0], dd[1], dd[2], f);
  }
  __device__ __forceinline__ void set_dd(uint32_t f, double v) {
    dd[0] = f == 0 ? v : dd[0];
    dd[1] = f == 1 ? v : dd[1];
    dd[2] = f == 2 ? v : dd[2];
  }

  // One level of searchLevel at inner node nd: choose the child to visit
  // first and push the other (or, on the first descent, put it in the LDS
  // ring).  Returns the child to visit.
  __device__ __forceinline__ uint32_t level(KStack &st, const DNodeV &nd) {
    const uint32_t f = nd.dim;
    double cut;
    const bool low_first = choose(nd, qd(f), cut);
    const double m2 = mind + cut - ddd(f);
    const uint32_t tag = far_tag(low_first ? nd.c2 : nd.c1, f);
    if (depth >= 0) {
      // First descent (no point found yet, worst = the search's bound): every
      // far child that can still be visited goes into the LDS ring, the
      // depth-th at slot depth mod kLdsDepth; one beyond the bound never can.
      if (m2 <= rs.worst()) {
        const int sl = st.ring_slot(depth);
        if (depth >= kLdsDepth) {
          uint32_t ot;
          double old;
          float op;
          unpack_entry(lds_entry(sl), ot, old, op);
          lost_min = old < lost_min ? old : lost_min;
        }
        set_lds_entry(sl, pack_entry(tag, m2, ptop));
        ptop = fminf(ptop, bound_below(m2));
        ++depth;
      }
    } else if (m2 <= rs.worst()) {
      st.push(tag, m2, ptop);
      ptop = fminf(ptop, bound_below(m2));
    }
    return low_first ? nd.c1 : nd.c2;
  }

  // One step at inner node `cur` (modes kTmNode / kTmEnter): one pair record
  // is read; an even-depth node and its near child are searched from it (two
  // levels), an odd-depth node (entered off the stack) on its own.
  __device__ __forceinline__ void node_step(const DevCloud &C, KStack &st) {
    const bool odd = (cur & kRefOdd) != 0u;
    const DRecV R = load_rec(C.recs, odd ? (cur & (kRefOdd - 1u)) >> 1 : cur);
    const DNodeV n0 = odd ? rec_child(R, cur & 1u) : R.x;
    if (mode == kTmEnter) {
      // Entering the far child of a popped entry: dists[idx] = cut_dist,
      // mindistsq from the entry; the old dists[idx] is restored on return.
      const double c = qd(pdim) - n0.pb;
      st.push(kUndo | pdim, ddd(pdim), ptop);
      set_dd(pdim, c * c);
      mind = pv;
    }
    uint32_t near = level(st, n0);
    if (!odd && !(near & kRefLeaf)) near = level(st, rec_child(R, near & 1u));
    cur = near;
    mode = kTmNode;
    if (near & kRefLeaf) {
      mode = kTmLeaf;
      if (depth < 0 && near == first_leaf) mode = kTmPop;  // re-descent is back at the first leaf
      if (depth >= 0) first_leaf = near;
    }
  }

  // Takes entries off the stack until one is visited (mode kTmEnter or
  // kTmLeaf) or nothing left can be (kTmDone).  The top kPopBatch entries are
  // read at once and resolved in registers, top first.
  static constexpr int kPopBatch = 2;
  __device__ __forceinline__ void pop(KStack &st) {
    const double worst = rs.worst();
    while (st.sp > 0) {
      uint4 e[kPopBatch];
#pragma unroll
      for (int j = 0; j < kPopBatch; ++j)
        if (j < st.sp) e[j] = st.read(st.sp - 1 - j);
      int used = 0;
      bool found = false, stop = false;
#pragma unroll
      for (int j = 0; j < kPopBatch; ++j) {
        if (found || stop || j >= st.sp) continue;
        used = j + 1;
        uint32_t tag;
        double v;
        float pb;
        unpack_entry(e[j], tag, v, pb);
        ptop = pb;
        if (tag == kSkip) continue;
        if ((tag & kUndo) == kUndo) {
          set_dd(tag & 3u, v);
        } else if (v <= worst) {  // mindistsq * epsError <= worstDist()
          if (tag & kRefLeaf) {
            cur = tag;
            mode = kTmLeaf;
          } else {
            cur = tag & kIdxMask;
            pdim = (tag >> 29) & 3u;
            pv = v;
            mode = kTmEnter;
          }
          found = true;
          continue;
        }
        // Nothing below can be visited any more: nanoflann prunes all of it.
        if ((double)pb > worst) stop = true;
      }
      st.drop(used);
      if (found) return;
      if (stop) break;
    }
    st.reset();
    mode = kTmDone;
  }

  // One leaf (mode kTmLeaf) then the stack (kTmPop).
  __device__ __forceinline__ void leaf_step(const DevCloud &C, KStack &st) {
    if (mode == kTmLeaf) {
      // searchLevel's leaf loop (include/nanoflann.hpp:1349-1366).  Testing
      // against the current worst distance instead of the one read at the
      // leaf's start inserts the same points: addPoint ignores a distance
      // that is not below the worst once the set is full.  Generators are
      // loaded kLeafChunk at a time; loads past the leaf's end stay inside
      // the padding records after the last slot.
      const uint32_t s0 = cur & kSlotMask;
      const uint32_t end = (uint32_t)gld(&C.gen[s0].end);
      uint32_t t0 = s0;
      do {
        double2 xy[kLeafChunk];
        double z[kLeafChunk];
        const Gen *g = C.gen + t0;  // one address; the chunk's records at fixed offsets
#pragma unroll
        for (int j = 0; j < kLeafChunk; ++j) {
          xy[j] = gld(reinterpret_cast<const double2 *>(g + j));
          z[j] = gld(&g[j].z);
        }
#pragma unroll
        for (int j = 0; j < kLeafChunk; ++j) {
          const double e0 = q[0] - xy[j].x;
          const double e1 = q[1] - xy[j].y;
          const double e2 = q[2] - z[j];
          double dist = e0 * e0;
          dist = dist + e1 * e1;
          dist = dist + e2 * e2;
          if (t0 + j < end && dist < rs.worst()) rs.add(dist, t0 + j);
        }
        t0 += kLeafChunk;
      } while (t0 < end);
      mode = kTmPop;
      if (depth >= 0) {
        // End of the first descent: the ring holds the far children of the
        // deepest kLdsDepth levels.
        const double w0 = rs.worst();
        if (depth <= kLdsDepth || !(lost_min <= w0)) {
          st.sp = depth < kLdsDepth ? depth : kLdsDepth;
          st.base = depth > kLdsDepth ? (depth & (kLdsDepth - 1)) : 0;
          st.maxsp = st.sp > st.maxsp ? st.sp : st.maxsp;
        } else {
          // An entry that fell out of the ring can still be visited: walk
          // the same path again, pushing the far children that can.
          st.reset();
          ptop = INFINITY;
          cur = C.rroot;
          mode = kTmNode;
        }
        depth = -1;
      }
    }
    if (mode == kTmPop) pop(st);
  }

  // Called when a search has finished.  Returns true with the query's result
  // in ids / r2 (entries >= k are kNone / DBL_MAX), or false after starting
  // the next image search.
  __device__ __forceinline__ bool finish(const DevCloud &C, KStack &st, uint32_t (&ids)[3], double (&r2)[3]) {
    const int k = rs.k;
    if constexpr (kPer) {
      if (img < 0) {
        ImageBox ib;
        ib.init(C, qb);
        if (ib.any_within(rs.worst() + kTol)) {
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            mi[j] = rs.id[j];
            md[j] = rs.d[j];
          }
          m = k;
          img = 0;
        }
      } else {
        // Merge: de-duplicate by id keeping the smaller r^2, stable sort by
        // r^2 (std::sort on at most 16 elements is libstdc++'s insertion
        // sort), truncate to k.  Between images at most k <= 3 entries
        // survive; up to 2k are live during the merge.
        uint32_t xi[6];
        double xd[6];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          xi[j] = mi[j];
          xd[j] = md[j];
          xi[j + 3] = kNone;
          xd[j + 3] = DBL_MAX;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (i < k && rs.id[i] != kNone) {
            bool found = false;
#pragma unroll
            for (int j = 0; j < 6; ++j) {
              if (!found && j < m && xi[j] == rs.id[i]) {
                xd[j] = (rs.d[i] < xd[j]) ? rs.d[i] : xd[j];
                found = true;
              }
            }
            if (!found) {
#pragma unroll
              for (int j = 0; j < 6; ++j)
                if (j == m) {
                  xi[j] = rs.id[i];
                  xd[j] = rs.d[i];
                }
              ++m;
            }
          }
        }
#pragma unroll
        for (int i = 1; i < 6; ++i) {
#pragma unroll
          for (int j = i; j > 0; --j) {
            if (i < m && xd[j] < xd[j - 1]) {
              const double td = xd[j];
              xd[j] = xd[j - 1];
              xd[j - 1] = td;
              const uint32_t ti = xi[j];
              xi[j] = xi[j - 1];
              xi[j - 1] = ti;
            }
          }
        }
        if (m > k) m = k;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          mi[j] = xi[j];
          md[j] = xd[j];
        }
        ++img;
      }
      if (img >= 0) {
        // Images img..26 in loop order, skipping at once those with a shifted
        // coordinate whose own term already exceeds wk + kTol (their
        // minDistSqToBox does too); the rest get the reference's exact test,
        // minDistSqToBox(shifted) > wk + kTol, as ib.d2, which adds the same
        // per-dimension terms in the same order.
        ImageBox ib;
        ib.init(C, qb);
        uint32_t cand;
        {
          const double w = sel3(md[0], md[1], md[2], (uint32_t)(k - 1)) + kTol;
          uint32_t ms[3];
#pragma unroll
          for (int d = 0; d < 3; ++d)
            ms[d] = 2u | (ib.lo[d] <= w ? 1u : 0u) | (ib.hi[d] <= w ? 4u : 0u);
          cand = 0u;
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
              if (((ms[0] >> a) & (ms[1] >> b) & 1u) != 0u) cand |= ms[2] << (a * 9 + b * 3);
          cand &= ~(1u << 13) & ~((1u << img) - 1u);
        }
        while (cand) {
          img = __ffs(cand) - 1;
          cand &= cand - 1u;
          const int ix = img / 9 - 1, iy = (img / 3) % 3 - 1, iz = img % 3 - 1;
          const double wk = sel3(md[0], md[1], md[2], (uint32_t)(k - 1));
          if (ib.d2(ix, iy, iz) > wk + kTol) continue;
          const double sh[3] = {qb[0] + ix * C.box[0], qb[1] + iy * C.box[1],
                                qb[2] + iz * C.box[2]};
          // An image's neighbours at r^2 >= wk cannot change the merged k
          // (a duplicate keeps its smaller r^2; a newcomer sorts after the
          // k entries <= wk and is truncated), so the image search looks
          // only below wk: the same merge, with most of the tree pruned.
          begin_search(C, st, sh[0], sh[1], sh[2], wk);
          return false;
        }
        img = 27;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          ids[j] = j < k ? mi[j] : kNone;
          r2[j] = j < k ? md[j] : DBL_MAX;
        }
        return true;
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ids[j] = j < k ? rs.id[j] : kNone;
      r2[j] = j < k ? rs.d[j] : DBL_MAX;
    }
    return true;
  }
};

// Lane phases.
enum : int {
  kPhFetch = 0,  // needs a ray
  kPhWalk = 1,   // a search has finished: walk step due
  kPhSearch = 2,
  kPhDone = 3,   // nothing to do (no rays left): free to walk another lane's interval
  kPhReplay = 4,  // own list exhausted, waiting to replay intervals given away
  kPhHelped = 5,  // a given interval is walked, waiting until its giver replayed it
};

// The 8x8-tile ray order: with the rays the pixels of row-major images whose
// rows hold w rays, fetch slot j is taken as the pixel (8b + r, 8t + c) of band
// b (8 rows), column tile t (8 columns, fewer at the right edge) and row r of
// the tile, so that the 64 rays a wave holds at once are neighbours on the
// image and walk through the same part of the tree.  A bijection of [0, nfull)
// (nfull = the rays of whole bands, at most nrays); slots from nfull on keep
// their index.  Only the order rays are taken in changes, never a result.
__device__ __forceinline__ uint32_t tile_ray(uint32_t j, uint32_t w, uint32_t nfull) {
  if (j >= nfull) return j;
  const uint32_t band = 8u * w;
  const uint32_t b = j / band;
  const uint32_t q = j - b * band;
  const uint32_t t = q >> 6;
  const uint32_t cw = min(8u, w - 8u * t);
  const uint32_t u = q - (t << 6);
  const uint32_t r = u / cw;
  return (8u * b + r) * w + 8u * t + (u - r * cw);
}

struct EngineCounters {
  uint64_t nseg = 0, nq = 0, nslow = 0;
  int maxw = 0;
};

// ---- walking another lane's intervals (lanes of one wave) -------------------
// A lane with nothing left to fetch takes the bottom pending pair of a busy
// lane of its wave (Walker::steal_bottom) and walks that interval; it records
// its case-2 boundaries and zero-width switches as events, which the giver
// replays, most recent gift first, once its own list is exhausted, before it
// closes the ray (or, for a lane that is itself walking a gift, before it
// reports its own events).  Replayed in order, the events give the giver's
// sink exactly the segments of the sequential walk.  Lanes exchange only
// registers (wave shuffles) and their own device buffers; events are read
// through the device-coherent path.
constexpr uint32_t kEvBoundary = 1u, kEvSwitch = 2u;
// Status a lane walking a gift publishes: state (3 bits), events (12 bits),
// walk iterations (17 bits, saturated).
enum : uint32_t { kHsIdle = 0, kHsWork = 1, kHsDone = 2, kHsOver = 3, kHsDefer = 4 };
constexpr uint32_t kHsMaxEvents = 4095u, kHsMaxIter = 131071u;
__device__ __forceinline__ uint32_t hs_pack(uint32_t st, uint32_t nev, int it) {
  const uint32_t i = it < 0 ? 0u : ((uint32_t)it > kHsMaxIter ? kHsMaxIter : (uint32_t)it);
  return st | (min(nev, kHsMaxEvents) << 3) | (i << 15);
}
// Rounds a lane waits for another before it gives up (a guard against a
// protocol error, which would otherwise spin forever).
constexpr int kMaxWaitRounds = 1 << 24;

// The sink the walk writes to: the IO's own sink for a lane walking its own
// ray, the event buffer for a lane walking a gift.
template <class Inner>
struct LaneSink {
  Inner in;
  uint4 *ev;  // this lane's event buffer, entries strided by `stride`
  int stride, cap;
  uint32_t nev;
  bool helper, over;
  __device__ __forceinline__ void seg(uint32_t cell, double a, double b) { in.seg(cell, a, b); }
  __device__ __forceinline__ void put(uint32_t type, double s, uint32_t cell) {
    if ((int)nev < cap) {
      const uint64_t b = (uint64_t)__double_as_longlong(s);
      gst(ev + (size_t)nev * stride, make_uint4(cell, type, (uint32_t)b, (uint32_t)(b >> 32)));
      ++nev;
    } else {
      over = true;
    }
  }
  // A case-2 boundary at s closing `open` (entered at `entry`) and opening `next`.
  __device__ __forceinline__ void close(uint32_t open, double entry, double s, uint32_t next) {
    if (helper)
      put(kEvBoundary, s, next);
    else
      in.seg(open, entry, s);
  }
  // A zero-width step: the open cell becomes `next` without a boundary.
  __device__ __forceinline__ void sw(uint32_t next) {
    if (helper) put(kEvSwitch, 0.0, next);
  }
};

__device__ __forceinline__ int nth_set_bit(uint64_t m, int n) {
  for (int i = 0; i < n; ++i) m &= m - 1;
  return __ffsll((unsigned long long)m) - 1;
}

// Lane roles.
enum : int { kRoleIdle = 0, kRoleOwner = 1, kRoleHelper = 2 };

// io.setup(sink), when the IO has one: state every lane's sink holds for the
// whole launch, set once before the first ray (the same values in every
// lane, so the compiler keeps them in scalar registers).
template <class IO, class Sink>
__device__ __forceinline__ auto io_setup(const IO &io, Sink &sk, int)
    -> decltype(io.setup(sk), void()) {
  io.setup(sk);
}
template <class IO, class Sink>
__device__ __forceinline__ void io_setup(const IO &, Sink &, long) {}

// io.finish(sink), when the IO has one: run once by every lane after its
// last ray of the launch.
template <class IO, class Sink>
__device__ __forceinline__ auto io_finish(const IO &io, Sink &sk, int)
    -> decltype(io.finish(sk), void()) {
  io.finish(sk);
}
template <class IO, class Sink>
__device__ __forceinline__ void io_finish(const IO &, Sink &, long) {}

// Runs rays through the walk, handing segments to the sinks IO provides:
// IO::Sink, io.begin(ray, sink), io.end(ray, sink) -> error code (kErrNone or
// a failure of the ray's output).  The rays are [0, nrays), or ray_list[0,
// nrays) when ray_list is given.  Without kExpand, a ray that needs the
// K_EXPAND classification is abandoned and appended to `defer` (its outputs
// are rewritten by the launch that runs the deferred rays with kExpand).
// With kHelp (and help buffers in S), idle lanes walk intervals of busy lanes
// of their wave; without it the helping code is compiled out.
template <bool kPer, bool kExpand, bool kHelp, class IO>
__device__ __forceinline__ void run_engine(const DevCloud &C, const Scratch &S,
                                           const double *start, const double *end,
                                           int64_t nrays, const uint32_t *ray_list,
                                           uint32_t *defer, ErrRec *err, Stats *stats, IO &io,
                                           KStack &st, EngineCounters &cnt) {
  const int gtid = blockIdx.x * kBlock + threadIdx.x;
  const size_t wbase = (size_t)gtid;
  const int lane = (int)(threadIdx.x & 63);
  const size_t wave0 = wbase - (size_t)lane;  // first lane of this wave
  Walker<kPer> W;
  Search<kPer> Q;
  LaneSink<typename IO::Sink> sk;
  io_setup(io, sk.in, 0);
  sk.ev = S.help_ev ? S.help_ev + wbase : nullptr;
  sk.stride = S.lanes;
  sk.cap = S.help_cap;
  sk.nev = 0;
  sk.helper = false;
  sk.over = false;
  const bool can_help = kHelp && S.help_ev != nullptr;
  int role = kRoleIdle;
  int boss = 0, my_slot = 0;  // a gift's giver (lane) and its place in the giver's list
  uint32_t boss_epoch = 0;
  uint32_t epoch = 0;     // changes whenever this lane's work ends or is abandoned
  int nsteal = 0;         // gifts of the current work not replayed yet
  int extra_it = 0;       // walk iterations of replayed gifts
  uint32_t hstat = kHsIdle;
  int waited = 0;
  int ph = kPhFetch;
  int64_t ray = 0;
  const int batch = S.walk_batch;
  const int leaf_batch = S.leaf_batch;
  const int nparts = S.parts;
  const int64_t psize = (nrays + nparts - 1) / nparts;
  int part = blockIdx.x & (nparts - 1);
  int part_tries = 0;
  // A caller's ray order (a permutation of [0, nrays), set only for launches
  // of that many rays) stands in for the list of a launch that has none.
  if (!ray_list) ray_list = S.ray_order;
  const uint32_t tile_w = (ray_list || nrays > 0xFFFFFFFFll) ? 0u : S.tile_row;
  const uint32_t tile_n =
      tile_w ? (uint32_t)((uint64_t)nrays / (8ull * tile_w) * (8ull * tile_w)) : 0u;
  for (;;) {
    {  // walk steps of every waiting lane
      WalkErr e{{0.0, 0.0, 0.0}, {0u, 0u}};
      int rc = kErrNone;
      bool query = false;
      bool ended = false;  // the lane's list is exhausted (its own part of the work)
      double q[3];
      int k = 1;
      double bound = DBL_MAX;
      bool published = false;  // this lane reported the end of a gift's walk
      if (ph == kPhWalk) {
        uint32_t ids[3];
        double r2[3];
        if (Q.finish(C, st, ids, r2)) {
          ++cnt.nq;
          rc = W.template consume<kExpand>(C, S, wbase, (int)threadIdx.x, gtid, ids, r2, sk,
                                           cnt.maxw, cnt.nslow, e);
          if (!rc && W.state < kStFinal) rc = W.prepare(C, S, wbase, end + 3 * ray, q, k, bound, e, sk);
          if (!rc && W.state == kStFinal) ended = true;
          // A gift whose events no longer fit goes back to its giver whole:
          // stop walking it (and drop this lane's own gifts of it).
          if (!rc && role == kRoleHelper && sk.over) {
            ended = true;
            nsteal = 0;
          }
          if (!rc && !ended && W.state != kStDone) query = true;
        } else {
          ph = kPhSearch;  // the next image search has started
        }
      }
      // A lane waiting for its most recent gift: replay it once it is walked.
      if constexpr (kHelp) {
        uint4 rec = make_uint4(0u, 0u, 0u, 0u);
        if (ph == kPhReplay) rec = gld(S.help_steal + (size_t)(2 * nsteal - 2) * S.lanes + wbase);
        const uint32_t hs = __shfl(hstat, (int)rec.x);
        if (ph == kPhReplay) {
          const uint32_t hst = hs & 7u;
          if (hst == kHsDone) {
            const uint32_t nev = (hs >> 3) & kHsMaxEvents;
            const uint4 *ev = S.help_ev + wave0 + rec.x;
            for (uint32_t i = 0; i < nev; ++i) {
              const uint64_t *p = reinterpret_cast<const uint64_t *>(ev + (size_t)i * S.lanes);
              const uint64_t a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              const uint64_t b =
                  __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              const uint32_t cell = (uint32_t)a;
              const double s = __longlong_as_double((long long)b);
              if ((uint32_t)(a >> 32) == kEvBoundary) {
                sk.close(W.open_cell, W.entry_s, s, cell);
                W.entry_s = s;
              } else {
                sk.sw(cell);
              }
              W.open_cell = cell;
            }
            extra_it += (int)(hs >> 15);
            --nsteal;
            waited = 0;
            if (W.iteration + extra_it > kMaxIter) {
              rc = kErrMaxIter;
            } else if (nsteal == 0) {
              ended = true;
            }
          } else if (hst == kHsOver) {
            // The gift did not fit its walker's event buffer: walk it here.
            const uint4 sv = gld(S.help_steal + (size_t)(2 * nsteal - 1) * S.lanes + wbase);
            --nsteal;
            waited = 0;
            atomicAdd(&stats->intervals_back, 1ULL);
            W.take_back(rec.y, __longlong_as_double((long long)(((uint64_t)sv.y << 32) | sv.x)),
                        rec.z, __longlong_as_double((long long)(((uint64_t)sv.w << 32) | sv.z)));
            rc = W.prepare(C, S, wbase, end + 3 * ray, q, k, bound, e, sk);
            if (!rc && W.state == kStFinal) ended = true;
            if (!rc && !ended) query = true;
          } else if (hst == kHsDefer) {
            --nsteal;
            rc = kDeferRay;
          } else if (++waited > kMaxWaitRounds) {
            rc = kErrHelpWait;
          }
        }
      }
      if (ended && nsteal > 0) {
        ph = kPhReplay;
      } else if (ended || (!rc && ph == kPhWalk && W.state == kStDone)) {
        // This lane's work is complete.
        if (role == kRoleHelper) {
          hstat = hs_pack(sk.over ? kHsOver : kHsDone, sk.nev, W.iteration + extra_it);
          published = true;
          ph = kPhHelped;
          waited = 0;
        } else {
          if (W.state == kStFinal) {
            sk.seg(W.open_cell, W.entry_s, W.r.len);
            W.state = kStDone;
          }
          rc = st.overflow ? kErrKnnStack : io.end(ray, sk.in);
          if (rc) report_error(err, rc, ray, e.v[0], e.v[1], e.v[2], e.id[0], e.id[1]);
          cnt.nseg += sk.in.n;
          ph = kPhFetch;
        }
        ++epoch;
      } else if (rc) {
        // Deferred to the K_EXPAND launch, or failed: the work is abandoned
        // (its gifts are released by the epoch change).
        if (!kExpand && rc == kDeferRay) {
          if (role == kRoleHelper) {
            hstat = hs_pack(kHsDefer, 0u, 0);
          } else {
            const unsigned long long slot = atomicAdd(&stats->deferred, 1ULL);
            gst(defer + slot, (uint32_t)ray);
          }
        } else {
          report_error(err, rc, ray, e.v[0], e.v[1], e.v[2], e.id[0], e.id[1]);
          if (role == kRoleHelper) hstat = hs_pack(kHsDone, 0u, 0);
        }
        if (role == kRoleHelper) {
          published = true;
          ph = kPhHelped;
          waited = 0;
        } else {
          cnt.nseg += sk.in.n;
          ph = kPhFetch;
        }
        nsteal = 0;
        ++epoch;
      }
      // A lane walking a gift: released once its giver has replayed it, or
      // when the giver abandoned its work.
      if constexpr (kHelp) {
        const uint32_t bep = __shfl(epoch, boss);
        const int bns = __shfl(nsteal, boss);
        if (role == kRoleHelper) {
          const bool gone = bep != boss_epoch;
          if ((ph == kPhHelped && (gone || bns <= my_slot)) ||
              (gone && (ph == kPhSearch || ph == kPhWalk || ph == kPhReplay)) ||
              (ph == kPhHelped && ++waited > kMaxWaitRounds)) {
            if (ph == kPhHelped && waited > kMaxWaitRounds)
              report_error(err, kErrHelpWait, ray, 0.0, 0.0, 0.0, 0u, 0u);
            if (ph != kPhHelped) {
              ++epoch;  // abandoned mid-walk: release this lane's own gifts
              nsteal = 0;
              query = false;
            }
            role = kRoleIdle;
            hstat = kHsIdle;
            ph = kPhDone;
          }
        }
      }
      if (ph == kPhFetch) {
        // The lane's own partition first, then the others in turn.
        ray = -1;
        if ((int)(threadIdx.x & 63) >= S.fetch_lanes) part_tries = nparts;
        while (part_tries < nparts) {
          const int64_t lo = part * psize;
          const int64_t idx = (int64_t)atomicAdd(&stats->part_counter[part], 1ULL);
          if (idx < psize && lo + idx < nrays) {
            ray = lo + idx;
            break;
          }
          part = (part + 1) & (nparts - 1);
          ++part_tries;
        }
        if (ray < 0 ||
            __hip_atomic_load(&err->flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ph = kPhDone;
          role = kRoleIdle;
        } else {
          if (ray_list) ray = (int64_t)gld(ray_list + ray);
          else if (tile_n) ray = (int64_t)tile_ray((uint32_t)ray, tile_w, tile_n);
          role = kRoleOwner;
          nsteal = 0;
          extra_it = 0;
          sk.helper = false;
          sk.over = false;
          sk.nev = 0;
          W.begin(start + 3 * ray, end + 3 * ray);
          io.begin(ray, sk.in);
          W.prepare(C, S, wbase, end + 3 * ray, q, k, bound, e, sk);  // kStStart: the ray's start
          query = true;
        }
      }
      // Lanes with nothing to do take the bottom pending pair of busy lanes of
      // their wave: the k-th idle lane the k-th lane that can give one.
      const uint64_t idle = can_help ? __ballot(ph == kPhDone && !query) : 0ull;
      if (kHelp && idle) {
        // A lane that is searching, or has just prepared its next search,
        // can give: a search reads only the top pair of the pending list.
        const bool can = (ph == kPhSearch || query) && role != kRoleIdle &&
                         W.can_give(S.help_depth) &&
                         nsteal < S.help_steal_cap && W.wsp + 4 < S.wcap;
        const uint64_t donors = __ballot(can);
        if (donors) {
          const uint64_t below = (1ull << lane) - 1ull;
          const int nidle = __popcll(idle), ndon = __popcll(donors);
          const bool take = ((idle >> lane) & 1ull) && __popcll(idle & below) < ndon;
          const bool give = ((donors >> lane) & 1ull) && __popcll(donors & below) < nidle;
          int src = lane;
          uint32_t pc = 0, qc = 0;
          double ps = 0.0, qs = 0.0;
          int slot = 0;
          if (take) src = nth_set_bit(donors, __popcll(idle & below));
          if (give) {
            const int to = nth_set_bit(idle, __popcll(donors & below));
            W.steal_bottom(S, wbase, pc, ps, qc, qs);
            const uint64_t pb = (uint64_t)__double_as_longlong(ps);
            const uint64_t qb = (uint64_t)__double_as_longlong(qs);
            uint4 *rec = S.help_steal + (size_t)(2 * nsteal) * S.lanes + wbase;
            gst(rec, make_uint4((uint32_t)to, pc, qc, 0u));
            gst(rec + S.lanes, make_uint4((uint32_t)pb, (uint32_t)(pb >> 32), (uint32_t)qb,
                                          (uint32_t)(qb >> 32)));
            slot = nsteal++;
            atomicAdd(&stats->intervals, 1ULL);
          }
          const uint32_t t_pc = __shfl(pc, src), t_qc = __shfl(qc, src);
          const double t_ps = __shfl(ps, src), t_qs = __shfl(qs, src);
          const uint32_t t_ray = __shfl((uint32_t)ray, src), t_ep = __shfl(epoch, src);
          const int t_slot = __shfl(slot, src);
          if (take) {
            role = kRoleHelper;
            boss = src;
            boss_epoch = t_ep;
            my_slot = t_slot;
            ray = (int64_t)t_ray;
            ++epoch;
            nsteal = 0;
            extra_it = 0;
            waited = 0;
            sk.helper = true;
            sk.over = false;
            sk.nev = 0;
            W.begin_interval(start + 3 * ray, end + 3 * ray, t_pc, t_ps, t_qc, t_qs);
            hstat = kHsWork;
            const int rc2 = W.prepare(C, S, wbase, end + 3 * ray, q, k, bound, e, sk);
            if (rc2) report_error(err, rc2, ray, e.v[0], e.v[1], e.v[2], e.id[0], e.id[1]);
            if (rc2 || W.state == kStFinal) {  // nothing to walk (zero-width) or failed
              hstat = hs_pack(sk.over ? kHsOver : kHsDone, sk.nev, W.iteration);
              published = true;
              ph = kPhHelped;
            } else {
              query = true;
            }
          }
        }
      }
      // A gift's events are in device memory before its giver can see its
      // status (read through a shuffle in a later iteration).
      if (kHelp && __ballot(published)) __threadfence();
      if (query) {
        Q.start(C, st, q, k, bound);
        ph = kPhSearch;
      }
    }
    if (!__ballot(ph != kPhDone)) break;
    // Search steps only, until `batch` lanes wait for a walk step or no lane
    // is searching.  Node steps run every iteration; leaf steps once
    // `leaf_batch` lanes (or every searching lane) are at a leaf.
    for (;;) {
      const bool at_leaf = ph == kPhSearch && Q.mode >= kTmLeaf;
      const uint64_t nodes = __ballot(ph == kPhSearch && Q.mode <= kTmEnter);
      const uint64_t leaves = __ballot(at_leaf);
      if (!(nodes | leaves)) break;
      if (__popcll(__ballot(ph == kPhWalk || ph == kPhFetch)) >= batch) break;
      if (leaves && (!nodes || __popcll(leaves) >= leaf_batch) && at_leaf) {
        Q.leaf_step(C, st);
        if (Q.mode == kTmDone) ph = kPhWalk;
      }
      if (ph == kPhSearch && Q.mode <= kTmEnter) Q.node_step(C, st);
    }
  }
  io_finish(io, sk.in, 0);
}

}  // namespace vt
