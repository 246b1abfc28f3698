// Host builder of the device kd-tree.
//
// The decisions are nanoflann 1.3.2's (include/nanoflann.hpp):
//   root box      computeBoundingBox            1316-1339
//   recursion     divideTree                    858-907
//   cut choice    middleSplit_                  910-959
//   partition     planeSplit (two Hoare passes) 967-1003
// so slot order, cut dimensions and divlow/divhigh are bit-identical to the
// reference's index.  The implementation differs: coordinates are kept in
// slot order (structure of arrays, permuted together with vind) so every
// min/max scan and partition pass streams contiguous memory, independent
// subtrees are built on separate host threads, and nodes are emitted directly
// in the device's pre-order layout.
#include "tree_build.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <utility>

namespace vt {
namespace {

struct Box {
  double lo[3], hi[3];
};

struct Ctx {
  double *c[3];    // coordinates in slot order
  uint32_t *vind;  // slot -> filtered index
};

struct SubTree {
  std::vector<Node> nodes;  // local pre-order, child2 indices local
  int64_t leaves = 0, max_depth = 0;
};

inline void range_min_max(const double *v, size_t l, size_t r, double &mn, double &mx) {
  mn = v[l];
  mx = v[l];
  for (size_t i = l + 1; i < r; ++i) {
    double x = v[i];
    if (x < mn) mn = x;
    if (x > mx) mx = x;
  }
}

inline void swap_slots(const Ctx &x, size_t i, size_t j) {
  std::swap(x.c[0][i], x.c[0][j]);
  std::swap(x.c[1][i], x.c[1][j]);
  std::swap(x.c[2][i], x.c[2][j]);
  std::swap(x.vind[i], x.vind[j]);
}

// Hoare partition of slots [base, base+count): the first pass moves every
// coordinate < cut to the front, the second every coordinate == cut after it.
// lim1/lim2 are relative to base.  Same index arithmetic as planeSplit,
// including its unsigned "!right" guard.
void partition(const Ctx &x, size_t base, size_t count, int f, double cut, size_t &lim1,
               size_t &lim2) {
  const double *v = x.c[f] + base;
  size_t lo = 0, hi = count - 1;
  for (;;) {
    while (lo <= hi && v[lo] < cut) ++lo;
    while (hi && lo <= hi && v[hi] >= cut) --hi;
    if (lo > hi || !hi) break;
    swap_slots(x, base + lo, base + hi);
    ++lo;
    --hi;
  }
  lim1 = lo;
  hi = count - 1;
  for (;;) {
    while (lo <= hi && v[lo] <= cut) ++lo;
    while (hi && lo <= hi && v[hi] > cut) --hi;
    if (lo > hi || !hi) break;
    swap_slots(x, base + lo, base + hi);
    ++lo;
    --hi;
  }
  lim2 = lo;
}

// Chooses cut dimension and value for slots [l, r) inside the loose box and
// partitions them; returns the split offset.
size_t split_node(const Ctx &x, size_t l, size_t r, const Box &box, int &f, double &cut) {
  const size_t count = r - l;
  const double eps = 0.00001;
  double max_span = box.hi[0] - box.lo[0];
  for (int d = 1; d < 3; ++d) max_span = std::max(max_span, box.hi[d] - box.lo[d]);
  double best_spread = -1;
  f = 0;
  double mn[3], mx[3];
  bool have[3] = {false, false, false};
  for (int d = 0; d < 3; ++d) {
    if (box.hi[d] - box.lo[d] > (1 - eps) * max_span) {
      range_min_max(x.c[d], l, r, mn[d], mx[d]);
      have[d] = true;
      double spread = mx[d] - mn[d];
      if (spread > best_spread) {
        f = d;
        best_spread = spread;
      }
    }
  }
  const double mid = (box.lo[f] + box.hi[f]) / 2;
  if (!have[f]) range_min_max(x.c[f], l, r, mn[f], mx[f]);
  cut = mid < mn[f] ? mn[f] : (mid > mx[f] ? mx[f] : mid);
  size_t lim1, lim2;
  partition(x, l, count, f, cut, lim1, lim2);
  if (lim1 > count / 2) return lim1;
  if (lim2 < count / 2) return lim2;
  return count / 2;
}

void leaf_box(const Ctx &x, size_t l, size_t r, Box &box) {
  for (int d = 0; d < 3; ++d) range_min_max(x.c[d], l, r, box.lo[d], box.hi[d]);
}

// Builds slots [l, r) under a loose box; on return `box` is the subtree's
// tight box, as divideTree leaves it.  Nodes are appended in pre-order.
void build_rec(const Ctx &x, size_t l, size_t r, Box &box, int64_t depth, SubTree &out,
               int spawn_levels) {
  const size_t me = out.nodes.size();
  out.nodes.emplace_back();
  out.max_depth = std::max(out.max_depth, depth);
  if (r - l <= static_cast<size_t>(kLeafMaxSize)) {
    Node &n = out.nodes[me];
    n.lo = n.hi = 0.0;
    n.a = static_cast<uint32_t>(l);
    n.b = static_cast<uint32_t>(r) | kLeafBit;
    n.pb = 0.0;
    leaf_box(x, l, r, box);
    out.leaves++;
    return;
  }
  int f;
  double cut;
  const size_t idx = split_node(x, l, r, box, f, cut);
  Box lb = box, rb = box;
  lb.hi[f] = cut;
  rb.lo[f] = cut;
  uint32_t child2;
  if (spawn_levels > 0) {
    SubTree right;
    std::thread t([&] { build_rec(x, l + idx, r, rb, depth + 1, right, spawn_levels - 1); });
    build_rec(x, l, l + idx, lb, depth + 1, out, spawn_levels - 1);
    t.join();
    child2 = static_cast<uint32_t>(out.nodes.size());
    const uint32_t shift = child2;
    for (Node n : right.nodes) {
      if (!(n.b & kLeafBit)) n.a += shift;
      out.nodes.push_back(n);
    }
    out.leaves += right.leaves;
    out.max_depth = std::max(out.max_depth, right.max_depth);
  } else {
    build_rec(x, l, l + idx, lb, depth + 1, out, 0);
    child2 = static_cast<uint32_t>(out.nodes.size());
    build_rec(x, l + idx, r, rb, depth + 1, out, 0);
  }
  Node &n = out.nodes[me];
  n.lo = lb.hi[f];
  n.hi = rb.lo[f];
  n.a = child2;
  n.b = static_cast<uint32_t>(f);
  n.pb = 0.0;
  out.nodes[me + 1].pb = n.lo;
  out.nodes[child2].pb = n.hi;
  for (int d = 0; d < 3; ++d) {
    box.lo[d] = std::min(lb.lo[d], rb.lo[d]);
    box.hi[d] = std::max(lb.hi[d], rb.hi[d]);
  }
}

}  // namespace

void build_tree(const double *pts, size_t n, int threads, HostTree &out) {
  if (n == 0) throw std::runtime_error("There are no points in the cloud; cannot build tree");
  if (n >= static_cast<size_t>(kLeafBit))
    throw std::invalid_argument("too many points for 32-bit slot indices");
  std::vector<double> cx(n), cy(n), cz(n);
  out.vind.resize(n);
  for (size_t i = 0; i < n; ++i) {
    cx[i] = pts[3 * i];
    cy[i] = pts[3 * i + 1];
    cz[i] = pts[3 * i + 2];
    out.vind[i] = static_cast<uint32_t>(i);
  }
  Ctx x{{cx.data(), cy.data(), cz.data()}, out.vind.data()};
  Box box;
  for (int d = 0; d < 3; ++d) range_min_max(x.c[d], 0, n, box.lo[d], box.hi[d]);
  for (int d = 0; d < 3; ++d) {
    out.bbox[d][0] = box.lo[d];
    out.bbox[d][1] = box.hi[d];
  }
  int spawn = 0;
  while ((1 << spawn) < threads && spawn < 6) ++spawn;
  if (n < 200000) spawn = 0;
  SubTree st;
  st.nodes.reserve(n / 3 + 16);
  build_rec(x, 0, n, box, 0, st, spawn);
  out.nodes = std::move(st.nodes);
  out.leaves = st.leaves;
  out.max_depth = st.max_depth;
}

void make_device_tree(const HostTree &t, DeviceTree &out) {
  const size_t nh = t.nodes.size();
  const size_t n = t.vind.size();
  if (n >= kMaxSlots) throw std::invalid_argument("too many points for one cloud (2^30 slots)");
  std::vector<uint32_t> dev(nh, 0);  // host node -> device index (inner nodes)
  uint32_t ninner = 0;
  for (size_t h = 0; h < nh; ++h)
    if (!(t.nodes[h].b & kLeafBit)) dev[h] = ninner++;
  if (ninner >= kMaxNodes)
    throw std::invalid_argument("too many points for one cloud (2^29 inner nodes)");
  auto ref = [&](size_t h) -> uint32_t {
    const Node &c = t.nodes[h];
    return (c.b & kLeafBit) ? (kRefLeaf | c.a) : dev[h];
  };
  out.nodes.resize(ninner);
  out.leaf_start.assign(n, 0);
  out.leaf_end.assign(n, 0);
  for (size_t h = 0; h < nh; ++h) {
    const Node &nd = t.nodes[h];
    if (nd.b & kLeafBit) {
      const uint32_t end = nd.b & ~kLeafBit;
      for (uint32_t s = nd.a; s < end; ++s) {
        out.leaf_start[s] = nd.a;
        out.leaf_end[s] = end;
      }
      continue;
    }
    DNode &d = out.nodes[dev[h]];
    d.lo = nd.lo;
    d.hi = nd.hi;
    d.pb = nd.pb;
    d.c1 = ref(h + 1) | ((nd.b & 1u) ? kDimBit : 0u);
    d.c2 = ref(nd.a) | ((nd.b & 2u) ? kDimBit : 0u);
  }
  out.root = ref(0);
}

}  // namespace vt
