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
#include <string>
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

  // Pair records: depth parity, record index of every even inner node, and
  // the record-relative reference of every node.
  std::vector<uint8_t> odd(nh, 0);
  std::vector<uint32_t> rec(nh, 0), parent(nh, 0);
  uint32_t nrec = 0;
  for (size_t h = 0; h < nh; ++h) {
    const Node &nd = t.nodes[h];
    if (nd.b & kLeafBit) continue;
    odd[h + 1] = odd[nd.a] = (uint8_t)!odd[h];
    parent[h + 1] = parent[nd.a] = (uint32_t)h;
    if (!odd[h]) rec[h] = nrec++;
  }
  if (nrec >= kMaxRecs) throw std::invalid_argument("too many points for one cloud (2^27 pair records)");
  auto rref = [&](size_t h) -> uint32_t {
    const Node &c = t.nodes[h];
    if (c.b & kLeafBit) return kRefLeaf | c.a;
    if (!odd[h]) return rec[h];
    const size_t p = parent[h];
    return kRefOdd | rec[p] << 1 | (h == p + 1 ? 0u : 1u);
  };
  auto dims = [&](size_t h, uint32_t &a1, uint32_t &a2) {
    const Node &nd = t.nodes[h];
    a1 = rref(h + 1) | ((nd.b & 1u) ? kDimBit : 0u);
    a2 = rref(nd.a) | ((nd.b & 2u) ? kDimBit : 0u);
  };
  out.recs.assign(nrec, DRec{});
  for (size_t h = 0; h < nh; ++h) {
    const Node &nd = t.nodes[h];
    if ((nd.b & kLeafBit) || odd[h]) continue;
    DRec &r = out.recs[rec[h]];
    r.lo = nd.lo;
    r.hi = nd.hi;
    r.pb = nd.pb;
    dims(h, r.c1, r.c2);
    const size_t h1 = h + 1, h2 = nd.a;
    if (!(t.nodes[h1].b & kLeafBit)) {
      r.lo1 = t.nodes[h1].lo;
      r.hi1 = t.nodes[h1].hi;
      dims(h1, r.c11, r.c12);
    }
    if (!(t.nodes[h2].b & kLeafBit)) {
      r.lo2 = t.nodes[h2].lo;
      r.hi2 = t.nodes[h2].hi;
      dims(h2, r.c21, r.c22);
    }
  }
  out.rroot = nh ? rref(0) : 0u;
}

namespace {

template <class T>
void put(std::vector<uint8_t> &out, T v) {
  const size_t o = out.size();
  out.resize(o + sizeof(T));
  std::memcpy(out.data() + o, &v, sizeof(T));
}

struct Reader {
  const uint8_t *b;
  size_t size, p = 0;
  template <class T>
  T get() {
    if (size - p < sizeof(T)) throw std::invalid_argument("tree bytes: truncated");
    T v;
    std::memcpy(&v, b + p, sizeof(T));
    p += sizeof(T);
    return v;
  }
};

}  // namespace

void save_tree_bytes(const HostTree &t, std::vector<uint8_t> &out) {
  out.clear();
  const uint64_t n = t.vind.size();
  put<uint64_t>(out, n);
  put<int32_t>(out, 3);
  for (int d = 0; d < 3; ++d) {
    put<double>(out, t.bbox[d][0]);
    put<double>(out, t.bbox[d][1]);
  }
  put<uint64_t>(out, (uint64_t)kLeafMaxSize);
  put<uint64_t>(out, n);
  for (uint32_t v : t.vind) put<uint64_t>(out, v);
  for (const Node &nd : t.nodes) {
    if (nd.b & kLeafBit) {
      put<uint64_t>(out, nd.a);
      put<uint64_t>(out, nd.b & ~kLeafBit);
      put<uint64_t>(out, 0);
      put<uint64_t>(out, 0);
      put<uint64_t>(out, 0);
    } else {
      put<int32_t>(out, (int32_t)nd.b);
      put<int32_t>(out, 0);
      put<double>(out, nd.lo);
      put<double>(out, nd.hi);
      put<uint64_t>(out, 1);
      put<uint64_t>(out, 1);
    }
  }
}

void load_tree_bytes(const uint8_t *b, size_t size, size_t npart, HostTree &out) {
  Reader r{b, size};
  const uint64_t m_size = r.get<uint64_t>();
  const int32_t dim = r.get<int32_t>();
  double bbox[3][2];
  for (int d = 0; d < 3; ++d) {
    bbox[d][0] = r.get<double>();
    bbox[d][1] = r.get<double>();
  }
  const uint64_t leaf = r.get<uint64_t>();
  if (m_size != npart)
    throw std::invalid_argument("tree bytes: the tree indexes " + std::to_string(m_size) +
                                " points, the cloud has " + std::to_string(npart));
  if (dim != 3) throw std::invalid_argument("tree bytes: dimension is not 3");
  if (leaf != (uint64_t)kLeafMaxSize) throw std::invalid_argument("tree bytes: leaf size is not 10");
  const uint64_t nv = r.get<uint64_t>();
  if (nv != npart) throw std::invalid_argument("tree bytes: vind size differs from the point count");
  HostTree t;
  t.vind.resize(npart);
  std::vector<uint8_t> seen(npart, 0);
  for (size_t i = 0; i < npart; ++i) {
    const uint64_t v = r.get<uint64_t>();
    if (v >= npart || seen[v]) throw std::invalid_argument("tree bytes: vind is not a permutation");
    seen[v] = 1;
    t.vind[i] = (uint32_t)v;
  }
  std::memcpy(t.bbox, bbox, sizeof bbox);
  // Nodes in pre-order.  `open` holds the inner nodes whose child2 has not
  // started; a leaf ends a subtree, so the next node is child2 of the top.
  std::vector<size_t> open;
  std::vector<int64_t> depth_of;
  uint64_t next_slot = 0;
  int64_t depth = 0;
  for (;;) {
    const size_t i = t.nodes.size();
    if (i >= (size_t)kMaxNodes * 2) throw std::invalid_argument("tree bytes: too many nodes");
    uint8_t raw[40];
    for (int k = 0; k < 40; ++k) raw[k] = r.get<uint8_t>();
    uint64_t c1, c2;
    std::memcpy(&c1, raw + 24, 8);
    std::memcpy(&c2, raw + 32, 8);
    Node nd{};
    t.max_depth = std::max(t.max_depth, depth);
    depth_of.push_back(depth);
    if (c1 == 0 && c2 == 0) {
      uint64_t left, right;
      std::memcpy(&left, raw, 8);
      std::memcpy(&right, raw + 8, 8);
      if (left != next_slot || right <= left || right > npart ||
          right - left > (uint64_t)kLeafMaxSize)
        throw std::invalid_argument("tree bytes: leaves do not tile the points in order");
      next_slot = right;
      nd.a = (uint32_t)left;
      nd.b = (uint32_t)right | kLeafBit;
      t.nodes.push_back(nd);
      t.leaves++;
      if (open.empty()) break;
      const size_t j = open.back();
      open.pop_back();
      t.nodes[j].a = (uint32_t)(i + 1);
      depth = depth_of[j] + 1;
    } else {
      if (c1 == 0 || c2 == 0) throw std::invalid_argument("tree bytes: inner node with one child");
      int32_t feat;
      std::memcpy(&feat, raw, 4);
      if (feat < 0 || feat > 2) throw std::invalid_argument("tree bytes: cut dimension out of range");
      std::memcpy(&nd.lo, raw + 8, 8);
      std::memcpy(&nd.hi, raw + 16, 8);
      nd.b = (uint32_t)feat;
      t.nodes.push_back(nd);
      open.push_back(i);
      ++depth;
    }
  }
  if (next_slot != npart) throw std::invalid_argument("tree bytes: leaves do not cover every point");
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    Node &nd = t.nodes[i];
    if (nd.b & kLeafBit) continue;
    t.nodes[i + 1].pb = nd.lo;
    t.nodes[nd.a].pb = nd.hi;
  }
  out = std::move(t);
}

}  // namespace vt
