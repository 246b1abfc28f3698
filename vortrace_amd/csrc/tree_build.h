// Host builder of the device kd-tree (see layout.h).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "layout.h"

namespace vt {

struct HostTree {
  std::vector<Node> nodes;       // pre-order
  std::vector<uint32_t> vind;    // slot -> filtered index
  double bbox[3][2];             // root bounding box (nanoflann root_bbox)
  int64_t leaves = 0;
  int64_t max_depth = 0;
};

// Builds the tree over pts (n x 3, filtered order).  Every decision (cut
// dimension, cut value, partition and split index) follows nanoflann 1.3.2's
// buildIndex, so the slot order and every divlow/divhigh equal the
// reference's.  Uses up to `threads` host threads for independent subtrees.
void build_tree(const double *pts, size_t n, int threads, HostTree &out);

// The device tree: the inner nodes of t in pre-order as DNodes, the root's
// reference, and for every slot the slot range of its leaf.  Throws
// std::invalid_argument when the tree exceeds the reference encoding.
struct DeviceTree {
  std::vector<DNode> nodes;
  std::vector<DRec> recs;  // pair records of the even-depth inner nodes, pre-order
  std::vector<uint32_t> leaf_start, leaf_end;  // per slot: its leaf's slot range
  uint32_t root = 0;   // root reference among DNodes
  uint32_t rroot = 0;  // root reference among pair records
};
void make_device_tree(const HostTree &t, DeviceTree &out);

// nanoflann 1.3.2's index dump (saveIndex / loadIndex,
// include/nanoflann.hpp:1023-1071): m_size (u64), dim (i32), root_bbox (3 x
// {low, high} f64), leaf size (u64), vind (u64 count + u64 entries), then the
// nodes in pre-order as 40-byte Node structs whose child pointers only say
// null (leaf) or not.  save writes 1 for a non-null child and zeros for
// padding.  load checks that the dump describes a leaf-size-10 tree over
// exactly npart points (throws std::invalid_argument otherwise).
void save_tree_bytes(const HostTree &t, std::vector<uint8_t> &out);
void load_tree_bytes(const uint8_t *b, size_t size, size_t npart, HostTree &out);

}  // namespace vt
