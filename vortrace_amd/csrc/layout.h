// Device layout of a built point cloud, shared by the host builder and the
// HIP kernels.
//
// The kd-tree is nanoflann 1.3.2's static index with leaf size 10 (reference
// src/pointcloud.cpp:204, include/nanoflann.hpp:755-767), flattened into a
// pre-order array of 32-byte nodes so that child1 of an inner node is the next
// node and only child2 is stored.  Generators are stored in the tree's leaf
// order ("slots": slot t holds the point vind[t]), so a leaf scan reads one
// contiguous run of 32-byte records and the cell identity used inside the
// kernels is the slot.  Fields follow the same slot order.  slot -> filtered
// index is kept for output.
#pragma once
#include <cstdint>

namespace vt {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kLeafMaxSize = 10;  // src/pointcloud.cpp:204

struct alignas(16) Node {
  double lo;   // inner: divlow  (tight max of child1 along the cut dim)
  double hi;   // inner: divhigh (tight min of child2 along the cut dim)
  uint32_t a;  // inner: index of child2;            leaf: first slot
  uint32_t b;  // inner: cut dimension (0, 1 or 2);  leaf: end slot | kLeafBit
  uint32_t depth;
  uint32_t pad;
};
static_assert(sizeof(Node) == 32, "node must be 32 bytes");

struct alignas(16) Gen {  // one generator, slot order
  double x, y, z, w;      // w: unused padding (keeps 16-byte loads aligned)
};
static_assert(sizeof(Gen) == 32, "generator record must be 32 bytes");

}  // namespace vt
