// Layout of a built point cloud: the host tree (nanoflann's node list) and
// the device tree the HIP kernels search.
//
// The kd-tree is nanoflann 1.3.2's static index with leaf size 10 (reference
// src/pointcloud.cpp:204, include/nanoflann.hpp:755-767).  The host keeps it
// as a pre-order array of nodes, leaves included (Node).  The device copy
// keeps only the inner nodes, in pre-order, as 32-byte records (DNode) whose
// child references name either another inner node or a leaf's run of slots,
// so a search reads a leaf's generators directly from its parent.
// Generators are stored in the tree's leaf order ("slots": slot t holds the
// point vind[t]), so a leaf is one contiguous run of 32-byte records, and the
// cell identity used inside the kernels is the slot.  Fields follow the same
// slot order.  slot -> filtered index is kept for output.
#pragma once
#include <cstdint>

namespace vt {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kLeafMaxSize = 10;  // src/pointcloud.cpp:204

// Child references of device nodes (and kNN stack tags):
//   inner node i:  i                      (bit 31 clear, i < kMaxNodes)
//   leaf:          kRefLeaf | first slot  (first slot < kMaxSlots)
// In a DNode's c1 / c2 fields bit 30 carries bit 0 / bit 1 of the node's cut
// dimension.  Stack tags use bits 29-30 of an inner reference for the parent's
// cut dimension and kRefLeaf | kRefUndo | dim for undo entries.
constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefUndo = 0x40000000u;
constexpr uint32_t kDimBit = 0x40000000u;
constexpr uint32_t kMaxNodes = 1u << 29;
constexpr uint32_t kMaxSlots = 1u << 30;
constexpr uint32_t kIdxMask = kMaxNodes - 1;
constexpr uint32_t kSlotMask = kMaxSlots - 1;

// Host node (pre-order, leaves included), as nanoflann's Node.
struct alignas(16) Node {
  double lo;   // inner: divlow  (tight max of child1 along the cut dim)
  double hi;   // inner: divhigh (tight min of child2 along the cut dim)
  uint32_t a;  // inner: index of child2;            leaf: first slot
  uint32_t b;  // inner: cut dimension (0, 1 or 2);  leaf: end slot | kLeafBit
  double pb;   // the parent's bound facing this node along the parent's cut
               // dimension: parent divlow for child1, parent divhigh for child2
               // (0 for the root).  A search entering this node as its parent's
               // far child computes cut_dist = (q[dim] - pb)^2 from it.
};
static_assert(sizeof(Node) == 32, "node must be 32 bytes");

// Device inner node.  The first 16 bytes decide the descent; the second 16
// bytes hold the children and the bound a search entering this node as its
// parent's far child needs.
struct alignas(16) DNode {
  double lo;    // divlow  (tight max of child1 along the cut dimension)
  double hi;    // divhigh (tight min of child2 along the cut dimension)
  uint32_t c1;  // child1 reference | (dim & 1) << 30
  uint32_t c2;  // child2 reference | (dim >> 1) << 30
  double pb;    // the parent's bound facing this node along the parent's cut
                // dimension: parent divlow for child1, parent divhigh for
                // child2; cut_dist = (q[parent dim] - pb)^2 on entering it
};
static_assert(sizeof(DNode) == 32, "device node must be 32 bytes");

// Pair record of the ray engine's searches: an inner node X at even depth
// (the root has depth 0) together with the split data of its two children, so
// one 80-byte read serves two levels of a descent.  Inner nodes at odd depth
// have no record of their own: they are read from their parent's.
// References in pair records and in the engine's stack tags:
//   even inner node:  record index r                         (r < 2^27)
//   odd inner node:   kRefOdd | r << 1 | b, child b (0: child1, 1: child2)
//                     of the node of record r
//   leaf:             kRefLeaf | first slot
// with bit 30 of the c* fields carrying cut-dimension bits as in DNode.
constexpr uint32_t kRefOdd = 1u << 28;
constexpr uint32_t kMaxRecs = 1u << 27;
struct alignas(16) DRec {
  double lo, hi;     // X: divlow, divhigh
  uint32_t c1, c2;   // X: child references | (dim & 1) << 30, | (dim >> 1) << 30
  double pb;         // X: as DNode::pb (entering X as its parent's far child)
  double lo1, hi1;   // child1 of X when inner: divlow, divhigh
  double lo2, hi2;   // child2 of X when inner
  uint32_t c11, c12, c21, c22;  // children of child1 / child2 (with dim bits)
};
static_assert(sizeof(DRec) == 80, "pair record must be 80 bytes");

struct alignas(16) Gen {  // one generator, slot order
  double x, y, z;
  uint32_t start, end;    // slot range of the generator's leaf (kNone-like
                          // 0xFFFFFFFF in the padding records after the last slot)
};
static_assert(sizeof(Gen) == 32, "generator record must be 32 bytes");

}  // namespace vt
