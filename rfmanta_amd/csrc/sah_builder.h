// Host-side DynBVH construction for the MI355X traversal library.
//
// Produces the reference DynBVH tree (Model/Groups/DynBVH.cc:1699-1857, 2277-2440):
// the same full-sweep SAH split choices, the same child order and node numbering, the
// same leaf id order, so a tree built here is node-for-node the tree DynBVH builds
// for the same mesh. The work is spread over host threads; results do not depend on
// the thread count.
#pragma once

#include <cstdint>
#include <vector>

namespace mh {

struct Aabb {
  float lo[3], hi[3];
};

// The reference BVHNode record (Model/Groups/DynBVH.h:37-77), 32 bytes.
struct RefNode {
  Aabb box;
  int32_t child;     // first child (inner) or first object_ids slot (leaf)
  uint8_t axis;      // split axis (inner), 0 (leaf)
  uint8_t bits;      // bit0 isLeftCheaper, bit1 isLargeSubtree
  int16_t children;  // 0 inner, >0 leaf primitive count
  bool leaf() const { return children != 0; }
};
static_assert(sizeof(RefNode) == 32, "reference BVHNode layout");

struct DynBvh {
  std::vector<RefNode> nodes;  // 2*T records, as DynBVH::allocate (DynBVH.cc:1680-1697)
  std::vector<int32_t> ids;    // object_ids
  int num_nodes = 0;           // records in use
};

// Full SAH build over an indexed triangle mesh. threads <= 0 uses all cores.
DynBvh build_dynbvh(const float* verts, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris,
                    int threads);

// isLeftCheaper for every inner node (DynBVH::computeTraversalCost, DynBVH.cc:1583-1678).
void compute_rtsah(std::vector<RefNode>& nodes, int threads);

// isLargeSubtree with largeSubtreeSize = nodes.size() / (20 * workers) (DynBVH.cc:1776,
// 1971-1996); workers is the reference's numProcs.
void mark_large_subtrees(std::vector<RefNode>& nodes, int workers = 1);

}  // namespace mh
