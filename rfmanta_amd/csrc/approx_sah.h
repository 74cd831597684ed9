// The approximate SAH build of one subtree by one thread, shared by the rotating update
// (leaf splits inside rotateNode) and the device build (subtrees of few objects):
// partitionApproxSAH (Model/Groups/DynBVH.cc:2125-2275; LONGEST_AXIS is defined,
// DynBVH.cc:39), partition2Objs (DynBVH.cc:2277-2308), partitionObjects
// (DynBVH.cc:1369-1396) and DynBVH::build with useApproximateBuild (DynBVH.cc:1804-1857).
// Arithmetic is the reference's: Real is float, BBox::computeArea is float arithmetic
// returned as double, min/max are the select forms, no contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "mh_device.h"

namespace mh {
namespace {

constexpr float kCIsec = 1.f;  // BVH_C_isec (DynBVH.cc:29)
constexpr float kCTrav = 1.f;  // BVH_C_trav (DynBVH.cc:30)
constexpr int kSamples = 16;   // BVH_num_samples (DynBVH.cc:34)
constexpr uint8_t kPending = 3;  // axis value of a record whose subtree is still to build

struct Box {
  float lo[3], hi[3];
};

__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ void reset(Box& b) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = FLT_MAX;
    b.hi[k] = -FLT_MAX;
  }
}
__device__ __forceinline__ void extend(Box& b, const Box& o) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = mn(b.lo[k], o.lo[k]);
    b.hi[k] = mx(b.hi[k], o.hi[k]);
  }
}
__device__ __forceinline__ void extend_pt(Box& b, const float* p) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = mn(b.lo[k], p[k]);
    b.hi[k] = mx(b.hi[k], p[k]);
  }
}
// BBox::computeArea (Core/Geometry/BBox.h:117-121): float arithmetic.
__device__ __forceinline__ float area(const Box& b) {
  const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
  return 2 * (dx * dy + dy * dz + dx * dz);
}
__device__ __forceinline__ Box rbox(const RefNode& n) {
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = n.box.lo[k];
    b.hi[k] = n.box.hi[k];
  }
  return b;
}
__device__ __forceinline__ void set_box(RefNode& n, const Box& b) {
  for (int k = 0; k < 3; ++k) {
    n.box.lo[k] = b.lo[k];
    n.box.hi[k] = b.hi[k];
  }
}
// Object box: from u.obox when the caller filled it (k_obj_bounds), else from the
// triangle's vertices with the same arithmetic (Mesh::getBBox, Model/Groups/Mesh.cc:228-237).
__device__ __forceinline__ Box obj_box(const DevUpdate& u, int obj) {
  Box b;
  if (u.obox) {
    const float* p = u.obox + 6 * (size_t)obj;
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = p[k];
      b.hi[k] = p[3 + k];
    }
    return b;
  }
  const uint32_t* t = u.idx + 3 * (size_t)obj;
  const float* p0 = u.verts + 3 * (size_t)t[0];
  for (int k = 0; k < 3; ++k) b.lo[k] = b.hi[k] = p0[k];
  extend_pt(b, u.verts + 3 * (size_t)t[1]);
  extend_pt(b, u.verts + 3 * (size_t)t[2]);
  return b;
}
// BBox::center = Interpolate(min, max, 0.5f) (BBox.h:63-65, MiscMath.h:136-139)
__device__ __forceinline__ float box_cen(const Box& b, int axis) {
  return b.hi[axis] * 0.5f + b.lo[axis] * (1 - 0.5f);
}
// Object centroid: from u.ocen when filled, else from the object box.
__device__ __forceinline__ float obj_cen(const DevUpdate& u, int obj, int axis) {
  if (u.ocen) return u.ocen[3 * (size_t)obj + axis];
  return box_cen(obj_box(u, obj), axis);
}
// Vector::indexOfMaxComponent (Core/Geometry/Vector.h:382-393)
__device__ __forceinline__ int index_of_max(float a, float b, float c) {
  if (a > b) return a > c ? 0 : 2;
  return b > c ? 1 : 2;
}
// BBox::longestAxis (Core/Geometry/BBox.h:123-129)
__device__ __forceinline__ int longest_axis(const Box& b) {
  return index_of_max(b.hi[0] - b.lo[0], b.hi[1] - b.lo[1], b.hi[2] - b.lo[2]);
}

struct Partition {
  Box l, r;
  int axis, split;
};

// partition2Objs's decision (DynBVH.cc:2277-2290): whether two objects with boxes left,
// right under a node with box node_box are split.
__device__ inline bool pair_splits(const Box& node_box, const Box& left, const Box& right) {
  const float nodeArea = area(node_box);
  const float leftArea = area(left);
  const float rightArea = area(right);
  const float leafCost = kCIsec * 2;
  const float newCost = kCTrav + (leftArea * kCIsec + rightArea * kCIsec) / nodeArea;
  return newCost < leafCost;
}
__device__ inline bool pair_splits(const DevUpdate& u, const Box& node_box, int b) {
  return pair_splits(node_box, obj_box(u, u.ids[b]), obj_box(u, u.ids[b + 1]));
}

// partition2Objs (DynBVH.cc:2277-2308)
__device__ inline Partition partition2(const DevUpdate& u, int node, int b) {
  Partition p;
  reset(p.l);
  reset(p.r);
  int32_t* ids = u.ids;
  if (pair_splits(u, rbox(u.refs[node]), b)) {
    const Box nb = rbox(u.refs[node]);
    const int axis = index_of_max(nb.hi[0] - nb.lo[0], nb.hi[1] - nb.lo[1], nb.hi[2] - nb.lo[2]);
    p.axis = axis;
    p.split = b + 1;
    if (obj_box(u, ids[b]).lo[axis] > obj_box(u, ids[b + 1]).lo[axis]) {
      const int t = ids[b];
      ids[b] = ids[b + 1];
      ids[b + 1] = t;
    }
    p.l = obj_box(u, ids[b]);
    p.r = obj_box(u, ids[b + 1]);
    return p;
  }
  p.axis = -1;
  p.split = 0;
  return p;
}

// partitionObjects (DynBVH.cc:1369-1396)
__device__ inline int partition_objects(const DevUpdate& u, int b, int e, int axis, float position) {
  int32_t* ids = u.ids;
  int first = b - 1, last = e;
  for (;;) {
    for (++first; first < last; first++)
      if (obj_cen(u, ids[first], axis) >= position) break;
    for (--last; first < last; --last)
      if (obj_cen(u, ids[last], axis) < position) break;
    if (first < last) {
      const int t = ids[first];
      ids[first] = ids[last];
      ids[last] = t;
    } else {
      return first;
    }
  }
}

// The binned SAH choice of partitionApproxSAH (DynBVH.cc:2125-2246) for n >= 3 objects
// [b, e) under a node with box node_box: the split axis (-1: the node stays a leaf) and
// position. LONGEST_AXIS is defined (DynBVH.cc:39), so only the centroid box's longest
// axis is binned. Reads only.
struct SplitChoice {
  int axis;
  float pos;
};
__device__ inline SplitChoice approx_choice(const DevUpdate& u, const Box& node_box, int b, int e) {
  const int n = e - b;
  const int32_t* ids = u.ids;
  Box cb;
  reset(cb);
  for (int i = b; i < e; ++i) {
    float c[3];
    if (u.ocen) {
      for (int k = 0; k < 3; ++k) c[k] = u.ocen[3 * (size_t)ids[i] + k];
    } else {
      const Box ob = obj_box(u, ids[i]);
      for (int k = 0; k < 3; ++k) c[k] = box_cen(ob, k);
    }
    extend_pt(cb, c);
  }
  const float inv_parent_area = (float)(1.0 / (double)area(node_box));
  float best_cost = n * kCIsec;
  int best_axis = -1;
  float best_pos = FLT_MAX;
  const int axis = longest_axis(cb);
  const float width = cb.hi[axis] - cb.lo[axis];
  const float scale = ((float)kSamples / width) * .999f;
  Box bins[kSamples];
  int counts[kSamples];
  for (int i = 0; i < kSamples; ++i) {
    reset(bins[i]);
    counts[i] = 0;
  }
  for (int i = b; i < e; ++i) {
    const Box ob = obj_box(u, ids[i]);
    const float cen = u.ocen ? u.ocen[3 * (size_t)ids[i] + axis] : box_cen(ob, axis);
    const float f = (cen - cb.lo[axis]) * scale;
    // int(f) as the reference's x86 build converts it (cvttss2si): NaN and values outside
    // the int range give INT_MIN, i.e. bin 0 after the clamp.
    int which = (f >= -2147483648.f && f < 2147483648.f) ? (int)f : INT_MIN;
    which = which < 0 ? 0 : (which > kSamples - 1 ? kSamples - 1 : which);
    counts[which]++;
    extend(bins[which], ob);
  }
  float left_areas[kSamples], right_areas[kSamples];
  int left_counts[kSamples], right_counts[kSamples];
  Box lb, rb;
  reset(lb);
  reset(rb);
  int nl = 0, nr = n;
  for (int i = 0; i < kSamples; ++i) {
    nl += counts[i];
    nr -= counts[i];
    extend(lb, bins[i]);
    left_counts[i] = nl;
    right_counts[i] = nr;
    left_areas[i] = area(lb);
  }
  for (int i = kSamples - 1; i >= 0; --i) {
    right_areas[i] = (i == kSamples - 1) ? FLT_MAX : area(rb);
    extend(rb, bins[i]);
  }
  for (int i = 0; i < kSamples - 1; ++i) {
    const float cost = (left_areas[i] * left_counts[i] + right_areas[i] * right_counts[i]) *
                           inv_parent_area * kCIsec + kCTrav;
    if (cost < best_cost) {
      best_cost = cost;
      best_axis = axis;
      best_pos = cb.lo[axis] + width * (static_cast<float>(i + 1) / kSamples);
    }
  }
  return SplitChoice{best_axis, best_pos};
}

// partitionApproxSAH (DynBVH.cc:2125-2275)
__device__ inline Partition partition_approx(const DevUpdate& u, int node, int b, int e) {
  Partition p;
  reset(p.l);
  reset(p.r);
  const int n = e - b;
  if (n == 1) {
    p.axis = -1;
    p.split = -1;
    return p;
  }
  if (n == 2) return partition2(u, node, b);
  const int32_t* ids = u.ids;
  const SplitChoice ch = approx_choice(u, rbox(u.refs[node]), b, e);
  const int best_axis = ch.axis;
  const float best_pos = ch.pos;
  p.axis = best_axis;
  if (p.axis != -1) {
    int middle = partition_objects(u, b, e, p.axis, best_pos);
    if (middle == b || middle == e) {
      if (n < 3) {
        p.axis = -1;
        p.split = 0;
        return p;
      }
      middle = b + n / 2;
    }
    for (int i = b; i < middle; ++i) extend(p.l, obj_box(u, ids[i]));
    for (int i = middle; i < e; ++i) extend(p.r, obj_box(u, ids[i]));
    p.split = middle;
    return p;
  }
  p.split = 0;
  return p;
}

// Record pairs for new inner nodes: from the shared counter (leaf splits of the update,
// which are rare), or from a range reserved beforehand (one lane building a whole
// subtree), which keeps the lanes of the device build off one hot counter.
struct CounterPairs {
  __device__ unsigned take(const DevUpdate& u) { return atomicAdd(u.next_free, 2u); }
};
struct RangePairs {
  unsigned next, end;
  __device__ unsigned take(const DevUpdate&) {
    if (next + 2 > end) return 0xFFFFFFFEu;  // cannot happen: a subtree of n objects needs <= 2(n-1)
    const unsigned s = next;
    next += 2;
    return s;
  }
};

// Turns `node` into an inner node over a new record pair whose records wait (axis =
// kPending) to be built over [b, split) and [split, e). False if no pair is left.
template <class Pairs>
__device__ inline bool make_inner(const DevUpdate& u, Pairs& pairs, int node, const Partition& p,
                                  int b, int e) {
  const unsigned spot = pairs.take(u);
  if (spot >= (unsigned)u.n_records || spot + 1 >= (unsigned)u.n_records) {
    atomicExch(u.err, 1);
    return false;
  }
  RefNode& nd = u.refs[node];
  nd.child = (int32_t)spot;
  nd.children = 0;
  nd.axis = (uint8_t)p.axis;
  RefNode& c0 = u.refs[spot];
  RefNode& c1 = u.refs[spot + 1];
  set_box(c0, p.l);
  set_box(c1, p.r);
  c0.child = b;
  c0.children = (int16_t)(p.split - b);
  c0.axis = kPending;
  c1.child = p.split;
  c1.children = (int16_t)(e - p.split);
  c1.axis = kPending;
  u.parent[spot] = node;
  u.parent[spot + 1] = node;
  return true;
}

// DynBVH::build(node, begin, end, useApproximateBuild = true) (DynBVH.cc:1804-1857) over
// every pending record of the subtree under `top`, starting at the pending record
// `start` (start is top itself, or top's left child), depth-first without a stack: parent
// links lead back up, and the climb stops at `top`, whose subtree size it then sets.
template <class Pairs>
__device__ inline void build_loop(const DevUpdate& u, Pairs& pairs, int start, int top) {
  RefNode* refs = u.refs;
  int cur = start;
  for (;;) {
    // cur waits to be built over [child, child + children)
    const int b = refs[cur].child, e = b + refs[cur].children;
    const Partition p = partition_approx(u, cur, b, e);
    if (p.axis != -1) {
      if (!make_inner(u, pairs, cur, p, b, e)) return;  // no pair left: reported through err
      cur = refs[cur].child;
      continue;
    }
    RefNode& nd = refs[cur];  // makeLeaf, its cost and size, std::sort of its ids
    nd.child = b;
    nd.children = (int16_t)(e - b);
    nd.axis = 0;
    u.costs[cur] = nd.children * kCIsec;
    u.ssize[cur] = 1;
    for (int i = b + 1; i < e; ++i) {
      const int32_t v = u.ids[i];
      int j = i - 1;
      while (j >= b && u.ids[j] > v) {
        u.ids[j + 1] = u.ids[j];
        --j;
      }
      u.ids[j + 1] = v;
    }
    // climb: finish parents whose right subtree is done, move to the next right sibling
    for (;;) {
      if (cur == top) return;
      const int par = u.parent[cur];
      if (cur == refs[par].child) {
        cur = cur + 1;
        break;
      }
      u.ssize[par] = 1 + u.ssize[refs[par].child] + u.ssize[refs[par].child + 1];
      cur = par;
    }
  }
}

// build(leaf, begin, end, useApproximateBuild = true, nodeAlreadyExists = true), as
// rotateNode calls it on a leaf child (DynBVH.cc:3057-3069): the leaf stays unless the
// approximate SAH splits it.
__device__ inline void split_leaf(const DevUpdate& u, int leaf) {
  RefNode* refs = u.refs;
  const int b0 = refs[leaf].child, e0 = b0 + refs[leaf].children;
  const Partition p0 = partition_approx(u, leaf, b0, e0);
  if (p0.axis == -1) return;
  CounterPairs pairs;
  if (!make_inner(u, pairs, leaf, p0, b0, e0)) return;
  build_loop(u, pairs, refs[leaf].child, leaf);
}

}  // namespace
}  // namespace mh
