// Rotating update for gfx950: DynBVH::update with TREE_ROT set (Model/Groups/DynBVH.h:377),
// the reference's default, on the device.
//
// The reference refits in post-order and calls rotateNode on every inner node after its
// subtree is done (updateBounds<false>, DynBVH.cc:2078-2123). rotateNode first splits
// leaf children with the approximate SAH build (DynBVH.cc:3057-3069 -> build with
// useApproximateBuild, partitionApproxSAH DynBVH.cc:2125-2275), then picks the cheapest
// of six rotations by its cost model and swaps records, and re-guesses split axes
// (DynBVH.cc:3051-3372). A node only touches records of its own subtree, so all nodes of
// one depth (depth taken at the start of the frame) run in parallel, deepest first, one
// thread per node:
//   k_obj_bounds     per object: box and centroid (DynBVH.cc:2018-2030), for the device
//                    build; the update forms them from the vertices where it needs them
//   k_upd_leaves     one lane per object slot, the first slot of each leaf doing the leaf:
//                    leaf boxes (DynBVH.cc:2091-2097), and a list of the leaves the
//                    parent's rotation may split
//   k_split_leaves   those splits, ahead of the rotations (same subtrees)
//   k_upd_level      one launch per depth: union of the child boxes, then rotateNode
//   k_upd_top        the same for the small depths at the top, in one workgroup
//   k_frontier       one launch per depth, top down: the next frame's level lists and
//                    leaf list, and the PairNodes (and big leaf counts) of the new tree
// (k_rebuild_pairs and the depth pass k_parents, k_jump, k_depth_hist, k_depth_scatter
// serve the device build, whose records are not yet in tree order.)
// Arithmetic is the reference's: Real is float, BBox::computeArea is float arithmetic
// returned as double, min/max are the select forms, -ffp-contract=off. New record pairs
// are taken with an atomic counter, so their numbering differs from the reference's
// serial order while the tree is the same.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "approx_sah.h"
#include "mh_device.h"

namespace mh {
namespace {

constexpr int kBlock = 256;
__device__ __forceinline__ void swap_rec(const DevUpdate& u, int a, int b) {
  int4* pa = reinterpret_cast<int4*>(u.refs + a);
  int4* pb = reinterpret_cast<int4*>(u.refs + b);
  const int4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
  pa[0] = b0;
  pa[1] = b1;
  pb[0] = a0;
  pb[1] = a1;
  const float c = u.costs[a];
  u.costs[a] = u.costs[b];
  u.costs[b] = c;
  const uint32_t z = u.ssize[a];
  u.ssize[a] = u.ssize[b];
  u.ssize[b] = z;
}

__device__ __forceinline__ Box unite(const DevUpdate& u, int a, int b) {
  Box x = rbox(u.refs[a]);
  extend(x, rbox(u.refs[b]));
  return x;
}

// Step 3 of rotateNode: the axis of largest overlap and child order (DynBVH.cc:3305-3352).
__device__ void guess_axis(const DevUpdate& u, int parent, int a, int b) {
  const Box x = rbox(u.refs[a]), y = rbox(u.refs[b]);
  float o[3];
  for (int k = 0; k < 3; ++k) o[k] = mx(x.lo[k], y.lo[k]) - mn(x.hi[k], y.hi[k]);
  const int ax = index_of_max(o[0], o[1], o[2]);
  u.refs[parent].axis = (uint8_t)ax;
  const float ha = ax == 0 ? x.hi[0] : (ax == 1 ? x.hi[1] : x.hi[2]);
  const float hb = ax == 0 ? y.hi[0] : (ax == 1 ? y.hi[1] : y.hi[2]);
  if (ha > hb) swap_rec(u, a, b);
}

// DynBVH::rotateNode (DynBVH.cc:3051-3372)
__device__ void rotate_node(const DevUpdate& u, int node) {
  RefNode* refs = u.refs;
  float* costs = u.costs;
  uint32_t* ss = u.ssize;
  const int L = refs[node].child, R = L + 1;
  // rotateNode first splits leaf children of more than one object with the approximate SAH
  // (DynBVH.cc:3057-3069); k_split_leaves has done that for every leaf before the levels.
  const int LL = refs[L].child, LR = LL + 1, RL = refs[R].child, RR = RL + 1;
  const bool l_inner = refs[L].children == 0, r_inner = refs[R].children == 0;
  const float nodeArea = area(rbox(refs[node]));
  const float leftArea = area(rbox(refs[L]));
  const float rightArea = area(rbox(refs[R]));
  float llA = 0.f, lrA = 0.f, rlA = 0.f, rrA = 0.f;
  Box boxA, boxB;
  reset(boxA);
  reset(boxB);
  float areaA = 0.f, areaB = 0.f;
  int best = 0;
  float bestCost = nodeArea * kCTrav + leftArea * costs[L] + rightArea * costs[R];
  if (l_inner) {
    llA = area(rbox(refs[LL]));
    lrA = area(rbox(refs[LR]));
    const float partial = nodeArea * kCTrav + rightArea * costs[R] + lrA * costs[LR] + llA * costs[LL];
    const Box b1 = unite(u, R, LR);
    const float a1 = area(b1);
    const float c1 = a1 * kCTrav + partial;
    if (c1 < bestCost) {
      best = 1;
      bestCost = c1;
      boxA = b1;
      areaA = a1;
    }
    const Box b2 = unite(u, R, LL);
    const float a2 = area(b2);
    const float c2 = a2 * kCTrav + partial;
    if (c2 < bestCost) {
      best = 2;
      bestCost = c2;
      boxA = b2;
      areaA = a2;
    }
  }
  if (r_inner) {
    rlA = area(rbox(refs[RL]));
    rrA = area(rbox(refs[RR]));
    const float partial = nodeArea * kCTrav + leftArea * costs[L] + rrA * costs[RR] + rlA * costs[RL];
    const Box b3 = unite(u, L, RR);
    const float a3 = area(b3);
    const float c3 = a3 * kCTrav + partial;
    if (c3 < bestCost) {
      best = 3;
      bestCost = c3;
      boxA = b3;
      areaA = a3;
    }
    const Box b4 = unite(u, L, RL);
    const float a4 = area(b4);
    const float c4 = a4 * kCTrav + partial;
    if (c4 < bestCost) {
      best = 4;
      bestCost = c4;
      boxA = b4;
      areaA = a4;
    }
    if (l_inner) {
      const float partial2 = nodeArea * kCTrav + llA * costs[LL] + lrA * costs[LR] + rlA * costs[RL] + rrA * costs[RR];
      const Box b5l = unite(u, LL, RL), b5r = unite(u, LR, RR);
      const float a5l = area(b5l), a5r = area(b5r);
      const float c5 = (a5l + a5r) * kCTrav + partial2;
      if (c5 < bestCost) {
        best = 5;
        bestCost = c5;
        boxA = b5l;
        areaA = a5l;
        boxB = b5r;
        areaB = a5r;
      }
      const Box b6l = unite(u, LL, RR), b6r = unite(u, RL, LR);
      const float a6l = area(b6l), a6r = area(b6r);
      const float c6 = (a6l + a6r) * kCTrav + partial2;
      if (c6 < bestCost) {
        best = 6;
        bestCost = c6;
        boxA = b6l;
        areaA = a6l;
        boxB = b6r;
        areaB = a6r;
      }
    }
  }
  costs[node] = bestCost / nodeArea;
  if (best == 0) return;
  const float invA = 1 / areaA;
  switch (best) {
    case 1:
      set_box(refs[L], boxA);
      costs[L] = kCTrav + ((lrA * costs[LR] + rightArea * costs[R]) * invA);
      swap_rec(u, LL, R);
      ss[L] = ss[LL] + ss[LR] + 1;
      break;
    case 2:
      set_box(refs[L], boxA);
      costs[L] = kCTrav + ((llA * costs[LL] + rightArea * costs[R]) * invA);
      swap_rec(u, LR, R);
      ss[L] = ss[LL] + ss[LR] + 1;
      break;
    case 3:
      set_box(refs[R], boxA);
      costs[R] = kCTrav + ((rrA * costs[RR] + leftArea * costs[L]) * invA);
      swap_rec(u, RL, L);
      ss[R] = ss[RL] + ss[RR] + 1;
      break;
    case 4:
      set_box(refs[R], boxA);
      costs[R] = kCTrav + ((rlA * costs[RL] + leftArea * costs[L]) * invA);
      swap_rec(u, RR, L);
      ss[R] = ss[RL] + ss[RR] + 1;
      break;
    case 5:
      set_box(refs[L], boxA);
      set_box(refs[R], boxB);
      costs[L] = kCTrav + ((llA * costs[LL] + rlA * costs[RL]) * invA);
      costs[R] = kCTrav + ((lrA * costs[LR] + rrA * costs[RR]) / areaB);
      swap_rec(u, LR, RL);
      ss[L] = ss[LL] + ss[LR] + 1;
      ss[R] = ss[RL] + ss[RR] + 1;
      break;
    default:  // 6
      set_box(refs[L], boxA);
      set_box(refs[R], boxB);
      costs[L] = kCTrav + ((llA * costs[LL] + rrA * costs[RR]) * invA);
      costs[R] = kCTrav + ((lrA * costs[LR] + rlA * costs[RL]) / areaB);
      swap_rec(u, LR, RR);
      ss[L] = ss[LL] + ss[LR] + 1;
      ss[R] = ss[RL] + ss[RR] + 1;
      break;
  }
  if (best <= 2 || best >= 5) guess_axis(u, L, LL, LR);
  if (best >= 3) guess_axis(u, R, RL, RR);
  guess_axis(u, node, L, R);
  if (refs[node].bits & 2) {
    auto mark = [&](int r) {
      if (ss[r] > u.large)
        refs[r].bits |= 2;
      else
        refs[r].bits &= ~2;
    };
    mark(L);
    mark(R);
    if (best <= 2 || best >= 5) {
      mark(LL);
      mark(LR);
    }
    if (best >= 3) {
      mark(RL);
      mark(RR);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_obj_bounds(DevUpdate u) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= u.n_objects) return;
  Box b;
  const float* p0 = u.verts + 3 * (size_t)u.idx[3 * i + 0];
  for (int k = 0; k < 3; ++k) b.lo[k] = b.hi[k] = p0[k];  // Mesh::getBBox (Mesh.cc:228-237)
  extend_pt(b, u.verts + 3 * (size_t)u.idx[3 * i + 1]);
  extend_pt(b, u.verts + 3 * (size_t)u.idx[3 * i + 2]);
  float* o = u.obox + 6 * (size_t)i;
  for (int k = 0; k < 3; ++k) {
    o[k] = b.lo[k];
    o[3 + k] = b.hi[k];
    // BBox::center = Interpolate(min, max, 0.5f) (BBox.h:63-65, MiscMath.h:136-139)
    u.ocen[3 * (size_t)i + k] = b.hi[k] * 0.5f + b.lo[k] * (1 - 0.5f);
  }
}

// The leaf that starts at each object slot, from a leaf list.
__global__ __launch_bounds__(kBlock) void k_leaf_starts(DevUpdate u, const int32_t* leaves,
                                                        int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int rec = leaves[i];
  u.leaf_at[u.refs[rec].child] = rec;
}

// One lane per object slot; the lane at a leaf's first slot does the leaf, so neighbouring
// lanes take neighbouring leaves and gather neighbouring vertices. Object boxes come from
// the vertices (Mesh::getBBox).
__global__ __launch_bounds__(kBlock) void k_upd_leaves(DevUpdate u, int64_t n_slots) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_slots) return;
  const int rec = u.leaf_at[i];
  if (rec < 0) return;  // not the first slot of a leaf
  RefNode& nd = u.refs[rec];
  const int first = nd.child, cnt = nd.children;
  Box b;
  reset(b);
  auto object = [&](int slot) { return obj_box(u, u.ids[slot]); };
  // The parent's rotation tries to split this leaf with the approximate SAH. Two objects:
  // decide it now, from the same object boxes (partition2Objs); more: leave the binned
  // decision to the split pass, whose registers it needs. A root leaf has no parent, a
  // one-object leaf never splits.
  bool maybe;
  if (cnt == 2) {  // the common leaf: both boxes once, for the leaf box and the decision
    const Box o0 = object(first), o1 = object(first + 1);
    extend(b, o0);
    extend(b, o1);
    maybe = pair_splits(b, o0, o1);
  } else {
    for (int k = 0; k < cnt; ++k) extend(b, object(first + k));
    maybe = cnt > 2;
  }
  set_box(nd, b);
  nd.axis = 0;
  // the candidates go to the split list, one atomic per wave
  const bool cand = rec != 0 && maybe;
  const unsigned long long m = __ballot(cand);
  if (m) {
    const int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(u.split_count, (unsigned)__popcll(m));
    base = (unsigned)__shfl((int)base, leader);
    if (cand) u.split_list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = rec;
  }
}

// The leaf splits of this frame's rotations, ahead of them, for the leaves the leaf pass
// listed (split_leaf decides again and keeps a leaf the SAH does not split). rotateNode
// splits a leaf child when its parent rotates (DynBVH.cc:3057-3069), from the leaf's own
// objects and box only; nothing before that rotation changes the leaf (rotations below the
// parent do not reach a sibling leaf), and the nodes a split creates are in none of this
// frame's level lists, so splitting the listed leaves here builds the same subtrees. The
// split code's registers stay out of the level kernels.
__global__ __launch_bounds__(kBlock) void k_split_leaves(DevUpdate u) {
  const unsigned n = *u.split_count;
  for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    split_leaf(u, u.split_list[i]);
}

// One inner node of the update: the union of its child boxes, then rotateNode.
__device__ __forceinline__ void upd_node(const DevUpdate& u, int node) {
  const int c = u.refs[node].child;
  Box b;
  reset(b);
  extend(b, rbox(u.refs[c]));
  extend(b, rbox(u.refs[c + 1]));
  set_box(u.refs[node], b);
  rotate_node(u, node);
}

__global__ __launch_bounds__(kBlock) void k_upd_level(DevUpdate u, const int32_t* recs,
                                                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  upd_node(u, recs[i]);
}

// The top depths [0, t.n), deepest first, in one workgroup: the barrier between depths
// orders them as separate launches would (the workgroup's waves share one compute unit
// and its L1, so stores before the barrier are seen by loads after it).
__global__ __launch_bounds__(kBlock) void k_upd_top(DevUpdate u, const int32_t* recs, TopLevels t) {
  for (int d = t.n - 1; d >= 0; --d) {
    for (int i = t.off[d] + (int)threadIdx.x; i < t.off[d + 1]; i += kBlock) upd_node(u, recs[i]);
    __syncthreads();
  }
}

// Records in use: the counter, capped at the record array (a split that found the array
// full has already raised err, and the host discards the frame).
__device__ __forceinline__ int64_t used(const unsigned* next_free, int32_t n_records) {
  const unsigned n = *next_free;
  return n < (unsigned)n_records ? (int64_t)n : (int64_t)n_records;
}

// A record as two 16-byte words: lo.x lo.y lo.z hi.x | hi.y hi.z child word, where word
// holds axis (byte 0), bits (byte 1) and children (int16, bytes 2-3).
struct RecWords {
  float4 a, b;
};
__device__ __forceinline__ RecWords load_rec(const RefNode* refs, int64_t r) {
  const float4* p = reinterpret_cast<const float4*>(refs + r);
  return RecWords{p[0], p[1]};
}
__device__ __forceinline__ int rec_child(const RecWords& w) { return __float_as_int(w.b.z); }
__device__ __forceinline__ int rec_children(const RecWords& w) {
  return __float_as_int(w.b.w) >> 16;  // arithmetic shift: the int16 with its sign
}

// Child reference of a PairNode for a loaded record: its pair index (inner) or its leaf
// code, with big_count set for leaves of 15 or more objects.
__device__ __forceinline__ int32_t encode_rec(const DevUpdate& u, const RecWords& w) {
  const int children = rec_children(w), child = rec_child(w);
  if (children == 0) return (child - 1) >> 1;
  const int c = children < kLeafCountInline ? children : kLeafCountInline;
  if (children >= kLeafCountInline) u.big_count[child] = children;
  return ~((child << kLeafCountBits) | c);
}

// The word of a record holding axis (byte 0), bits (byte 1) and children (bytes 2-3).
__device__ __forceinline__ int rec_word(const RecWords& w) { return __float_as_int(w.b.w); }

// The PairNode of inner record r (first child c, word `word`) from its two child records
// c0, c1, as four 16-byte stores.
__device__ __forceinline__ void store_pair(const DevUpdate& u, int64_t r, int c, int word,
                                           const RecWords& c0, const RecWords& c1) {
  float4* P = reinterpret_cast<float4*>(u.pairs + ((c - 1) >> 1));
  P[0] = make_float4(c0.a.x, c0.a.w, c0.a.y, c0.b.x);  // child 0: lo.x hi.x lo.y hi.y
  P[1] = make_float4(c0.a.z, c0.b.y, c1.a.x, c1.a.w);  // lo.z hi.z | child 1: lo.x hi.x
  P[2] = make_float4(c1.a.y, c1.b.x, c1.a.z, c1.b.y);  // lo.y hi.y lo.z hi.z
  reinterpret_cast<int4*>(P)[3] =
      make_int4(encode_rec(u, c0), encode_rec(u, c1), (word & 3) | (((word >> 8) & 1) << 2), (int)r);
}

// The PairNode of every inner record in use.
__global__ __launch_bounds__(kBlock) void k_rebuild_pairs(DevUpdate u) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= used(u.next_free, u.n_records)) return;
  const RecWords nd = load_rec(u.refs, r);
  if (rec_children(nd) != 0) return;
  const int c = rec_child(nd);
  store_pair(u, r, c, rec_word(nd), load_rec(u.refs, c), load_rec(u.refs, c + 1));
}

// Level lists of the updated tree, top down, and its PairNodes. fr holds kFrontierWords
// counters: fr[d] inner records of depth d (d < kMaxDepth - 1), fr[kMaxDepth] leaves,
// fr[kMaxDepth + 1] set when a list would pass its capacity, and from fr + kMaxDepth + 2
// the list offsets: depth d is levels[off[d], off[d] + fr[d]). Beside each list entry,
// aux holds its record's first child and word, copied from the record when its parent
// listed it, so that no depth loads its own records again. The root is an inner record.
//
// One workgroup's share of depth d, whose records are levels[first, first + n): items
// i_begin, i_begin + stride, ... Each record writes the PairNode of its two children and
// appends them, in child order, to depth d + 1 (inner; that list starts at first + n) or
// to the leaf list. Per loop iteration the workgroup ranks its appends with ballots and
// takes its ranges with one call of reserve(inner, leaves, inner base, leaf base) by its
// first lane. Order inside a list is free: the nodes of one depth are independent.
template <class Reserve>
__device__ __forceinline__ void frontier_depth(const DevUpdate& u, unsigned* fr, int32_t* levels,
                                               int2* aux, int32_t* leaves, unsigned cap,
                                               unsigned first, unsigned n, unsigned i_begin,
                                               unsigned stride, Reserve reserve) {
  __shared__ unsigned w_in[kBlock / 64], w_lf[kBlock / 64];
  __shared__ unsigned b_in, b_lf;
  const unsigned next = first + n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  auto put = [&](int32_t* list, unsigned pos, int v) {
    if (pos < cap)
      list[pos] = v;
    else
      fr[kMaxDepth + 1] = 1u;
  };
  for (unsigned i0 = i_begin; i0 < n; i0 += stride) {  // uniform over the workgroup
    const unsigned i = i0 + threadIdx.x;
    const bool live = i < n && first + i < cap;  // past cap only after an overflow (flagged)
    int c = 0;
    bool in0 = false, in1 = false;
    int2 ax0 = make_int2(0, 0), ax1 = make_int2(0, 0);
    if (live) {
      const int r = levels[first + i];
      const int2 ax = aux[first + i];
      c = ax.x;
      const RecWords c0 = load_rec(u.refs, c), c1 = load_rec(u.refs, c + 1);
      store_pair(u, r, c, ax.y, c0, c1);
      in0 = rec_children(c0) == 0;
      in1 = rec_children(c1) == 0;
      ax0 = make_int2(rec_child(c0), rec_word(c0));
      ax1 = make_int2(rec_child(c1), rec_word(c1));
    }
    const int k_in = live ? (int)in0 + (int)in1 : 0, k_lf = live ? 2 - k_in : 0;
    const unsigned long long i1 = __ballot(k_in >= 1), i2 = __ballot(k_in == 2);
    const unsigned long long l1 = __ballot(k_lf >= 1), l2 = __ballot(k_lf == 2);
    if (lane == 0) {
      w_in[wave] = (unsigned)(__popcll(i1) + __popcll(i2));
      w_lf[wave] = (unsigned)(__popcll(l1) + __popcll(l2));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned t_in = 0, t_lf = 0;
      for (int w = 0; w < kBlock / 64; ++w) {
        const unsigned x = w_in[w], y = w_lf[w];
        w_in[w] = t_in;
        w_lf[w] = t_lf;
        t_in += x;
        t_lf += y;
      }
      unsigned bi = 0, bl = 0;
      reserve(t_in, t_lf, bi, bl);
      b_in = bi;
      b_lf = bl;
    }
    __syncthreads();
    if (live) {
      unsigned pi = next + b_in + w_in[wave] + (unsigned)(__popcll(i1 & below) + __popcll(i2 & below));
      unsigned pl = b_lf + w_lf[wave] + (unsigned)(__popcll(l1 & below) + __popcll(l2 & below));
      if (in0) {
        if (pi < cap) aux[pi] = ax0;
        put(levels, pi++, c);
      } else {
        put(leaves, pl++, c);
        if ((unsigned)ax0.x < cap) u.leaf_at[ax0.x] = c;  // the leaf starting at its first slot
      }
      if (in1) {
        if (pi < cap) aux[pi] = ax1;
        put(levels, pi, c + 1);
      } else {
        put(leaves, pl, c + 1);
        if ((unsigned)ax1.x < cap) u.leaf_at[ax1.x] = c + 1;
      }
    }
    __syncthreads();  // w_in, w_lf, b_in, b_lf are rewritten by the next iteration
  }
}

// Depths [0, n_top) in one workgroup, from the root: it clears fr, counts in LDS (a
// barrier ends each depth; the workgroup's waves share one compute unit and its L1, so
// the list entries stored before it are seen after it) and leaves fr[0..n_top], the
// offsets up to depth n_top and the leaf count for the launches of the deeper depths.
__global__ __launch_bounds__(kBlock) void k_frontier_top(DevUpdate u, unsigned* fr, int32_t* levels,
                                                         int2* aux, int32_t* leaves, unsigned cap,
                                                         int n_top) {
  __shared__ unsigned s_cnt[kTopLevels + 1], s_off[kTopLevels + 1], s_leaf;
  for (int i = threadIdx.x; i < kFrontierWords; i += kBlock) fr[i] = 0u;
  if (threadIdx.x == 0) {
    s_cnt[0] = 1u;  // depth 0: the root, at levels[0]
    s_off[0] = 0u;
    s_leaf = 0u;
    levels[0] = 0;
    const RecWords root = load_rec(u.refs, 0);
    aux[0] = make_int2(rec_child(root), rec_word(root));
  }
  __syncthreads();
  for (int d = 0; d < n_top; ++d) {
    const unsigned n = s_cnt[d], first = s_off[d];
    __syncthreads();  // every lane has read the depth's count before lane 0 opens the next
    if (threadIdx.x == 0) {
      s_off[d + 1] = first + n;
      s_cnt[d + 1] = 0u;
    }
    __syncthreads();
    frontier_depth(u, fr, levels, aux, leaves, cap, first, n, 0u, kBlock,
                   [&](unsigned t_in, unsigned t_lf, unsigned& bi, unsigned& bl) {
                     bi = s_cnt[d + 1];
                     s_cnt[d + 1] = bi + t_in;
                     bl = s_leaf;
                     s_leaf = bl + t_lf;
                   });
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    unsigned* off = fr + kMaxDepth + 2;
    for (int d = 0; d <= n_top; ++d) {
      fr[d] = s_cnt[d];
      off[d] = s_off[d];
    }
    fr[kMaxDepth] = s_leaf;
  }
}

// Depth d (>= the depths of k_frontier_top) over the whole grid, counters in fr.
__global__ __launch_bounds__(kBlock) void k_frontier(DevUpdate u, unsigned* fr, int32_t* levels,
                                                     int2* aux, int32_t* leaves, unsigned cap,
                                                     int d) {
  unsigned* off = fr + kMaxDepth + 2;
  const unsigned n = fr[d], first = off[d];
  if (blockIdx.x == 0 && threadIdx.x == 0) off[d + 1] = first + n;
  frontier_depth(u, fr, levels, aux, leaves, cap, first, n, blockIdx.x * kBlock,
                 gridDim.x * kBlock,
                 [&](unsigned t_in, unsigned t_lf, unsigned& bi, unsigned& bl) {
                   bi = t_in ? atomicAdd(fr + d + 1, t_in) : 0u;
                   bl = t_lf ? atomicAdd(fr + kMaxDepth, t_lf) : 0u;
                 });
}

// anc = parent record (root: -1), dep = 1 (root: 0) for every record in use.
__global__ __launch_bounds__(kBlock) void k_parents(DevUpdate u, int32_t* anc, int32_t* dep) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= used(u.next_free, u.n_records)) return;
  if (r == 0) {
    anc[0] = -1;
    dep[0] = 0;
  }
  const RefNode nd = u.refs[r];
  if (nd.children == kHole) {  // a record of the device build that no node uses
    anc[r] = -1;
    dep[r] = 0;
    return;
  }
  if (nd.children != 0) return;
  anc[nd.child] = (int32_t)r;
  anc[nd.child + 1] = (int32_t)r;
  dep[nd.child] = 1;
  dep[nd.child + 1] = 1;
}

// One pointer-jumping step: depth sums along ancestor links that double in length.
__global__ __launch_bounds__(kBlock) void k_jump(const unsigned* next_free, int32_t n_records,
                                                 const int32_t* anc_in, const int32_t* dep_in,
                                                 int32_t* anc_out, int32_t* dep_out) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= used(next_free, n_records)) return;
  const int32_t a = anc_in[r];
  if (a < 0) {
    anc_out[r] = a;
    dep_out[r] = dep_in[r];
    return;
  }
  anc_out[r] = anc_in[a];
  dep_out[r] = dep_in[r] + dep_in[a];
}

// hist[d] = inner records at depth d (d < kMaxDepth), counted in LDS first (one global
// atomic per bin and workgroup). hist[kMaxDepth] is set when an inner record lies at
// kMaxDepth or deeper. Holes are skipped.
// Adds one to bins[b] for every lane of the wave with b >= 0, with one LDS atomic per
// distinct value; returns the lane's place among the lanes counted into its bin by this
// call and earlier ones (lane order inside a call). Called by every lane of the wave.
__device__ __forceinline__ unsigned wave_bin_add(unsigned* bins, int b) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(b >= 0);
  unsigned place = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int v = __shfl(b, leader);
    const unsigned long long m = __ballot(b == v);
    unsigned before = 0;
    if (lane == leader) before = atomicAdd(&bins[v], (unsigned)__popcll(m));
    before = (unsigned)__shfl((int)before, leader);
    if (b == v) place = before + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    todo &= ~m;
  }
  return place;
}

// Records [blockIdx.x * chunk, + chunk) of a workgroup (chunk a multiple of kBlock), cut
// at the records in use.
__device__ __forceinline__ void depth_range(const DevUpdate& u, int64_t chunk, int64_t& lo,
                                            int64_t& hi) {
  const int64_t n = used(u.next_free, u.n_records);
  lo = (int64_t)blockIdx.x * chunk;
  hi = lo + chunk < n ? lo + chunk : n;
}

// Inner records per depth. Each workgroup counts a range of records in LDS and adds its
// counts to the global histogram once per depth, so the few shared counters see one
// atomic per workgroup and depth.
__global__ __launch_bounds__(kBlock) void k_depth_hist(DevUpdate u, const int32_t* anc,
                                                       const int32_t* dep, unsigned* hist,
                                                       int64_t chunk) {
  __shared__ unsigned lh[kMaxDepth];
  for (int i = threadIdx.x; i < kMaxDepth; i += kBlock) lh[i] = 0;
  __syncthreads();
  int64_t lo, hi;
  depth_range(u, chunk, lo, hi);
  for (int64_t r0 = lo; r0 < hi; r0 += kBlock) {
    const int64_t r = r0 + threadIdx.x;
    int b = -1;
    if (r < hi && anc[r] != -1) hist[kMaxDepth + 2] = 1u;  // the jumps did not reach the root
    if (r < hi && u.refs[r].children == 0) {
      const int d = dep[r];
      if (d < kMaxDepth)
        b = d;
      else
        hist[kMaxDepth] = 1u;
    }
    wave_bin_add(lh, b);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMaxDepth; i += kBlock)
    if (lh[i]) atomicAdd(hist + i, lh[i]);
}

// A full binary tree has one leaf more than inner records, and its deepest record is a
// leaf one below the deepest inner record: hist[kMaxDepth] = leaves, hist[kMaxDepth + 1] =
// deepest depth (kMaxDepth when the tree is too deep). One workgroup.
__global__ __launch_bounds__(kBlock) void k_depth_finish(unsigned* hist) {
  __shared__ unsigned inner[kBlock], deep[kBlock];
  unsigned n = 0, dmax = 0;
  for (int i = threadIdx.x; i < kMaxDepth; i += kBlock) {
    n += hist[i];
    if (hist[i]) dmax = (unsigned)i + 1 > dmax ? (unsigned)i + 1 : dmax;
  }
  inner[threadIdx.x] = n;
  deep[threadIdx.x] = dmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tn = 0, td = 0;
    for (int k = 0; k < kBlock; ++k) {
      tn += inner[k];
      td = deep[k] > td ? deep[k] : td;
    }
    const bool too_deep = hist[kMaxDepth] != 0;
    hist[kMaxDepth] = tn + 1;
    hist[kMaxDepth + 1] = too_deep ? (unsigned)kMaxDepth : td;
  }
}

// cursor[d]: next position for depth d in levels; cursor[kMaxDepth]: next leaf position.
// Each workgroup counts its range of records per depth (leaves in bin kMaxDepth), reserves
// one range per bin, then passes over its records again to write them into their ranges
// (order within a list does not matter). Inner depths are < kMaxDepth (checked by the host
// before this pass).
__global__ __launch_bounds__(kBlock) void k_depth_scatter(DevUpdate u, const int32_t* dep,
                                                          unsigned* cursor, int32_t* levels,
                                                          int32_t* leaves, int64_t chunk) {
  __shared__ unsigned cnt[kMaxDepth + 1], base[kMaxDepth + 1];
  for (int i = threadIdx.x; i <= kMaxDepth; i += kBlock) cnt[i] = 0;
  __syncthreads();
  int64_t lo, hi;
  depth_range(u, chunk, lo, hi);
  auto bin_of = [&](int64_t r) {
    if (r >= hi) return -1;
    const int16_t c = u.refs[r].children;
    if (c > 0) return kMaxDepth;
    if (c == 0 && dep[r] < kMaxDepth) return (int)dep[r];
    return -1;  // a hole
  };
  for (int64_t r0 = lo; r0 < hi; r0 += kBlock) wave_bin_add(cnt, bin_of(r0 + threadIdx.x));
  __syncthreads();
  for (int i = threadIdx.x; i <= kMaxDepth; i += kBlock) {
    base[i] = cnt[i] ? atomicAdd(cursor + i, cnt[i]) : 0u;
    cnt[i] = 0;
  }
  __syncthreads();
  for (int64_t r0 = lo; r0 < hi; r0 += kBlock) {
    const int64_t r = r0 + threadIdx.x;
    const int b = bin_of(r);
    const unsigned place = wave_bin_add(cnt, b);
    if (b == kMaxDepth)
      leaves[base[b] + place] = (int32_t)r;
    else if (b >= 0)
      levels[base[b] + place] = (int32_t)r;
  }
}

unsigned grid(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Workgroups of the depth histogram and scatter: at most kDepthBlocks, each over a range
// of a multiple of kBlock records covering the n record slots.
constexpr unsigned kDepthBlocks = 2048;
int64_t depth_chunk(int64_t n, unsigned& blocks) {
  const unsigned g = grid(n > 0 ? n : 1);
  const unsigned want = g < kDepthBlocks ? g : kDepthBlocks;
  const int64_t chunk = ((int64_t)g + want - 1) / want * kBlock;
  blocks = (unsigned)((n + chunk - 1) / chunk);
  if (blocks == 0) blocks = 1;
  return chunk;
}

}  // namespace

int launch_obj_bounds(const DevUpdate& u, void* stream) {
  if (u.n_objects <= 0) return 0;
  hipLaunchKernelGGL(k_obj_bounds, dim3(grid(u.n_objects)), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), u);
  return (int)hipGetLastError();
}

int launch_leaf_starts(const DevUpdate& u, const int32_t* leaves, int64_t n_leaves,
                       void* stream) {
  if (n_leaves > 0)
    hipLaunchKernelGGL(k_leaf_starts, dim3(grid(n_leaves)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), u, leaves, n_leaves);
  return (int)hipGetLastError();
}

int launch_update_tree(const DevUpdate& u, int64_t n_leaves, const int32_t* levels,
                       const int64_t* level_off, int n_levels, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (n_leaves > 0 && u.n_objects > 0) {
    hipLaunchKernelGGL(k_upd_leaves, dim3(grid(u.n_objects)), dim3(kBlock), 0, s, u,
                       (int64_t)u.n_objects);
    const unsigned g = grid(n_leaves);  // the list holds at most n_leaves entries
    hipLaunchKernelGGL(k_split_leaves, dim3(g < 1024 ? g : 1024), dim3(kBlock), 0, s, u);
  }
  TopLevels t;
  const int top = top_levels(level_off, n_levels, t);
  for (int d = n_levels - 1; d >= top; --d) {
    const int64_t first = level_off[d], n = level_off[d + 1] - level_off[d];
    if (n > 0)
      hipLaunchKernelGGL(k_upd_level, dim3(grid(n)), dim3(kBlock), 0, s, u, levels + first, n);
  }
  if (top > 0) hipLaunchKernelGGL(k_upd_top, dim3(1), dim3(kBlock), 0, s, u, levels, t);
  return (int)hipGetLastError();
}

int launch_frontier_top(const DevUpdate& u, unsigned* fr, int32_t* levels, int2* aux,
                        int32_t* leaves, unsigned cap, int n_top, void* stream) {
  if (n_top < 1 || n_top > kTopLevels || n_top >= kMaxDepth - 1) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_frontier_top, dim3(1), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), u, fr, levels, aux, leaves, cap, n_top);
  return (int)hipGetLastError();
}

int launch_frontier(const DevUpdate& u, unsigned* fr, int32_t* levels, int2* aux,
                    int32_t* leaves, unsigned cap, int d, unsigned blocks, void* stream) {
  if (d < 0 || d >= kMaxDepth - 1) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_frontier, dim3(blocks ? blocks : 1), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), u, fr, levels, aux, leaves, cap, d);
  return (int)hipGetLastError();
}

int launch_rebuild_pairs(const DevUpdate& u, void* stream) {
  hipLaunchKernelGGL(k_rebuild_pairs, dim3(grid(u.n_records)), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), u);
  return (int)hipGetLastError();
}

int launch_depths(const DevUpdate& u, int32_t* anc0, int32_t* dep0, int32_t* anc1, int32_t* dep1,
                  unsigned* hist, int steps, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const unsigned g = grid(u.n_records);
  hipLaunchKernelGGL(k_parents, dim3(g), dim3(kBlock), 0, s, u, anc0, dep0);
  // 2^8 > kMaxDepth: eight doublings reach the root from any record; fewer (an even count,
  // so that the results end in anc0/dep0) do for shallower trees, and the histogram pass
  // flags a record they left short of the root.
  if (steps < 2 || steps > 8 || (steps & 1)) steps = 8;
  for (int it = 0; it < steps; ++it) {
    const bool even = (it & 1) == 0;
    hipLaunchKernelGGL(k_jump, dim3(g), dim3(kBlock), 0, s, u.next_free, u.n_records,
                       even ? anc0 : anc1,
                       even ? dep0 : dep1, even ? anc1 : anc0, even ? dep1 : dep0);
  }
  unsigned blocks;
  const int64_t chunk = depth_chunk(u.n_records, blocks);
  hipLaunchKernelGGL(k_depth_hist, dim3(blocks), dim3(kBlock), 0, s, u, anc0, dep0, hist, chunk);
  hipLaunchKernelGGL(k_depth_finish, dim3(1), dim3(kBlock), 0, s, hist);
  return (int)hipGetLastError();
}

int launch_depth_scatter(const DevUpdate& u, const int32_t* dep, unsigned* cursor,
                         int32_t* levels, int32_t* leaves, void* stream) {
  unsigned blocks;
  const int64_t chunk = depth_chunk(u.n_records, blocks);
  hipLaunchKernelGGL(k_depth_scatter, dim3(blocks), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), u, dep, cursor, levels, leaves, chunk);
  return (int)hipGetLastError();
}

}  // namespace mh
