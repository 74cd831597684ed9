// Device data layout shared by the HIP kernels and the host library.
//
// HBM layout of one scene (all arrays 256-byte aligned, one allocation each):
//   refs      RefNode[2T]      reference DynBVH records, 32 B (packet traversal, refit)
//   pairs     PairNode[T]      one 64-B record per pair of sibling records (2p+1, 2p+2):
//                              both boxes, so one fetch decides both children of a node
//                              (single-ray traversal). Sibling records are allocated in
//                              pairs (nextFree += 2 from 1), so record r sits in pair
//                              (r - 1) / 2 on side (r - 1) & 1, and a node whose children
//                              are (c, c + 1) refers to pair (c - 1) / 2.
//   tris      float4[3T]       per object slot (object_ids order), 48 B:
//                              Kensler-Shirley: p0, e0 = p1-p0, e1 = p0-p2, n = e0 x e1
//                              Wald: n_u n_v n_d k | b_nu b_nv b_d n_k | c_nu c_nv c_d 0
//   obj_ids   int32[T]         object_ids
//   big_count int32[T]         leaf size at a leaf's first slot when it is >= 15
//   verts     float[3V], idx uint32[3T], base_verts float[3V]    mesh (refit, animation)
//   leaves int32[L], levels int32[I]   leaf records; inner records grouped by depth
//   costs float[2T], ssize uint32[2T], parent int32[2T], obj boxes/centroids   rotating
//                              update state (allocated at the first mh_scene_update)
#pragma once

#include <cstdint>

#include "sah_builder.h"

namespace mh {

// Child reference inside a PairNode: >= 0 is an inner-node (pair) index; < 0 is a leaf,
// ~((first_slot << 4) | c) with c = min(count, 15); c == 15 means "read big_count".
constexpr int kLeafCountBits = 4;
constexpr int kLeafCountInline = 15;
// Fewer than 2^26 objects: leaf codes hold first_slot << 4 in an int, and the pair table
// (64 B per object) stays within the 4 GiB that a buffer descriptor's byte offsets reach.
constexpr uint32_t kMaxObjects = 1u << 26;

struct alignas(16) PairNode {
  // child 0: lo.x hi.x lo.y hi.y | lo.z hi.z ; child 1: lo.x hi.x | lo.y hi.y lo.z hi.z
  float b[12];
  int32_t ref[2];
  int32_t meta;  // bits 0-1 split axis, bit 2 isLeftCheaper
  int32_t node;  // reference record id of this inner node
};
static_assert(sizeof(PairNode) == 64, "pair node is 64 bytes");


struct DevScene {
  const PairNode* pairs;
  const RefNode* refs;
  const float* tris;  // 12 floats per slot
  const int32_t* obj_ids;
  const int32_t* big_count;
  int32_t root_ref;   // pair-encoded reference of record 0
  int32_t prim_kind;  // 0 KS, 1 Wald
  uint32_t n_pairs;   // PairNodes allocated (the bound of the pair table's descriptor)
  const void* pairs_soa;  // kLayoutSoa: 16-byte field k of pair p at [k * n_pairs + p], or null
  uint32_t n_slots;   // triangle records allocated (the bound of the triangle table's)
};

struct DevRays {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  const uint8_t* masked;
};

struct DevHits {
  int32_t* prim;
  float *t, *u, *v;
  uint8_t* occluded;
};

struct LaunchCfg {
  int any_hit;
  int packet_size;    // 1 = single-ray kernel
  int stack_lds;      // 8/16/32 (single-ray)
  int blocks;         // grid size in blocks of 256
  int64_t* counters;  // device int64[2] or null
  int32_t* spill;     // per-thread stack spill (single-ray)
  int64_t spill_stride;
};

// Node layouts of the single-ray traversal (MH_LAYOUT_*). A non-root record r's box is
// also held by PairNode (r - 1) / 2, side (r - 1) & 1.
constexpr int kLayoutPair = 0;  // 64-byte PairNodes, both child boxes per fetch
constexpr int kLayoutRef = 1;   // 32-byte reference records, own box tested at visit
constexpr int kLayoutSoa = 2;   // the PairNodes as four arrays of their 16-byte fields

// One ray batch of a persistent launch: rays [start, start + count) of the launch's
// global numbering are rays [0, count) of this batch.
struct BatchDev {
  DevRays rays;
  DevHits hits;
  int64_t count;
  int64_t start;
  uint32_t flags;  // bit 0: any-hit
  uint32_t pad;
};

// Device-resident batch table: the kernel reads the batch count and total from here, so a
// captured graph stays valid when the table's contents change.
struct BatchTable {
  int64_t total;
  int32_t nb;
  int32_t pad;
  BatchDev batches[1];  // nb entries
};

// Bytes of a batch launch's counters (ray counters, exit count; zero between launches).
constexpr size_t kBatchCounterBytes = 10 * sizeof(unsigned long long);

struct BatchLaunch {
  const BatchTable* table;
  int nb;  // host copy, for argument checks only
  unsigned long long* counters;  // device u64[10], zero between launches (self-resetting)
  int32_t* spill;
  int64_t spill_stride;  // blocks * block
  int blocks;
  int block;      // 64, 128 or 256 lanes
  int stack_lds;  // 8, 16 or 32
  int layout;     // kLayoutPair, kLayoutRef or kLayoutSoa
  int refill;     // idle lanes that trigger a refill, 1..63; 64 = whole-wave chunks
  int chunk;      // rays per counter grab (multiple of 64); 0 = static stride, no counter
  int xcd_remap;  // contiguous ray shares per XCD
  int reverse;    // hand chunks out from the end of the table
  int kinds;      // 0 mixed table, 1 closest-hit batches only, 2 any-hit batches only
};

int launch_trace_batch(const DevScene& sc, const BatchLaunch& bl, void* stream);

// Launchers (trace.hip). Return hipError_t as int.
int launch_trace(const DevScene& sc, const DevRays& rays, const DevHits& hits, int64_t n,
                 const LaunchCfg& cfg, void* stream);

struct DevRefit {
  RefNode* refs;
  PairNode* pairs;
  float* tris;
  const int32_t* obj_ids;
  const float* verts;
  const uint32_t* idx;
  const int32_t* leaves;  // leaf record ids
  const int32_t* levels;  // inner record ids grouped by depth (level_off gives the groups)
  int32_t n_leaves;
  int32_t n_slots;        // T
  int32_t prim_kind;
};

// rotate.hip: DynBVH::update with tree rotations and leaf splits.
constexpr int kMaxDepth = 256;  // deepest tree the level lists hold

// The top depths of a tree that one workgroup runs in one launch, with a barrier between
// depths (refit and update): depths [0, n), each of at most kTopNodes records, depth d at
// [off[d], off[d + 1]) of the level list.
constexpr int kTopLevels = 32;
constexpr int kTopNodes = 512;
struct TopLevels {
  int32_t off[kTopLevels + 1];
  int32_t n;
};
// The top depths of level_off (n_levels + 1 host offsets) that fit: fills t, returns t.n.
inline int top_levels(const int64_t* level_off, int n_levels, TopLevels& t) {
  int n = 0;
  while (n < n_levels && n < kTopLevels && level_off[n + 1] - level_off[n] <= kTopNodes) ++n;
  for (int d = 0; d <= n; ++d) t.off[d] = (int32_t)level_off[d];
  t.n = n;
  return n;
}

struct DevUpdate {
  RefNode* refs;
  PairNode* pairs;
  int32_t* big_count;
  float* costs;          // DynBVH::costs (Real), per record
  uint32_t* ssize;       // DynBVH::subtree_size, per record
  int32_t* parent;       // parent links of records built by leaf splits
  int32_t* ids;          // object_ids (leaf splits permute ranges)
  float* obox;           // per object: lo xyz, hi xyz (null: formed from verts and idx)
  float* ocen;           // per object: centroid xyz (null: formed from the object box)
  const float* verts;
  const uint32_t* idx;
  unsigned* next_free;   // records in use (nextFree)
  int* err;              // set when the record array is full
  int32_t* split_list;   // update: leaves whose parent's rotation may split them
  unsigned* split_count; // update: entries of split_list (cleared by the host)
  int32_t* leaf_at;      // update: per slot, the leaf whose first slot it is, else -1
  int32_t n_records;     // capacity, 2T
  int32_t n_objects;     // T
  uint32_t large;        // largeSubtreeSize
};
int launch_obj_bounds(const DevUpdate& u, void* stream);
// u.leaf_at from a leaf list (the caller sets every entry to -1 first).
int launch_leaf_starts(const DevUpdate& u, const int32_t* leaves, int64_t n_leaves,
                       void* stream);
// Leaf pass over the object slots (u.leaf_at), leaf splits, then one launch per depth
// (deepest first).
int launch_update_tree(const DevUpdate& u, int64_t n_leaves, const int32_t* levels,
                       const int64_t* level_off, int n_levels, void* stream);
// PairNodes and big leaf counts of every record pair in use, from the reference records.
int launch_rebuild_pairs(const DevUpdate& u, void* stream);
// Depth of every record in use by `steps` pointer-jumping steps (even, at most 8; 8 covers
// any depth below kMaxDepth); hist[d] inner records at depth d, hist[kMaxDepth] leaves,
// hist[kMaxDepth + 1] the deepest depth, hist[kMaxDepth + 2] nonzero when the steps did
// not reach the root (hist, kDepthHist entries, cleared by the caller). Depths in dep0.
constexpr int kDepthHist = kMaxDepth + 3;
int launch_depths(const DevUpdate& u, int32_t* anc0, int32_t* dep0, int32_t* anc1, int32_t* dep1,
                  unsigned* hist, int steps, void* stream);
int launch_depth_scatter(const DevUpdate& u, const int32_t* dep, unsigned* cursor,
                         int32_t* levels, int32_t* leaves, void* stream);
// Level lists and PairNodes of an updated tree, top down from an inner root: counters fr
// (kFrontierWords: inner records per depth, leaves at [kMaxDepth], an overflow flag at
// [kMaxDepth + 1], list offsets from [kMaxDepth + 2]). launch_frontier_top clears them and
// runs depths [0, n_top) (1 <= n_top <= kTopLevels) in one workgroup; launch_frontier(d)
// then lists depth d + 1 (d < kMaxDepth - 1) and the leaf children of depth d, in
// `blocks` workgroups. Each list holds at most cap entries; aux (cap entries) holds each
// level entry's first child and record word for the next depth.
constexpr int kFrontierWords = 2 * kMaxDepth + 3;
int launch_frontier_top(const DevUpdate& u, unsigned* fr, int32_t* levels, int2* aux,
                        int32_t* leaves, unsigned cap, int n_top, void* stream);
int launch_frontier(const DevUpdate& u, unsigned* fr, int32_t* levels, int2* aux,
                    int32_t* leaves, unsigned cap, int d, unsigned blocks, void* stream);

// refit.hip
int launch_prim_setup(const DevRefit& r, void* stream);  // tris from verts (all slots)
// tris + leaf boxes, then one launch per level of inner records from the deepest up;
// level_off (host, n_levels + 1 entries) delimits each depth's records in r.levels.
int launch_refit(const DevRefit& r, const int64_t* level_off, int n_levels, void* stream);
int launch_displace_sine(float* verts, const float* base, int64_t n_verts, float amp, float freq,
                         float phase, void* stream);

// genrays.hip: camera rays and secondary rays from hit records.
struct RayOut {
  float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  uint8_t* masked;  // may be null for camera rays
};
struct CameraDev {
  float eye[3], U[3], V[3], W[3];  // up * tan(vfov/2) * |W|, right * tan(hfov/2) * |W|, lookat-eye
  int xres, yres, spp;             // xres, yres multiples of 8
  unsigned long long seed;
};
struct SecondaryArgs {
  DevRays in;            // the primary rays
  const int32_t* hit_prim;
  const float* hit_t;
  const float* verts;
  const uint32_t* idx;
  float light[3];
  unsigned long long seed;
  int64_t n;
  RayOut out;
};
int launch_camera(const CameraDev& c, const RayOut& out, void* stream);
int launch_secondary(const SecondaryArgs& a, int kind /*0 shadow, 1 bounce*/, void* stream);

// util.hip: streaming copy with 16-byte vector accesses (plus a byte tail).
int launch_copy(void* dst, const void* src, size_t bytes, int cu_count, void* stream);


// Device build (build.hip): DynBVH::rebuild with the approximate SAH at every node.
constexpr int16_t kHole = -1;     // children of a temporary record no node uses
constexpr int kBuildSerial = 64;  // nodes of at most this many objects: one lane each
constexpr int kBuildVB = 2048;     // slots per virtual block of the level passes
constexpr int kBuildBins = 16;     // BVH_num_samples (DynBVH.cc:34)

// Accumulators of one level node, stored field-major ([field][node]) so the per-level
// clear streams.
enum : int {
  kFcMin = 0,                    // 3 centroid min keys
  kFcMax = 3,                    // 3 centroid max keys
  kFbMin = 6,                    // 16 x 3 bin min keys
  kFbMax = kFbMin + 3 * kBuildBins,   // 16 x 3 bin max keys
  kFlMin = kFbMax + 3 * kBuildBins,   // 3 left-side min keys
  kFlMax = kFlMin + 3,
  kFrMin = kFlMax + 3,
  kFrMax = kFrMin + 3,
  kKeyFields = kFrMax + 3,
};
enum : int { kFcnt = 0, kFless = kBuildBins, kCntFields = kBuildBins + 1 };

struct BuildSeg {  // a node of the current level: its temporary record and slot range
  int32_t rec, b, e, vb0;
};
struct BuildPlan {
  int32_t axis, leaf, split, nless;
  float lo_c, scale, pos, width;
};

enum : int { kCtrNext = 0, kCtrSmall = 1, kCtrBig = 2, kCtrNvb = 3, kCtrCount = 4 };

struct DevBuild {
  DevUpdate u;                    // temporary records, ids, object boxes, counters
  float* slots;                   // 12 floats per slot: centroid, id, box
  const float* verts;
  int64_t n_verts;
  BuildSeg* seg;
  BuildSeg* seg_next;
  BuildSeg* small;
  BuildSeg* big;
  BuildPlan* plan;
  unsigned long long* acck;       // [kKeyFields][cap]
  unsigned* accc;                 // [kCntFields][cap]
  int64_t cap;                    // capacity of the node lists
  unsigned* ctr;                  // kCtr*
  int32_t* vbseg;                 // per virtual block: its node
  unsigned* vcnt;                 // per virtual block: left, right misplaced counts
  unsigned* gscan;                // exclusive scans of vcnt (left at [0, nvb], right after)
  int64_t vcap;                   // capacity of the virtual block arrays
  int32_t* lpos;
  int32_t* rpos;
  unsigned long long* rootk;      // 6 keys: vertex box
  int32_t nseg;
};


struct DevFinish {
  const RefNode* trefs;    // temporary records
  RefNode* refs;           // final records
  const unsigned* tcount;  // records in use
  int32_t* ic;             // inner records per temporary subtree
  int32_t* rank;           // preorder rank of a temporary record among inner records
  int32_t* fin;            // final number of a temporary record
  int32_t* levels;
  int32_t* leaves;
  float* costs;
  uint32_t* ssize;
  float* rt;               // RTSAH pairs (c0, c1) per final record
  const float* logtab;
  int32_t n_records;
  int32_t n_final;         // records of the tree (inner + leaves)
  uint32_t large;
  int rtsah;
};

int build_launch_setup(const DevBuild& d, int serial, void* stream);
int build_launch_level(const DevBuild& d, int64_t nvb_bound, void* stream);
int build_launch_serial(const DevBuild& d, int64_t n_small, int64_t n_big, int32_t* scratch,
                        void* stream);
int build_launch_ranks(const DevFinish& f, const int64_t* level_off, int n_levels,
                       int64_t n_leaves, void* stream);

// shade.hip: hit positions, normals and texture coordinates of traced rays.
struct DevShade {
  DevRays rays;
  const int32_t* prim;
  const float *t, *u, *v;
  const float* verts;
  const uint32_t* idx;
  const float* normals;        // null: face normals
  const uint32_t* normal_idx;
  const float* tex;            // null: barycentric weights
  const uint32_t* tex_idx;
  float* out[12];              // px py pz, nx ny nz, gx gy gz, tu tv tw (each may be null)
  int64_t n;
  int32_t n_tris;
  int32_t prim_kind;
};
int launch_shade(const DevShade& s, void* stream);

}  // namespace mh
