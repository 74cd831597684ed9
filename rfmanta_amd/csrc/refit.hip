// Per-frame refit kernels for gfx950.
//
// The refit is DynBVH::updateBounds<false> without rotations (DynBVH.cc:2084-2117), done
// level-synchronously:
//   k_refit_leaves: one lane per leaf record rebuilds the triangle records of its leaf
//     from the current vertices and forms the leaf box (Mesh::getBBox +
//     BBox::extendByBox, Model/Groups/Mesh.cc:228-237, DynBVH.cc:2091-2097);
//   k_refit_level: one launch per depth, deepest first; each lane unions the two child
//     boxes of one inner record (DynBVH.cc:2110-2116);
//   k_refit_top: the small depths at the top of the tree (at most kTopNodes records each)
//     in one workgroup, deepest first, with a barrier between depths.
// Every box is stored in the reference record and in its PairNode. Kernel boundaries on
// the stream order the levels, so no cross-workgroup synchronisation is
// needed (agent-scope acquire/release would write back and invalidate the per-XCD L2s).
// Box arithmetic is exact (min/max of the same fp32 values), so the result is
// bit-identical to a serial recomputation.

#include <hip/hip_runtime.h>

#include <cfloat>

#include "mh_device.h"
#include "prim_setup.h"

namespace mh {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float pick_lo(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float pick_hi(float a, float b) { return a > b ? a : b; }


__device__ __forceinline__ void write_record(const DevRefit& r, int slot, const Tri& t) {
  store_tri_record(r.tris, r.prim_kind, slot, t);
}

__global__ __launch_bounds__(kBlock) void k_prim_setup(DevRefit r) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= r.n_slots) return;
  write_record(r, (int)s, load_tri(r.verts, r.idx, r.obj_ids[s]));
}

__device__ __forceinline__ void store_box(const DevRefit& r, int rec, const float lo[3],
                                          const float hi[3]) {
  RefNode* nd = r.refs + rec;
  nd->box.lo[0] = lo[0];
  nd->box.lo[1] = lo[1];
  nd->box.lo[2] = lo[2];
  nd->box.hi[0] = hi[0];
  nd->box.hi[1] = hi[1];
  nd->box.hi[2] = hi[2];
  if (rec > 0) {  // record rec sits in PairNode (rec - 1) / 2 on side (rec - 1) & 1
    float* b = r.pairs[(rec - 1) >> 1].b + 6 * ((rec - 1) & 1);
    b[0] = lo[0];
    b[1] = hi[0];
    b[2] = lo[1];
    b[3] = hi[1];
    b[4] = lo[2];
    b[5] = hi[2];
  }
}

__global__ __launch_bounds__(kBlock) void k_refit_leaves(DevRefit r) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= r.n_leaves) return;
  const int rec = r.leaves[i];
  const RefNode leaf = r.refs[rec];
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int k = 0; k < leaf.children; ++k) {
    const int slot = leaf.child + k;
    const Tri t = load_tri(r.verts, r.idx, r.obj_ids[slot]);
    write_record(r, slot, t);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float tlo = pick_lo(pick_lo(t.p[0][a], t.p[1][a]), t.p[2][a]);
      const float thi = pick_hi(pick_hi(t.p[0][a], t.p[1][a]), t.p[2][a]);
      lo[a] = pick_lo(lo[a], tlo);
      hi[a] = pick_hi(hi[a], thi);
    }
  }
  store_box(r, rec, lo, hi);
}

// Inner record rec: the union of its two child boxes.
__device__ __forceinline__ void refit_node(const DevRefit& r, int rec) {
  const int c = r.refs[rec].child;
  const RefNode a = r.refs[c];
  const RefNode b = r.refs[c + 1];
  float lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = pick_lo(pick_lo(FLT_MAX, a.box.lo[k]), b.box.lo[k]);
    hi[k] = pick_hi(pick_hi(-FLT_MAX, a.box.hi[k]), b.box.hi[k]);
  }
  store_box(r, rec, lo, hi);
}

__global__ __launch_bounds__(kBlock) void k_refit_level(DevRefit r, int64_t first, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  refit_node(r, r.levels[first + i]);
}

// The top depths [0, t.n), deepest first, in one workgroup with a barrier between depths
// (its waves share one compute unit and its L1, so the boxes stored before the barrier
// are seen by the loads after it).
__global__ __launch_bounds__(kBlock) void k_refit_top(DevRefit r, TopLevels t) {
  for (int d = t.n - 1; d >= 0; --d) {
    for (int i = t.off[d] + (int)threadIdx.x; i < t.off[d + 1]; i += kBlock)
      refit_node(r, r.levels[i]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_displace_sine(float* verts, const float* base,
                                                          int64_t n, float amp, float freq,
                                                          float phase) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float x = base[3 * i + 0];
  verts[3 * i + 0] = x;
  verts[3 * i + 1] = base[3 * i + 1] + amp * sinf(freq * x + phase);
  verts[3 * i + 2] = base[3 * i + 2];
}

}  // namespace

int launch_prim_setup(const DevRefit& r, void* stream) {
  if (r.n_slots <= 0) return 0;
  const unsigned blocks = (unsigned)((r.n_slots + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_prim_setup, dim3(blocks), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), r);
  return (int)hipGetLastError();
}

int launch_refit(const DevRefit& r, const int64_t* level_off, int n_levels, void* stream) {
  if (r.n_leaves <= 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  unsigned blocks = (unsigned)((r.n_leaves + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_refit_leaves, dim3(blocks), dim3(kBlock), 0, s, r);
  TopLevels t;
  const int top = top_levels(level_off, n_levels, t);
  for (int d = n_levels - 1; d >= top; --d) {
    const int64_t first = level_off[d], n = level_off[d + 1] - level_off[d];
    if (n <= 0) continue;
    blocks = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_refit_level, dim3(blocks), dim3(kBlock), 0, s, r, first, n);
  }
  if (top > 0) hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(kBlock), 0, s, r, t);
  return (int)hipGetLastError();
}

int launch_displace_sine(float* verts, const float* base, int64_t n_verts, float amp, float freq,
                         float phase, void* stream) {
  if (n_verts <= 0) return 0;
  const unsigned blocks = (unsigned)((n_verts + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_displace_sine, dim3(blocks), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), verts, base, n_verts, amp, freq,
                     phase);
  return (int)hipGetLastError();
}

}  // namespace mh
