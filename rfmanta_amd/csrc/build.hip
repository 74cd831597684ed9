// Device build for gfx950: DynBVH::rebuild (Model/Groups/DynBVH.cc:1699-1793) as the
// reference builds it with USE_APPROXIMATE_BUILD set (its MANTA_USE_DYNBVH_APPROXIMATE
// option, include/CMakeLists.txt:134-147, DynBVH.cc:1815-1822): every node is partitioned
// by partitionApproxSAH (16 bins over the longest centroid axis), leaves are sorted, then
// computeSubTreeSizes, the RTSAH pass and computeCost<true>. The tree is the reference's
// record for record, numbering included.
//
// Phase 1, level-synchronous, for nodes of more than kBuildSerial objects. The objects of a
// node are a contiguous range of slots; each slot carries its object's centroid, id and
// box (48 bytes, moved with the id), so every pass streams. A level's nodes are cut into
// virtual blocks of kVB slots; per level:
//   k_cbounds      centroid box of each node          (block reduce, then global atomics)
//   k_seg_axis     longest centroid axis, bin scale   (one lane per node)
//   k_bins         16 bins: counts and boxes
//   k_seg_split    the binned SAH sweep: split position or leaf
//   k_count_less   objects left of the position
//   k_seg_mid      split slot, or the middle when everything fell on one side
//   k_mis_count, k_scan_mis, k_mis_scatter, k_swap   partitionObjects
//   k_child_bounds boxes of the two sides
//   k_seg_children new record pair, child nodes to the next level or to phase 2
// Reductions must give what the reference's serial loops give. Its min/max are select
// forms (a < b ? a : b) applied in slot order, so of equal values (+0 and -0) the last one
// wins; each value is reduced as a 64-bit key (order-preserving value bits, then the slot
// number) with integer atomics, which makes the result exact and independent of order.
// partitionObjects (Hoare's scheme) exchanges the k-th slot from the left whose key is on
// the right side with the k-th slot from the right whose key is on the left side; the
// ranks come from a scan, and the exchanges are independent.
// Phase 2: one lane per node of at most kBuildSerial objects runs the reference's recursion
// (build_loop). Phase 3 renumbers the records in the reference's allocation order
// (pairs in preorder of their parents) and computes sizes, marks, costs and RTSAH bottom up.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "approx_sah.h"
#include "mh_device.h"
#include "rtsah.h"

namespace mh {
namespace {

constexpr int kBlock = 256;
constexpr int kPer = kBuildVB / kBlock;
constexpr int kVB = kBuildVB;
constexpr int kBins = kSamples;
static_assert(kBins == kBuildBins, "bins");
constexpr unsigned long long kMinInit = ~0ull;

// 32-bit key that orders floats as numbers, with -0 and +0 equal.
__device__ __forceinline__ uint32_t ord(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u << 1) == 0) u = 0;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t e) {
  return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}
constexpr uint32_t kOrdZero = 0x80000000u;
// min: smallest value, then the largest slot; max: largest value, then the largest slot.
__device__ __forceinline__ unsigned long long kmin(float v, uint32_t slot) {
  return ((unsigned long long)ord(v) << 32) | (0xFFFFFFFFu - slot);
}
__device__ __forceinline__ unsigned long long kmax(float v, uint32_t slot) {
  return ((unsigned long long)ord(v) << 32) | slot;
}
__device__ __forceinline__ uint32_t kmin_slot(unsigned long long k) {
  return 0xFFFFFFFFu - (uint32_t)k;
}
__device__ __forceinline__ uint32_t kmax_slot(unsigned long long k) { return (uint32_t)k; }

// Slot record: centroid, object id, box.
struct alignas(16) Slot {
  float c[3];
  int32_t id;
  float lo[3];
  float hi[3];
  float pad[2];
};
static_assert(sizeof(Slot) == 48, "slot record is 48 bytes");
__device__ __forceinline__ Slot* SL(const DevBuild& d) { return reinterpret_cast<Slot*>(d.slots); }

__device__ __forceinline__ void load_slot(const Slot* s, int64_t i, Slot& o) {
  const float4* p = reinterpret_cast<const float4*>(s + i);
  const float4 a = p[0], b = p[1], c = p[2];
  o.c[0] = a.x;
  o.c[1] = a.y;
  o.c[2] = a.z;
  o.id = __float_as_int(a.w);
  o.lo[0] = b.x;
  o.lo[1] = b.y;
  o.lo[2] = b.z;
  o.hi[0] = b.w;
  o.hi[1] = c.x;
  o.hi[2] = c.y;
  o.pad[0] = c.z;
  o.pad[1] = c.w;
}
__device__ __forceinline__ float slot_c(const Slot* s, int64_t i, int a) {
  return reinterpret_cast<const float*>(s + i)[a];
}

// The level node that owns virtual block v.
__device__ __forceinline__ int find_seg(const DevBuild& d, int64_t v) { return d.vbseg[v]; }

struct VRange {
  int s;
  int64_t first, end;  // slots of this virtual block
};
__device__ __forceinline__ bool vrange(const DevBuild& d, VRange& r) {
  const int64_t v = blockIdx.x;
  if (v >= (int64_t)d.ctr[kCtrNvb]) return false;
  r.s = find_seg(d, v);
  const BuildSeg sg = d.seg[r.s];
  r.first = sg.b + (v - sg.vb0) * (int64_t)kVB;
  r.end = r.first + kVB < sg.e ? r.first + kVB : sg.e;
  return true;
}

__device__ __forceinline__ unsigned long long& K(const DevBuild& d, int f, int s) {
  return d.acck[(int64_t)f * d.cap + s];
}
__device__ __forceinline__ unsigned& Cn(const DevBuild& d, int f, int s) {
  return d.accc[(int64_t)f * d.cap + s];
}

// ---------------------------------------------------------------------------------------
// Set-up

// Box of all vertices (Mesh::computeBounds, Model/Groups/Mesh.cc:195-225); a non-finite
// coordinate sets err = 2.
__global__ __launch_bounds__(kBlock) void k_root_box(DevBuild d) {
  __shared__ unsigned long long sk[6];
  if (threadIdx.x < 6) sk[threadIdx.x] = threadIdx.x < 3 ? kMinInit : 0ull;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < d.n_verts) {
    for (int a = 0; a < 3; ++a) {
      const float x = d.verts[3 * i + a];
      if (!(x >= -FLT_MAX && x <= FLT_MAX)) atomicExch(d.u.err, 2);
      atomicMin(&sk[a], kmin(x, (uint32_t)i));
      atomicMax(&sk[3 + a], kmax(x, (uint32_t)i));
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&d.rootk[threadIdx.x], sk[threadIdx.x]);
  if (threadIdx.x >= 3 && threadIdx.x < 6) atomicMax(&d.rootk[threadIdx.x], sk[threadIdx.x]);
}

// Slot records from the object boxes and centroids (in object order, as object_ids
// starts out).
__global__ __launch_bounds__(kBlock) void k_slots(DevBuild d) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.u.n_objects) return;
  const float* b = d.u.obox + 6 * i;
  const float* c = d.u.ocen + 3 * i;
  float4* p = reinterpret_cast<float4*>(SL(d) + i);
  p[0] = make_float4(c[0], c[1], c[2], __int_as_float((int)i));
  p[1] = make_float4(b[0], b[1], b[2], b[3]);
  p[2] = make_float4(b[4], b[5], 0.f, 0.f);
}

// Record 0 gets the vertex box; the root becomes the first node of phase 1 or 2.
__global__ void k_root_init(DevBuild d, int serial) {
  RefNode& r = d.u.refs[0];
  for (int a = 0; a < 3; ++a) {
    const unsigned long long kl = d.rootk[a], kh = d.rootk[3 + a];
    r.box.lo[a] = (uint32_t)(kl >> 32) == kOrdZero ? d.verts[3 * (int64_t)kmin_slot(kl) + a]
                                                   : unord((uint32_t)(kl >> 32));
    r.box.hi[a] = (uint32_t)(kh >> 32) == kOrdZero ? d.verts[3 * (int64_t)kmax_slot(kh) + a]
                                                   : unord((uint32_t)(kh >> 32));
  }
  r.child = 0;
  r.children = 0;
  r.axis = 0;
  r.bits = 0;
  *d.u.next_free = 1;
  const BuildSeg root{0, 0, d.u.n_objects, 0};
  if (serial) {
    d.small[0] = root;
    d.ctr[kCtrSmall] = 1;
  } else {
    d.seg[0] = root;
  }
}

// ---------------------------------------------------------------------------------------
// Phase 1: one level

__global__ __launch_bounds__(kBlock) void k_acc_clear(DevBuild d) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nk = (int64_t)kKeyFields * d.nseg;
  if (i < nk) {
    const int f = (int)(i / d.nseg), s = (int)(i % d.nseg);
    const bool is_min = (f >= kFcMin && f < kFcMax) || (f >= kFbMin && f < kFbMax) ||
                        (f >= kFlMin && f < kFlMax) || (f >= kFrMin && f < kFrMax);
    K(d, f, s) = is_min ? kMinInit : 0ull;
  } else if (i < nk + (int64_t)kCntFields * d.nseg) {
    const int64_t j = i - nk;
    Cn(d, (int)(j / d.nseg), (int)(j % d.nseg)) = 0u;
  }
}

// Exclusive prefix of the nodes' virtual block counts (one workgroup).
__global__ __launch_bounds__(1024) void k_scan_segments(DevBuild d) {
  __shared__ unsigned part[1024];
  __shared__ unsigned carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < d.nseg; base += 1024) {
    const int64_t s = base + threadIdx.x;
    const unsigned n = s < d.nseg ? (unsigned)((d.seg[s].e - d.seg[s].b + kVB - 1) / kVB) : 0u;
    part[threadIdx.x] = n;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan (Hillis-Steele)
      const unsigned x = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += x;
      __syncthreads();
    }
    if (s < d.nseg) d.seg[s].vb0 = (int32_t)(carry + part[threadIdx.x] - n);
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.ctr[kCtrNvb] = carry;
}

// vbseg[v] = the node of virtual block v, for every virtual block of the level.
__global__ __launch_bounds__(kBlock) void k_vb_map(DevBuild d) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= d.nseg) return;
  const BuildSeg sg = d.seg[s];
  const int nvb = (sg.e - sg.b + kVB - 1) / kVB;
  for (int k = 0; k < nvb; ++k) d.vbseg[sg.vb0 + k] = s;
}

__global__ __launch_bounds__(kBlock) void k_cbounds(DevBuild d) {
  __shared__ unsigned long long sk[6];
  VRange r;
  if (!vrange(d, r)) return;
  if (threadIdx.x < 6) sk[threadIdx.x] = threadIdx.x < 3 ? kMinInit : 0ull;
  __syncthreads();
  unsigned long long mn_[3] = {kMinInit, kMinInit, kMinInit}, mx_[3] = {0ull, 0ull, 0ull};
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = r.first + k * kBlock + threadIdx.x;
    if (i >= r.end) break;
    for (int a = 0; a < 3; ++a) {
      const float c = slot_c(SL(d), i, a);
      const unsigned long long x = kmin(c, (uint32_t)i), y = kmax(c, (uint32_t)i);
      mn_[a] = x < mn_[a] ? x : mn_[a];
      mx_[a] = y > mx_[a] ? y : mx_[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    atomicMin(&sk[a], mn_[a]);
    atomicMax(&sk[3 + a], mx_[a]);
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&K(d, kFcMin + threadIdx.x, r.s), sk[threadIdx.x]);
  if (threadIdx.x >= 3 && threadIdx.x < 6)
    atomicMax(&K(d, kFcMax + threadIdx.x - 3, r.s), sk[threadIdx.x]);
}

// Value of a reduced key; a zero is read back from the winning slot for its sign.
__device__ __forceinline__ float key_value(unsigned long long k, bool is_min, const Slot* slots,
                                           int field) {
  const uint32_t o = (uint32_t)(k >> 32);
  if (o != kOrdZero) return unord(o);
  const uint32_t slot = is_min ? kmin_slot(k) : kmax_slot(k);
  return reinterpret_cast<const float*>(slots + slot)[field];
}

__global__ __launch_bounds__(kBlock) void k_seg_axis(DevBuild d) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= d.nseg) return;
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = key_value(K(d, kFcMin + a, s), true, SL(d), a);
    hi[a] = key_value(K(d, kFcMax + a, s), false, SL(d), a);
  }
  BuildPlan p{};
  const float w0 = hi[0] - lo[0], w1 = hi[1] - lo[1], w2 = hi[2] - lo[2];
  p.axis = index_of_max(w0, w1, w2);  // BBox::longestAxis
  const float w = p.axis == 0 ? w0 : (p.axis == 1 ? w1 : w2);
  p.width = w;
  p.lo_c = lo[p.axis];
  p.scale = ((float)kBins / w) * .999f;
  d.plan[s] = p;
}

__device__ __forceinline__ int bin_of(float c, float lo_c, float scale) {
  const float f = (c - lo_c) * scale;
  // int(f) as the reference's x86 build converts it (cvttss2si): NaN and values outside
  // the int range give INT_MIN, i.e. bin 0 after the clamp.
  int which = (f >= -2147483648.f && f < 2147483648.f) ? (int)f : INT_MIN;
  return which < 0 ? 0 : (which > kBins - 1 ? kBins - 1 : which);
}

__global__ __launch_bounds__(kBlock) void k_bins(DevBuild d) {
  __shared__ unsigned cnt[kBins];
  __shared__ unsigned long long bmn[kBins * 3], bmx[kBins * 3];
  VRange r;
  if (!vrange(d, r)) return;
  for (int k = threadIdx.x; k < kBins * 3; k += kBlock) {
    bmn[k] = kMinInit;
    bmx[k] = 0ull;
  }
  if (threadIdx.x < kBins) cnt[threadIdx.x] = 0;
  __syncthreads();
  const BuildPlan p = d.plan[r.s];
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = r.first + k * kBlock + threadIdx.x;
    if (i >= r.end) break;
    Slot o;
    load_slot(SL(d), i, o);
    const int w = bin_of(o.c[p.axis], p.lo_c, p.scale);
    atomicAdd(&cnt[w], 1u);
    for (int a = 0; a < 3; ++a) {
      atomicMin(&bmn[3 * w + a], kmin(o.lo[a], (uint32_t)i));
      atomicMax(&bmx[3 * w + a], kmax(o.hi[a], (uint32_t)i));
    }
  }
  __syncthreads();
  if (threadIdx.x < kBins * 3) {
    const int w = threadIdx.x / 3;
    if (cnt[w]) {
      atomicMin(&K(d, kFbMin + threadIdx.x, r.s), bmn[threadIdx.x]);
      atomicMax(&K(d, kFbMax + threadIdx.x, r.s), bmx[threadIdx.x]);
      if (threadIdx.x % 3 == 0) atomicAdd(&Cn(d, kFcnt + w, r.s), cnt[w]);
    }
  }
}

// The sweep of partitionApproxSAH (DynBVH.cc:2196-2236) over the reduced bins.
__global__ __launch_bounds__(kBlock) void k_seg_split(DevBuild d) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= d.nseg) return;
  BuildPlan p = d.plan[s];
  const BuildSeg sg = d.seg[s];
  const int n = sg.e - sg.b;
  Box bins[kBins];
  int counts[kBins];
  for (int w = 0; w < kBins; ++w) {
    counts[w] = (int)Cn(d, kFcnt + w, s);
    if (counts[w] == 0) {
      reset(bins[w]);
      continue;
    }
    for (int a = 0; a < 3; ++a) {
      bins[w].lo[a] = key_value(K(d, kFbMin + 3 * w + a, s), true, SL(d), 4 + a);
      bins[w].hi[a] = key_value(K(d, kFbMax + 3 * w + a, s), false, SL(d), 7 + a);
    }
  }
  const float inv_parent_area = (float)(1.0 / (double)area(rbox(d.u.refs[sg.rec])));
  float best_cost = n * kCIsec;
  int best_axis = -1;
  float best_pos = FLT_MAX;
  float left_areas[kBins], right_areas[kBins];
  int left_counts[kBins], right_counts[kBins];
  Box lb, rb;
  reset(lb);
  reset(rb);
  int nl = 0, nr = n;
  for (int i = 0; i < kBins; ++i) {
    nl += counts[i];
    nr -= counts[i];
    extend(lb, bins[i]);
    left_counts[i] = nl;
    right_counts[i] = nr;
    left_areas[i] = area(lb);
  }
  for (int i = kBins - 1; i >= 0; --i) {
    right_areas[i] = (i == kBins - 1) ? FLT_MAX : area(rb);
    extend(rb, bins[i]);
  }
  for (int i = 0; i < kBins - 1; ++i) {
    const float cost = (left_areas[i] * left_counts[i] + right_areas[i] * right_counts[i]) *
                           inv_parent_area * kCIsec + kCTrav;
    if (cost < best_cost) {
      best_cost = cost;
      best_axis = p.axis;
      best_pos = p.lo_c + p.width * (static_cast<float>(i + 1) / kBins);
    }
  }
  p.leaf = best_axis == -1;
  p.pos = best_pos;
  d.plan[s] = p;
}

__global__ __launch_bounds__(kBlock) void k_count_less(DevBuild d) {
  __shared__ unsigned tot;
  VRange r;
  if (!vrange(d, r)) return;
  const BuildPlan p = d.plan[r.s];
  if (p.leaf) return;  // uniform over the block
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  unsigned c = 0;
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = r.first + k * kBlock + threadIdx.x;
    if (i >= r.end) break;
    c += slot_c(SL(d), i, p.axis) < p.pos ? 1u : 0u;
  }
  if (c) atomicAdd(&tot, c);
  __syncthreads();
  if (threadIdx.x == 0 && tot) atomicAdd(&Cn(d, kFless, r.s), tot);
}

// partitionObjects returns the number of objects left of the position; when that is 0
// or all of them the reference splits at the middle (DynBVH.cc:2250-2261).
__global__ __launch_bounds__(kBlock) void k_seg_mid(DevBuild d) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= d.nseg) return;
  BuildPlan p = d.plan[s];
  if (p.leaf) return;
  const BuildSeg sg = d.seg[s];
  const int n = sg.e - sg.b;
  const int m = (int)Cn(d, kFless, s);
  p.nless = m;
  p.split = (m == 0 || m == n) ? sg.b + n / 2 : sg.b + m;
  d.plan[s] = p;
}

// A slot is misplaced on the left if it lies before the split with a key >= the position,
// on the right if it lies at or after the split with a key < the position. No slot moves
// in the middle fallback.
__global__ __launch_bounds__(kBlock) void k_mis_count(DevBuild d) {
  __shared__ unsigned tl, tr;
  VRange r;
  if (!vrange(d, r)) return;
  const int64_t v = blockIdx.x;
  const BuildPlan p = d.plan[r.s];
  const BuildSeg sg = d.seg[r.s];
  const int n = sg.e - sg.b;
  const bool active = !p.leaf && p.nless != 0 && p.nless != n;
  if (threadIdx.x == 0) tl = tr = 0;
  __syncthreads();
  if (active) {
    unsigned cl = 0, cr = 0;
    for (int k = 0; k < kPer; ++k) {
      const int64_t i = r.first + k * kBlock + threadIdx.x;
      if (i >= r.end) break;
      const float c = slot_c(SL(d), i, p.axis);
      cl += (i < p.split && !(c < p.pos)) ? 1u : 0u;
      cr += (i >= p.split && c < p.pos) ? 1u : 0u;
    }
    if (cl) atomicAdd(&tl, cl);
    if (cr) atomicAdd(&tr, cr);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    d.vcnt[2 * v] = tl;
    d.vcnt[2 * v + 1] = tr;
  }
}

// Exclusive scans over all virtual blocks of the left and right counts (one workgroup):
// gscan[v] (left) and gscan[vcap + 1 + v] (right), v in [0, nvb].
__global__ __launch_bounds__(1024) void k_scan_mis(DevBuild d) {
  __shared__ unsigned pl[1024], pr[1024];
  __shared__ unsigned cl, cr;
  const int64_t nvb = d.ctr[kCtrNvb];
  unsigned* gl = d.gscan;
  unsigned* gr = d.gscan + d.vcap + 1;
  if (threadIdx.x == 0) cl = cr = 0;
  __syncthreads();
  for (int64_t base = 0; base < nvb; base += 1024) {
    const int64_t v = base + threadIdx.x;
    const unsigned a = v < nvb ? d.vcnt[2 * v] : 0u, b = v < nvb ? d.vcnt[2 * v + 1] : 0u;
    pl[threadIdx.x] = a;
    pr[threadIdx.x] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const unsigned x = threadIdx.x >= (unsigned)off ? pl[threadIdx.x - off] : 0u;
      const unsigned y = threadIdx.x >= (unsigned)off ? pr[threadIdx.x - off] : 0u;
      __syncthreads();
      pl[threadIdx.x] += x;
      pr[threadIdx.x] += y;
      __syncthreads();
    }
    if (v < nvb) {
      gl[v] = cl + pl[threadIdx.x] - a;
      gr[v] = cr + pr[threadIdx.x] - b;
    }
    __syncthreads();
    if (threadIdx.x == 1023) {
      cl += pl[1023];
      cr += pr[1023];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    gl[nvb] = cl;
    gr[nvb] = cr;
  }
}

// Block-wide exclusive scan of one value per thread in slot order (threads own kPer
// consecutive slots here); returns the thread's offset.
__device__ __forceinline__ unsigned block_exscan(unsigned x, unsigned* buf) {
  buf[threadIdx.x] = x;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const unsigned y = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0u;
    __syncthreads();
    buf[threadIdx.x] += y;
    __syncthreads();
  }
  const unsigned r = buf[threadIdx.x] - x;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kBlock) void k_mis_scatter(DevBuild d) {
  __shared__ unsigned buf[kBlock];
  VRange r;
  if (!vrange(d, r)) return;
  const int64_t v = blockIdx.x;
  const BuildPlan p = d.plan[r.s];
  const BuildSeg sg = d.seg[r.s];
  const int n = sg.e - sg.b;
  if (p.leaf || p.nless == 0 || p.nless == n) return;  // uniform over the block
  const unsigned* gl = d.gscan;
  const unsigned* gr = d.gscan + d.vcap + 1;
  const int nvb_s = (n + kVB - 1) / kVB;
  const unsigned baseL = gl[v] - gl[sg.vb0];
  const unsigned baseR = gr[v] - gr[sg.vb0];
  const unsigned totR = gr[sg.vb0 + nvb_s] - gr[sg.vb0];
  // this thread's kPer consecutive slots
  const int64_t t0 = r.first + (int64_t)threadIdx.x * kPer;
  unsigned cl = 0, cr = 0;
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = t0 + k;
    if (i >= r.end) break;
    const float c = slot_c(SL(d), i, p.axis);
    cl += (i < p.split && !(c < p.pos)) ? 1u : 0u;
    cr += (i >= p.split && c < p.pos) ? 1u : 0u;
  }
  unsigned rl = baseL + block_exscan(cl, buf);
  unsigned rr = baseR + block_exscan(cr, buf);
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = t0 + k;
    if (i >= r.end) break;
    const float c = slot_c(SL(d), i, p.axis);
    if (i < p.split && !(c < p.pos)) d.lpos[sg.b + rl++] = (int32_t)i;
    if (i >= p.split && c < p.pos) d.rpos[sg.b + (totR - 1 - rr++)] = (int32_t)i;
  }
}

__global__ __launch_bounds__(kBlock) void k_swap(DevBuild d) {
  VRange r;
  if (!vrange(d, r)) return;
  const BuildPlan p = d.plan[r.s];
  const BuildSeg sg = d.seg[r.s];
  const int n = sg.e - sg.b;
  if (p.leaf || p.nless == 0 || p.nless == n) return;
  const unsigned* gl = d.gscan;
  const int nvb_s = (n + kVB - 1) / kVB;
  const int64_t nswap = gl[sg.vb0 + nvb_s] - gl[sg.vb0];
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = r.first + k * kBlock + threadIdx.x;
    if (i >= r.end) break;
    const int64_t j = i - sg.b;
    if (j >= nswap) break;
    const int64_t a = d.lpos[sg.b + j], b = d.rpos[sg.b + j];
    float4* pa = reinterpret_cast<float4*>(SL(d) + a);
    float4* pb = reinterpret_cast<float4*>(SL(d) + b);
    const float4 a0 = pa[0], a1 = pa[1], a2 = pa[2];
    const float4 b0 = pb[0], b1 = pb[1], b2 = pb[2];
    pa[0] = b0;
    pa[1] = b1;
    pa[2] = b2;
    pb[0] = a0;
    pb[1] = a1;
    pb[2] = a2;
  }
}

__global__ __launch_bounds__(kBlock) void k_child_bounds(DevBuild d) {
  __shared__ unsigned long long sk[12];  // left min 3, left max 3, right min 3, right max 3
  VRange r;
  if (!vrange(d, r)) return;
  const BuildPlan p = d.plan[r.s];
  if (p.leaf) return;
  if (threadIdx.x < 12) sk[threadIdx.x] = (threadIdx.x % 6) < 3 ? kMinInit : 0ull;
  __syncthreads();
  unsigned long long m[12];
  for (int k = 0; k < 12; ++k) m[k] = (k % 6) < 3 ? kMinInit : 0ull;
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = r.first + k * kBlock + threadIdx.x;
    if (i >= r.end) break;
    Slot o;
    load_slot(SL(d), i, o);
    const int side = i < p.split ? 0 : 6;
    for (int a = 0; a < 3; ++a) {
      const unsigned long long x = kmin(o.lo[a], (uint32_t)i), y = kmax(o.hi[a], (uint32_t)i);
      m[side + a] = x < m[side + a] ? x : m[side + a];
      m[side + 3 + a] = y > m[side + 3 + a] ? y : m[side + 3 + a];
    }
  }
  for (int k = 0; k < 12; ++k) {
    if ((k % 6) < 3) {
      if (m[k] != kMinInit) atomicMin(&sk[k], m[k]);
    } else if (m[k] != 0ull) {
      atomicMax(&sk[k], m[k]);
    }
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    const int f = k < 3 ? kFlMin + k : k < 6 ? kFlMax + k - 3 : k < 9 ? kFrMin + k - 6 : kFrMax + k - 9;
    if ((k % 6) < 3) {
      if (sk[k] != kMinInit) atomicMin(&K(d, f, r.s), sk[k]);
    } else if (sk[k] != 0ull) {
      atomicMax(&K(d, f, r.s), sk[k]);
    }
  }
}

__device__ __forceinline__ void push_node(const DevBuild& d, int rec, int b, int e) {
  const BuildSeg sg{rec, b, e, 0};
  if (e - b > kBuildSerial) {
    d.seg_next[atomicAdd(&d.ctr[kCtrNext], 1u)] = sg;
  } else {
    d.small[atomicAdd(&d.ctr[kCtrSmall], 1u)] = sg;
  }
}

__global__ __launch_bounds__(kBlock) void k_seg_children(DevBuild d) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= d.nseg) return;
  const BuildPlan p = d.plan[s];
  const BuildSeg sg = d.seg[s];
  RefNode& nd = d.u.refs[sg.rec];
  const int n = sg.e - sg.b;
  if (p.leaf) {  // sorted later, with the other large leaves
    if (n > 32767) {
      atomicExch(d.u.err, 3);
      return;
    }
    nd.child = sg.b;
    nd.children = (int16_t)n;
    nd.axis = 0;
    d.big[atomicAdd(&d.ctr[kCtrBig], 1u)] = sg;
    return;
  }
  const unsigned spot = atomicAdd(d.u.next_free, 2u);
  if (spot + 1 >= (unsigned)d.u.n_records) {
    atomicExch(d.u.err, 1);
    return;
  }
  nd.child = (int32_t)spot;
  nd.children = 0;
  nd.axis = (uint8_t)p.axis;
  for (int side = 0; side < 2; ++side) {
    RefNode& c = d.u.refs[spot + side];
    const int fmin = side ? kFrMin : kFlMin, fmax = side ? kFrMax : kFlMax;
    for (int a = 0; a < 3; ++a) {
      c.box.lo[a] = key_value(K(d, fmin + a, s), true, SL(d), 4 + a);
      c.box.hi[a] = key_value(K(d, fmax + a, s), false, SL(d), 7 + a);
    }
    c.child = 0;
    c.children = 0;
    c.axis = 0;
    c.bits = 0;
  }
  push_node(d, spot, sg.b, p.split);
  push_node(d, spot + 1, p.split, sg.e);
}

// ---------------------------------------------------------------------------------------
// Phase 2 and large leaves

__global__ __launch_bounds__(kBlock) void k_unslot(DevBuild d) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.u.n_objects) return;
  d.u.ids[i] = __float_as_int(reinterpret_cast<const float4*>(SL(d) + i)[0].w);
}

__global__ __launch_bounds__(64) void k_serial(DevBuild d, int64_t n_small) {
  const int64_t k = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (k >= n_small) return;
  const BuildSeg sg = d.small[k];
  RefNode& nd = d.u.refs[sg.rec];
  const int n = sg.e - sg.b;
  nd.child = sg.b;
  nd.children = (int16_t)n;
  nd.axis = kPending;
  // A subtree of n objects has at most n - 1 inner nodes: reserve their pairs at once.
  const unsigned need = 2u * (unsigned)(n - 1);
  const unsigned base = need ? atomicAdd(d.u.next_free, need) : 0u;
  if (need && base + need > (unsigned)d.u.n_records) {
    atomicExch(d.u.err, 1);
    return;
  }
  RangePairs pairs{base, base + need};
  build_loop(d.u, pairs, sg.rec, sg.rec);
  for (unsigned r = pairs.next; r < base + need; ++r) {  // unused records: marked as holes
    RefNode& h = d.u.refs[r];
    h.child = 0;
    h.children = kHole;
    h.axis = 0;
    h.bits = 0;
  }
}

// std::sort of a large leaf's ids: bitonic sort of the range padded to a power of two
// in scratch (two ints per slot of the range, so leaves never share scratch).
__global__ __launch_bounds__(1024) void k_sort_big(DevBuild d, int32_t* scratch) {
  const BuildSeg sg = d.big[blockIdx.x];
  const int n = sg.e - sg.b;
  int P = 1;
  while (P < n) P <<= 1;
  int32_t* a = scratch + 2 * (int64_t)sg.b;
  for (int i = threadIdx.x; i < P; i += blockDim.x) a[i] = i < n ? d.u.ids[sg.b + i] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const int32_t x = a[i], y = a[l];
          const bool up = (i & k) == 0;
          if (up ? x > y : x < y) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) d.u.ids[sg.b + i] = a[i];
}

// ---------------------------------------------------------------------------------------
// Phase 3: reference numbering and the bottom-up passes


__global__ __launch_bounds__(kBlock) void k_ic_leaves(DevFinish f, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) f.ic[f.leaves[i]] = 0;
}
__global__ __launch_bounds__(kBlock) void k_ic_level(DevFinish f, int64_t first, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int x = f.levels[first + i];
  const int c = f.trefs[x].child;
  f.ic[x] = 1 + f.ic[c] + f.ic[c + 1];
}
__global__ void k_rank_root(DevFinish f) {
  f.rank[0] = 0;
  f.fin[0] = 0;
}
// Pairs are handed out in preorder of their parents (DynBVH::build recursion).
__global__ __launch_bounds__(kBlock) void k_rank_level(DevFinish f, int64_t first, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int x = f.levels[first + i];
  const int c = f.trefs[x].child;
  const int rk = f.rank[x];
  f.fin[c] = 2 * rk + 1;
  f.fin[c + 1] = 2 * rk + 2;
  f.rank[c] = rk + 1;
  f.rank[c + 1] = rk + 1 + f.ic[c];
}
// Records in use move to their reference numbers; records from the node count on are
// left as DynBVH::allocate leaves them after a reset. Holes (records the one-lane
// subtrees reserved but did not use) have no number and vanish.
__global__ __launch_bounds__(kBlock) void k_emit(DevFinish f) {
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (x >= f.n_records) return;
  if (x >= f.n_final) {
    RefNode b;
    for (int a = 0; a < 3; ++a) {
      b.box.lo[a] = FLT_MAX;
      b.box.hi[a] = -FLT_MAX;
    }
    b.child = 0;
    b.axis = 0;
    b.bits = 0;
    b.children = 0;
    f.refs[x] = b;
  }
  if (x >= (int64_t)*f.tcount) return;
  const int32_t to = f.fin[x];
  if (to < 0 || to >= f.n_final) return;  // a hole
  RefNode t = f.trefs[x];
  t.bits = 0;
  if (t.children == 0) t.child = 2 * f.rank[x] + 1;
  f.refs[to] = t;
}
__global__ __launch_bounds__(kBlock) void k_map_lists(DevFinish f, int64_t n_inner,
                                                      int64_t n_leaves) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n_inner) f.levels[i] = f.fin[f.levels[i]];
  if (i < n_leaves) f.leaves[i] = f.fin[f.leaves[i]];
}
// computeSubTreeSizes, computeCost<true> and the RTSAH pass (DynBVH.cc:1971-1996,
// 3388-3415, 1594-1678) for the leaves, then one launch per depth, deepest first.
__global__ __launch_bounds__(kBlock) void k_fin_leaves(DevFinish f, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int r = f.leaves[i];
  RefNode& nd = f.refs[r];
  nd.bits &= ~2;
  f.ssize[r] = 1;
  f.costs[r] = nd.children * kCIsec;
  const rtsah::CostPair c = rtsah::leaf_cost(nd.children);
  f.rt[2 * r] = c.c0;
  f.rt[2 * r + 1] = c.c1;
}
__global__ __launch_bounds__(kBlock) void k_fin_level(DevFinish f, int64_t first, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int r = f.levels[first + i];
  RefNode& nd = f.refs[r];
  const int c = nd.child;
  const uint32_t size = 1 + f.ssize[c] + f.ssize[c + 1];
  f.ssize[r] = size;
  uint8_t bits = nd.bits & ~3;
  if (size > f.large) bits |= 2;
  const RefNode L = f.refs[c], R = f.refs[c + 1];
  const float ar = area(rbox(nd)), al = area(rbox(L)), arr = area(rbox(R));
  f.costs[r] = kCTrav + (al * f.costs[c] + arr * f.costs[c + 1]) / ar;
  Aabb P, Lb, Rb;
  for (int a = 0; a < 3; ++a) {
    P.lo[a] = nd.box.lo[a];
    P.hi[a] = nd.box.hi[a];
    Lb.lo[a] = L.box.lo[a];
    Lb.hi[a] = L.box.hi[a];
    Rb.lo[a] = R.box.lo[a];
    Rb.hi[a] = R.box.hi[a];
  }
  if (f.rtsah) {
    bool left_cheaper = false;
    const rtsah::CostPair cl{f.rt[2 * c], f.rt[2 * c + 1]};
    const rtsah::CostPair cr{f.rt[2 * c + 2], f.rt[2 * c + 3]};
    const rtsah::CostPair o = rtsah::combine(f.logtab, P, Lb, Rb, cl, cr, left_cheaper);
    if (left_cheaper) bits |= 1;
    f.rt[2 * r] = o.c0;
    f.rt[2 * r + 1] = o.c1;
  }
  nd.bits = bits;
}

unsigned grid(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

}  // namespace

// ---------------------------------------------------------------------------------------
// Launchers (the host drives the level loop; see mh_scene_build_device)

int build_launch_setup(const DevBuild& d, int serial, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_root_box, dim3(grid(d.n_verts)), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_slots, dim3(grid(d.u.n_objects)), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_root_init, dim3(1), dim3(1), 0, s, d, serial);
  return (int)hipGetLastError();
}

int build_launch_level(const DevBuild& d, int64_t nvb_bound, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const unsigned gs = grid(d.nseg);
  const unsigned gv = (unsigned)nvb_bound;
  hipLaunchKernelGGL(k_acc_clear, dim3(grid((int64_t)(kKeyFields + kCntFields) * d.nseg)),
                     dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_scan_segments, dim3(1), dim3(1024), 0, s, d);
  hipLaunchKernelGGL(k_vb_map, dim3(gs), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_cbounds, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_seg_axis, dim3(gs), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_bins, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_seg_split, dim3(gs), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_count_less, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_seg_mid, dim3(gs), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_mis_count, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_scan_mis, dim3(1), dim3(1024), 0, s, d);
  hipLaunchKernelGGL(k_mis_scatter, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_swap, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_child_bounds, dim3(gv), dim3(kBlock), 0, s, d);
  hipLaunchKernelGGL(k_seg_children, dim3(gs), dim3(kBlock), 0, s, d);
  return (int)hipGetLastError();
}

int build_launch_serial(const DevBuild& d, int64_t n_small, int64_t n_big, int32_t* scratch,
                        void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_unslot, dim3(grid(d.u.n_objects)), dim3(kBlock), 0, s, d);
  if (n_small > 0)
    hipLaunchKernelGGL(k_serial, dim3(grid(n_small, 64)), dim3(64), 0, s, d, n_small);
  if (n_big > 0) hipLaunchKernelGGL(k_sort_big, dim3((unsigned)n_big), dim3(1024), 0, s, d, scratch);
  return (int)hipGetLastError();
}

int build_launch_ranks(const DevFinish& f, const int64_t* level_off, int n_levels,
                       int64_t n_leaves, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (n_leaves > 0)
    hipLaunchKernelGGL(k_ic_leaves, dim3(grid(n_leaves)), dim3(kBlock), 0, s, f, n_leaves);
  for (int d = n_levels - 1; d >= 0; --d) {
    const int64_t first = level_off[d], n = level_off[d + 1] - first;
    if (n > 0) hipLaunchKernelGGL(k_ic_level, dim3(grid(n)), dim3(kBlock), 0, s, f, first, n);
  }
  hipLaunchKernelGGL(k_rank_root, dim3(1), dim3(1), 0, s, f);
  for (int d = 0; d < n_levels; ++d) {
    const int64_t first = level_off[d], n = level_off[d + 1] - first;
    if (n > 0) hipLaunchKernelGGL(k_rank_level, dim3(grid(n)), dim3(kBlock), 0, s, f, first, n);
  }
  hipLaunchKernelGGL(k_emit, dim3(grid(f.n_records)), dim3(kBlock), 0, s, f);
  const int64_t n_inner = level_off[n_levels];
  const int64_t m = n_inner > n_leaves ? n_inner : n_leaves;
  if (m > 0)
    hipLaunchKernelGGL(k_map_lists, dim3(grid(m)), dim3(kBlock), 0, s, f, n_inner, n_leaves);
  if (n_leaves > 0)
    hipLaunchKernelGGL(k_fin_leaves, dim3(grid(n_leaves)), dim3(kBlock), 0, s, f, n_leaves);
  for (int d = n_levels - 1; d >= 0; --d) {
    const int64_t first = level_off[d], n = level_off[d + 1] - first;
    if (n > 0) hipLaunchKernelGGL(k_fin_level, dim3(grid(n)), dim3(kBlock), 0, s, f, first, n);
  }
  return (int)hipGetLastError();
}

}  // namespace mh
