// Persistent multi-batch traversal for gfx950 (wave64).
//
// k_trace_batch consumes a table of ray batches (each its own SoA arrays, size and
// closest-hit/any-hit flag) in one launch. Rays are numbered across batches by the
// table's prefix offsets. Each wave keeps a pool of ray numbers taken 64 at a time from a
// global counter (one atomic per 64 rays); a lane whose ray has finished takes the next
// number from the pool as soon as at least `refill` lanes of its wave are idle, so lanes
// do not wait for the wave's slowest ray (ballot + mbcnt hand-out, no LDS traffic).
// The last wave to leave resets the counters, so consecutive launches (and graph replays)
// need no clearing pass.
//
// Two node layouts, both giving the reference's one-ray results and visit order
// (DynBVH::intersectNode on a one-ray packet, Model/Groups/DynBVH.cc:274-404):
//   PAIR: 64-byte PairNodes hold both child boxes; the far child is pushed with its entry
//         distance and skipped at pop time if that exceeds the current minT.
//   REF:  the reference's 32-byte records; each node's own box is tested when it is
//         visited, exactly as firstIntersects does.
// Triangle tests and arithmetic are those of trace.hip (shared header).

#include <hip/hip_runtime.h>

#include <cfloat>

#include "mh_device.h"
#include "trace_common.h"

namespace mh {
namespace {

constexpr int kMaxBatchesLds = 256;

struct Lane {
  Ray r;
  float minT;
  int best;
  float bu, bv;
  int cur;
  int sp, base;
  int bi;       // batch index
  int64_t li;   // ray index inside the batch
  bool any;
};

template <int PRIM>
__device__ __forceinline__ void finish(const BatchDev* batches, const DevScene& sc, Lane& L) {
  const BatchDev& B = batches[L.bi];
  const bool hit = L.best >= 0;
  const int32_t prim = hit ? sc.obj_ids[L.best] : -1;
  store_hit(B.hits, L.li, prim, (L.any && hit) ? kEps : L.minT, L.bu, L.bv, hit);
}

template <int S, int BLOCK>
__device__ __forceinline__ void push(Lane& L, int* s_ref, float* s_tn, int2* spill,
                                     int64_t spill_stride, int64_t gtid, int tid, int ref,
                                     float tn) {
  if (L.sp - L.base == S) {
    const int sl = L.base & (S - 1);
    spill[(int64_t)L.base * spill_stride + gtid] =
        make_int2(s_ref[sl * BLOCK + tid], __float_as_int(s_tn[sl * BLOCK + tid]));
    ++L.base;
  }
  const int sl = L.sp & (S - 1);
  s_ref[sl * BLOCK + tid] = ref;
  s_tn[sl * BLOCK + tid] = tn;
  ++L.sp;
}

template <int S, int BLOCK>
__device__ __forceinline__ void pop(Lane& L, const int* s_ref, const float* s_tn,
                                    const int2* spill, int64_t spill_stride, int64_t gtid, int tid,
                                    int& ref, float& tn) {
  if (L.sp == L.base) {
    --L.sp;
    --L.base;
    const int2 e = spill[(int64_t)L.sp * spill_stride + gtid];
    ref = e.x;
    tn = __int_as_float(e.y);
  } else {
    --L.sp;
    const int sl = L.sp & (S - 1);
    ref = s_ref[sl * BLOCK + tid];
    tn = s_tn[sl * BLOCK + tid];
  }
}

// One traversal step; returns false when the ray is finished.
template <int PRIM, int LAYOUT, int S, int BLOCK>
__device__ __forceinline__ bool step(const DevScene& sc, Lane& L, int* s_ref, float* s_tn,
                                     int2* spill, int64_t spill_stride, int64_t gtid, int tid) {
  if (LAYOUT == kLayoutPair) {
    if (L.cur >= 0) {
      const float4* np = reinterpret_cast<const float4*>(sc.pairs + L.cur);
      const float4 A = np[0], B = np[1], C = np[2];
      const int4 M = reinterpret_cast<const int4*>(np)[3];
      float tn0, tn1;
      const bool h0 = slab(A.x, A.y, A.z, A.w, B.x, B.y, L.r, L.minT, tn0);
      const bool h1 = slab(B.z, B.w, C.x, C.y, C.z, C.w, L.r, L.minT, tn1);
      bool one_first;
      if (L.any) {
        one_first = ((M.z >> 2) & 1) == 0;
      } else {
        const int ax = M.z & 3;
        const float da = pick3(ax, L.r.d[0], L.r.d[1], L.r.d[2]);
        one_first = !(da > 0.f);
      }
      const int fref = one_first ? M.y : M.x;
      const int bref = one_first ? M.x : M.y;
      const bool fh = one_first ? h1 : h0;
      const bool bh = one_first ? h0 : h1;
      if (fh) {
        if (bh) push<S, BLOCK>(L, s_ref, s_tn, spill, spill_stride, gtid, tid, bref, one_first ? tn0 : tn1);
        L.cur = fref;
        return true;
      }
      if (bh) {
        L.cur = bref;
        return true;
      }
    } else {
      const int code = ~L.cur;
      const int first = code >> kLeafCountBits;
      int cnt = code & kLeafCountInline;
      if (cnt == kLeafCountInline) cnt = sc.big_count[first];
      for (int k = 0; k < cnt; ++k) {
        float t, u, v;
        if (prim_test<PRIM>(sc.tris, first + k, L.r, L.minT, t, u, v)) {
          L.minT = t;
          L.best = first + k;
          L.bu = u;
          L.bv = v;
          if (L.any) return false;
        }
      }
    }
    while (L.sp > 0) {
      int ref;
      float tn;
      pop<S, BLOCK>(L, s_ref, s_tn, spill, spill_stride, gtid, tid, ref, tn);
      if (tn <= L.minT) {
        L.cur = ref;
        return true;
      }
    }
    return false;
  } else {
    const RefNode nd = sc.refs[L.cur];
    float tn;
    if (slab(nd.box.lo[0], nd.box.hi[0], nd.box.lo[1], nd.box.hi[1], nd.box.lo[2], nd.box.hi[2],
             L.r, L.minT, tn)) {
      if (nd.children != 0) {
        for (int k = 0; k < nd.children; ++k) {
          float t, u, v;
          if (prim_test<PRIM>(sc.tris, nd.child + k, L.r, L.minT, t, u, v)) {
            L.minT = t;
            L.best = nd.child + k;
            L.bu = u;
            L.bv = v;
            if (L.any) return false;
          }
        }
      } else {
        int front;
        if (L.any) {
          front = (nd.bits & 1) ? 0 : 1;
        } else {
          const float da = pick3(nd.axis, L.r.d[0], L.r.d[1], L.r.d[2]);
          front = da > 0.f ? 0 : 1;
        }
        push<S, BLOCK>(L, s_ref, s_tn, spill, spill_stride, gtid, tid, nd.child + 1 - front, 0.f);
        L.cur = nd.child + front;
        return true;
      }
    }
    if (L.sp > 0) {
      int ref;
      float tn;
      pop<S, BLOCK>(L, s_ref, s_tn, spill, spill_stride, gtid, tid, ref, tn);
      L.cur = ref;
      return true;
    }
    return false;
  }
}

// Starts ray g (global number); returns false if it finished immediately.
template <int PRIM, int LAYOUT>
__device__ __forceinline__ bool start_ray(const DevScene& sc, const BatchDev* batches,
                                          const int64_t* s_start, int nb, int64_t g, Lane& L) {
  int lo = 0, hi = nb - 1;  // last batch with start <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_start[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  L.bi = lo;
  L.li = g - s_start[lo];
  const BatchDev& B = batches[lo];
  L.any = (B.flags & 1u) != 0;
  if (B.rays.masked && B.rays.masked[L.li]) {
    store_hit(B.hits, L.li, -2, -kMaxT, 0.f, 0.f, true);
    return false;
  }
  load_ray(B.rays, L.li, L.r);
  L.minT = B.rays.tmax[L.li];
  L.best = -1;
  L.bu = L.bv = 0.f;
  L.sp = L.base = 0;
  if (LAYOUT == kLayoutPair) {
    const RefNode root = sc.refs[0];
    float tn;
    if (!slab(root.box.lo[0], root.box.hi[0], root.box.lo[1], root.box.hi[1], root.box.lo[2],
              root.box.hi[2], L.r, L.minT, tn)) {
      store_hit(B.hits, L.li, -1, L.minT, 0.f, 0.f, false);
      return false;
    }
    L.cur = sc.root_ref;
  } else {
    L.cur = 0;
  }
  return true;
}

template <int PRIM, int LAYOUT, int S, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_batch(DevScene sc, const BatchTable* table,
                                                       unsigned long long* ctr, int2* spill,
                                                       int64_t spill_stride, int refill) {
  static_assert((S & (S - 1)) == 0, "ring size is a power of two");
  __shared__ int s_ref[S * BLOCK];
  __shared__ float s_tn[S * BLOCK];
  __shared__ int64_t s_start[kMaxBatchesLds];
  const int nb = table->nb < kMaxBatchesLds ? table->nb : kMaxBatchesLds;
  const int64_t total = table->total;
  const BatchDev* batches = table->batches;
  for (int i = threadIdx.x; i < nb; i += BLOCK) s_start[i] = batches[i].start;
  __syncthreads();

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int64_t gtid = (int64_t)blockIdx.x * BLOCK + tid;
  Lane L;
  bool active = false;
  bool exhausted = false;
  int64_t pool_next = 0, pool_end = 0;  // wave-uniform

  for (;;) {
    const unsigned long long idle = __ballot(!active);
    const int n_idle = __popcll(idle);
    if (n_idle == 64 && exhausted) break;
    if (n_idle >= refill && !exhausted) {
      const int rank = __popcll(idle & ((1ull << lane) - 1ull));
      const int64_t avail = pool_end - pool_next;
      int64_t g = -1;
      if (!active && rank < avail) g = pool_next + rank;
      if (n_idle > avail) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(ctr, 64ull);
        const int64_t base = (int64_t)__shfl((long long)b, 0, 64);
        if (base >= total) {
          exhausted = true;
          pool_next = pool_end = total;
        } else {
          const int64_t end = base + 64 < total ? base + 64 : total;
          if (!active && rank >= avail && base + (rank - avail) < end) g = base + (rank - avail);
          const int64_t taken = n_idle - avail;
          pool_next = base + taken < end ? base + taken : end;
          pool_end = end;
        }
      } else {
        pool_next += n_idle;
      }
      if (g >= 0) active = start_ray<PRIM, LAYOUT>(sc, batches, s_start, nb, g, L);
    }
    if (active) {
      if (!step<PRIM, LAYOUT, S, BLOCK>(sc, L, s_ref, s_tn, spill, spill_stride, gtid, tid)) {
        finish<PRIM>(batches, sc, L);
        active = false;
      }
    }
  }
  if (lane == 0) {
    const unsigned long long waves = (unsigned long long)gridDim.x * (BLOCK / 64);
    const unsigned long long old = atomicAdd(ctr + 1, 1ull);
    if (old == waves - 1) {
      atomicExch(ctr + 0, 0ull);
      atomicExch(ctr + 1, 0ull);
    }
  }
}

template <int PRIM, int LAYOUT, int BLOCK>
int launch_s(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  dim3 g(bl.blocks), b(BLOCK);
  int2* sp = reinterpret_cast<int2*>(bl.spill);
  switch (bl.stack_lds) {
    case 8:
      hipLaunchKernelGGL((k_trace_batch<PRIM, LAYOUT, 8, BLOCK>), g, b, 0, s, sc, bl.table,
                         bl.counters, sp, bl.spill_stride, bl.refill);
      break;
    case 32:
      hipLaunchKernelGGL((k_trace_batch<PRIM, LAYOUT, 32, BLOCK>), g, b, 0, s, sc, bl.table,
                         bl.counters, sp, bl.spill_stride, bl.refill);
      break;
    default:
      hipLaunchKernelGGL((k_trace_batch<PRIM, LAYOUT, 16, BLOCK>), g, b, 0, s, sc, bl.table,
                         bl.counters, sp, bl.spill_stride, bl.refill);
      break;
  }
  return (int)hipGetLastError();
}

template <int PRIM, int LAYOUT>
int launch_b(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  switch (bl.block) {
    case 64:
      return launch_s<PRIM, LAYOUT, 64>(sc, bl, s);
    case 128:
      return launch_s<PRIM, LAYOUT, 128>(sc, bl, s);
    default:
      return launch_s<PRIM, LAYOUT, 256>(sc, bl, s);
  }
}

}  // namespace

int launch_trace_batch(const DevScene& sc, const BatchLaunch& bl, void* stream) {
  if (bl.nb <= 0 || bl.nb > kMaxBatchesLds) return (int)hipErrorInvalidValue;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (sc.prim_kind == 1)
    return bl.layout == kLayoutRef ? launch_b<1, kLayoutRef>(sc, bl, s)
                                   : launch_b<1, kLayoutPair>(sc, bl, s);
  return bl.layout == kLayoutRef ? launch_b<0, kLayoutRef>(sc, bl, s)
                                 : launch_b<0, kLayoutPair>(sc, bl, s);
}

}  // namespace mh
