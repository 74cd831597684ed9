// Persistent multi-batch traversal for gfx950 (wave64). Included by the four
// batch_<triangle>_<layout>.hip files, which instantiate the launchers of one triangle
// test and node layout each, so that the kernel variants compile in parallel.
//
// k_trace_batch consumes a table of ray batches (each its own SoA arrays, size and
// closest-hit/any-hit flag) in one launch. Rays are numbered across batches by the
// table's prefix offsets. Three schedules (persistent workgroups in all of them):
//   whole-wave chunks (refill 64, chunk C > 0): a wave traces C consecutive rays, 64 at a
//     time, so the rays of a wave stay as coherent as their batch. Its first chunk is
//     assigned statically (wave id), later ones come from a shared counter, one atomic
//     per C rays, which balances the load between waves;
//   static stride (refill 64, chunk 0): the same without the counter;
//   per-lane refill (refill < 64): each wave keeps a pool of ray numbers taken C at a
//     time from the counter (or from its XCD's eighth of the rays first); a lane whose
//     ray has finished takes the next number, between traversal rounds, as soon as at
//     least `refill` lanes of its wave are idle, so lanes do not wait for the wave's
//     slowest ray (ballot + popcount hand-out, no LDS traffic).
// With `reverse` the chunks are handed out from the end of the table (last batch first,
// rays ascending inside each chunk), so that a table listing its costly batches last still
// ends on short rays. The last wave to leave resets the counters, so consecutive launches
// (and graph replays) need no clearing pass.
//
// Three node layouts, all giving the reference's one-ray results and visit order
// (DynBVH::intersectNode on a one-ray packet, Model/Groups/DynBVH.cc:274-404; the
// traversal loops are traverse.h's): PAIR (64-byte records with both child boxes), SOA
// (the same records as four arrays of their 16-byte fields) and REF (the reference's
// 32-byte records).
// Triangle tests and arithmetic are those of trace.hip (shared header).
//
// A table whose batches are all closest-hit (or all any-hit) runs a kernel specialised to
// that kind, without the other traversal's code and registers; mixed tables branch per
// batch span (whole-wave schedules) or per lane (per-lane refill).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "mh_device.h"
#include "trace_common.h"
#include "traverse.h"

namespace mh {
namespace {

constexpr int kMaxBatches = 256;  // MH_MAX_BATCHES
constexpr int kKindMixed = 0;
constexpr int kKindClosest = 1;
constexpr int kKindAny = 2;
constexpr int kModeLane = 0;        // per-lane refill from a wave pool
constexpr int kModeWave = 1;        // whole-wave chunks from the shared counter
constexpr int kModeWaveStatic = 2;  // whole-wave chunks on a static stride, no counter
// Launch counters: [0] shared ray counter, [1] waves that left, [2..9] per-XCD ray counters.
constexpr int kCounters = (int)(kBatchCounterBytes / sizeof(unsigned long long));

// Last batch whose first ray number is <= g (empty batches are skipped over).
__device__ __forceinline__ int find_batch(const int64_t* s_start, int nb, int64_t g) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_start[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Rays [pos, gend) of one batch whose first ray number is bstart, 64 per step of the wave.
template <bool ANY, int PRIM, int LAYOUT, int S, int BLOCK>
__device__ __forceinline__ void trace_span(const DevScene& sc, const RefNode& root,
                                           const GRays& rays, const GHits& hits,
                                           int64_t bstart, int64_t pos, int64_t gend, int lane,
                                           LaneStack<S, BLOCK>& st) {
  unsigned long long nbox = 0, ntri = 0;
  for (int64_t b = pos; b < gend; b += 64) {
    const int64_t g = b + lane;
    if (g < gend)
      trace_store<ANY, PRIM, LAYOUT, S, BLOCK, false>(sc, root, rays, hits, g - bstart, st, nbox,
                                                      ntri);
  }
}

// Per-lane state of the lane-refill schedule: one traversal advanced a round at a time.
struct Lane {
  Ray r;
  RayHit h;
  int cur;
  int bi;      // batch index
  int64_t li;  // ray index inside the batch
  bool any;
  bool regular;  // slab_fast is exact for this ray (regular_ray at its start)
};

// Per-lane refill schedule, REF layout: one traversal step of traverse.h's trace_ref with
// a run-time closest/any-hit flag; returns false when the ray is finished.
template <int PRIM, int S, int BLOCK>
__device__ __forceinline__ bool step_ref(const DevScene& sc, Lane& L, LaneStack<S, BLOCK>& st) {
  unsigned long long ntri = 0;
  const RefNode nd = sc.refs[L.cur];
  float tn;
  if (slab(nd.box.lo[0], nd.box.hi[0], nd.box.lo[1], nd.box.hi[1], nd.box.lo[2], nd.box.hi[2],
           L.r, L.h.minT, tn)) {
    if (nd.children != 0) {
      const bool done = L.any ? leaf<true, PRIM, false>(sc, nd.child, nd.children, L.r, L.h, ntri)
                              : leaf<false, PRIM, false>(sc, nd.child, nd.children, L.r, L.h, ntri);
      if (done) return false;
    } else {
      const int front = L.any ? ((nd.bits & 1) ? 0 : 1)
                              : (((L.r.dpos >> (nd.axis & 3)) & 1) ? 0 : 1);
      st.push(nd.child + 1 - front, 0.f);
      L.cur = nd.child + front;
      return true;
    }
  }
  if (st.empty()) return false;
  st.pop(L.cur, tn);
  return true;
}

// Per-lane refill schedule, PAIR and SOA layouts: one round of traverse.h's trace_pair loop,
// resumed from the lane's state: the walk down inner nodes to a leaf (or to a miss), the leaf's
// triangles and the pops to the next entry still worth visiting. The lanes of a wave walk
// the inner nodes in lockstep (a lane that reached its leaf waits for the others), so the
// wave issues each inner-node step once for all of them, as the plain kernel does; a
// finished lane is refilled between rounds. Returns false when the ray is finished.
template <bool ANY, int PRIM, int S, int BLOCK, bool FAST, int LAYOUT>
__device__ __forceinline__ bool round_pair(const DevScene& sc, Lane& L, LaneStack<S, BLOCK>& st) {
  unsigned long long ntri = 0;
  bool missed = false;
  while (L.cur >= 0) {
    float4 A, B, C;
    int4 M;
    fetch_pair<LAYOUT>(sc, L.cur, A, B, C, M);
    float tn0, tn1;
    const bool h0 = slab_v<FAST>(A.x, A.y, A.z, A.w, B.x, B.y, L.r, L.h.minT, tn0);
    const bool h1 = slab_v<FAST>(B.z, B.w, C.x, C.y, C.z, C.w, L.r, L.h.minT, tn1);
    const bool one_first = ANY ? ((M.z >> 2) & 1) == 0 : __builtin_amdgcn_ubfe(L.r.dpos, M.z, 1) == 0;
    if (h0 && h1) st.push(one_first ? M.x : M.y, one_first ? tn0 : tn1);
    if (!(h0 || h1)) {
      missed = true;
      break;
    }
    L.cur = (h1 && (one_first || !h0)) ? M.y : M.x;
  }
  if (!missed) {
    const int code = ~L.cur;
    const int first = code >> kLeafCountBits;
    int cnt = code & kLeafCountInline;
    if (cnt == kLeafCountInline) cnt = sc.big_count[first];
    if (leaf<ANY, PRIM, false>(sc, first, cnt, L.r, L.h, ntri)) return false;
  }
  while (!st.empty()) {
    int ref;
    float t;
    st.pop(ref, t);
    if (t <= L.h.minT) {
      L.cur = ref;
      return true;
    }
  }
  return false;
}

// Next chunk [base, end) of ray numbers for a wave of the per-lane refill schedule, or
// base = -1 once there are none. One range: a shared counter over all rays (ctr[0]). Per
// XCD: the rays are cut into eight contiguous ranges with a counter each (ctr[2 + k]); a
// wave starts on the range of its XCD (workgroups are dealt to the 8 XCDs round-robin), so
// each XCD's L2 serves the nodes of one region of the rays, and moves on to the next
// ranges in turn when its own is used up. xr (wave-uniform) counts the ranges this wave
// found used up.
__device__ __forceinline__ void grab_chunk(unsigned long long* ctr, int64_t C, int64_t total,
                                           bool per_xcd, int& xr, int lane, int64_t& base,
                                           int64_t& end) {
  long long b = -1, e = -1;
  int nx = xr;
  if (lane == 0) {
    if (!per_xcd) {
      const int64_t o = (int64_t)atomicAdd(ctr, (unsigned long long)C);
      if (o < total) {
        b = o;
        e = o + C < total ? o + C : total;
      }
    } else {
      const int x0 = (int)(blockIdx.x & 7);
      while (nx < 8) {
        const int k = (x0 + nx) & 7;
        const int64_t lo = total * k / 8, hi = total * (k + 1) / 8;
        const int64_t o = (int64_t)atomicAdd(ctr + 2 + k, (unsigned long long)C);
        if (lo + o < hi) {
          b = lo + o;
          e = b + C < hi ? b + C : hi;
          break;
        }
        ++nx;
      }
    }
  }
  base = (int64_t)__shfl(b, 0, 64);
  end = (int64_t)__shfl(e, 0, 64);
  xr = __shfl(nx, 0, 64);
}

// Starts ray li of batch B in lane L; returns false if it finished immediately.
template <int PRIM, int LAYOUT, int S, int BLOCK>
__device__ __forceinline__ bool start_in(const DevScene& sc, const RefNode& root, const BatchDev& B,
                                         int64_t li, Lane& L, LaneStack<S, BLOCK>& st) {
  L.li = li;
  L.any = (B.flags & 1u) != 0;
  if (B.rays.masked && B.rays.masked[L.li]) {
    store_hit(B.hits, L.li, -2, -kMaxT, 0.f, 0.f, true);
    return false;
  }
  load_ray(B.rays, L.li, L.r);
  L.h = RayHit{B.rays.tmax[L.li], -1, 0.f, 0.f};
  L.regular = regular_ray(L.r, L.h.minT);  // minT only falls to finite hit distances
  st.reset();
  if (LAYOUT != kLayoutRef) {
    float tn;
    if (!slab(root.box.lo[0], root.box.hi[0], root.box.lo[1], root.box.hi[1], root.box.lo[2],
              root.box.hi[2], L.r, L.h.minT, tn)) {
      store_hit(B.hits, L.li, -1, L.h.minT, 0.f, 0.f, false);
      return false;
    }
    L.cur = sc.root_ref;
  } else {
    L.cur = 0;
  }
  return true;
}

// Starts ray g (global number); returns false if it finished immediately. When every lane
// starting a ray takes it from the same batch (the usual case), the batch's record is read
// once for the wave by scalar loads instead of once per lane.
template <int PRIM, int LAYOUT, int S, int BLOCK>
__device__ __forceinline__ bool start_ray(const DevScene& sc, const RefNode& root,
                                          const BatchDev* batches, const int64_t* s_start, int nb,
                                          int64_t g, Lane& L, LaneStack<S, BLOCK>& st) {
  L.bi = find_batch(s_start, nb, g);
  const int b0 = __builtin_amdgcn_readfirstlane(L.bi);
  if (__ballot(L.bi != b0) == 0ull)
    return start_in<PRIM, LAYOUT, S, BLOCK>(sc, root, batches[b0], g - s_start[b0], L, st);
  return start_in<PRIM, LAYOUT, S, BLOCK>(sc, root, batches[L.bi], g - s_start[L.bi], L, st);
}

__device__ __forceinline__ void finish_in(const BatchDev& B, const DevScene& sc, const Lane& L) {
  const bool hit = L.h.best >= 0;
  store_hit(B.hits, L.li, hit ? sc.obj_ids[L.h.best] : -1, (L.any && hit) ? kEps : L.h.minT,
            L.h.u, L.h.v, hit);
}

// Stores a finished ray's hit (batch record by scalar loads when the lanes agree on it).
__device__ __forceinline__ void finish(const BatchDev* batches, const DevScene& sc, const Lane& L) {
  const int b0 = __builtin_amdgcn_readfirstlane(L.bi);
  if (__ballot(L.bi != b0) == 0ull)
    finish_in(batches[b0], sc, L);
  else
    finish_in(batches[L.bi], sc, L);
}

// Seven waves per SIMD (registers fitted to it, without spills) instead of the six the
// unconstrained allocation leaves: more node fetches in flight. Eight would spill to scratch
// and measured slower. The per-lane refill kernel for mixed tables keeps six: it carries
// both ray kinds' steps and would spill at seven.
template <int PRIM, int LAYOUT, int S, int BLOCK, int MODE, int KIND>
__global__ __launch_bounds__(BLOCK) __attribute__((
    amdgpu_waves_per_eu(MODE == kModeLane && KIND == kKindMixed ? 6 : 7, 8))) void k_trace_batch(
    DevScene sc, const BatchTable* __restrict__ table, unsigned long long* ctr, int2* spill,
    int64_t spill_stride, int nb_launch, int refill, int chunk, int xcd_remap, int reverse) {
  __shared__ unsigned long long s_ring[S * BLOCK];
  extern __shared__ int64_t s_start[];  // nb entries (dynamic LDS)
  // The dynamic LDS holds nb_launch offsets; a table with more entries is cut to that.
  const int nb = table->nb < nb_launch ? table->nb : nb_launch;
  const int64_t total = nb == table->nb ? table->total : table->batches[nb - 1].start +
                                                             table->batches[nb - 1].count;
  const BatchDev* batches = table->batches;
  for (int i = threadIdx.x; i < nb; i += BLOCK) s_start[i] = batches[i].start;
  __syncthreads();

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int64_t gtid = (int64_t)blockIdx.x * BLOCK + tid;  // spill column (physical)
  // Workgroups are dealt to the 8 XCDs round-robin; with xcd_remap the logical order
  // gives XCD x the x-th eighth of the workgroups, so its static shares are contiguous.
  int blk = blockIdx.x;
  if (xcd_remap && (gridDim.x & 7) == 0) blk = (blk & 7) * (gridDim.x >> 3) + (blk >> 3);
  const int64_t n_waves = (int64_t)gridDim.x * (BLOCK / 64);
  const int64_t wave_id = ((int64_t)blk * BLOCK + tid) >> 6;
  const RefNode root = sc.refs[0];
  LaneStack<S, BLOCK> st(s_ring, reinterpret_cast<unsigned long long*>(spill), spill_stride,
                         gtid, tid);

  if constexpr (MODE != kModeLane) {
    // A wave traces C consecutive ray numbers, 64 at a time, so its rays stay as coherent
    // as their batch. Its first chunk is wave_id * C; later chunks come from the counter
    // (dynamic) or the next stride (static). Each span of one batch inside a chunk runs
    // with that batch's arrays and its closest/any-hit kind fixed for the whole wave.
    const int64_t C = MODE == kModeWaveStatic ? 64 : (int64_t)chunk;
    const int64_t static_part = n_waves * C;
    int64_t next = wave_id * C;
    while (next < total) {
      // chunk [next, next + C) of the hand-out order; reversed, the chunk of ray numbers
      // that many from the end (rays still ascend inside it)
      int64_t pos, stop;
      if (reverse) {
        stop = total - next;
        pos = stop - C > 0 ? stop - C : 0;
      } else {
        pos = next;
        stop = next + C < total ? next + C : total;
      }
      while (pos < stop) {
        const int bi = __builtin_amdgcn_readfirstlane(find_batch(s_start, nb, pos));
        const BatchDev& B = batches[bi];  // uniform: scalar loads
        const GRays rays = global_view(B.rays);
        const GHits hits = global_view(B.hits);
        const int64_t bstart = s_start[bi];
        const int64_t bend = bstart + B.count;
        const int64_t gend = stop < bend ? stop : bend;
        const bool any = KIND == kKindAny || (KIND == kKindMixed && (B.flags & 1u));
        if (KIND != kKindClosest && any)
          trace_span<true, PRIM, LAYOUT, S, BLOCK>(sc, root, rays, hits, bstart, pos, gend, lane, st);
        else if (KIND != kKindAny)
          trace_span<false, PRIM, LAYOUT, S, BLOCK>(sc, root, rays, hits, bstart, pos, gend, lane, st);
        pos = gend;
      }
      if (MODE == kModeWaveStatic) {
        next += static_part;
      } else {
        unsigned long long x = 0;
        if (lane == 0) x = atomicAdd(ctr, (unsigned long long)C);
        next = (int64_t)__shfl((long long)x, 0, 64) + static_part;
      }
    }
  } else {
    const int64_t C = chunk;
    Lane L;
    bool active = false;
    // A finished ray's hit is stored when its lane takes the next ray (or at the end): on
    // gfx9 one counter tracks loads and stores, so a store after every finish would make
    // the next node fetch of the whole wave wait for it.
    bool pending = false;
    bool exhausted = false;
    int64_t pool_next = 0, pool_end = 0;  // wave-uniform
    int xr = 0;                           // ray ranges this wave found empty (per-XCD grabs)

    for (;;) {
      const unsigned long long idle = __ballot(!active);
      const int n_idle = __popcll(idle);
      if (n_idle == 64 && exhausted) {
        if (pending) finish(batches, sc, L);
        break;
      }
      if (n_idle >= refill && !exhausted) {
        const int rank = __popcll(idle & ((1ull << lane) - 1ull));
        const int64_t avail = pool_end - pool_next;
        int64_t g = -1;
        if (!active && rank < avail) g = pool_next + rank;
        if (n_idle > avail) {
          int64_t base, end;
          grab_chunk(ctr, C, total, xcd_remap != 0, xr, lane, base, end);
          if (reverse && base >= 0) {  // the same chunk counted from the end of the table
            const int64_t b = total - end;
            end = total - base;
            base = b;
          }
          if (base < 0) {
            exhausted = true;
            pool_next = pool_end = total;
          } else {
            if (!active && rank >= avail && base + (rank - avail) < end) g = base + (rank - avail);
            const int64_t taken = n_idle - avail;
            pool_next = base + taken < end ? base + taken : end;
            pool_end = end;
          }
        } else {
          pool_next += n_idle;
        }
        if (g >= 0) {
          if (pending) finish(batches, sc, L);
          pending = false;
          active = start_ray<PRIM, LAYOUT, S, BLOCK>(sc, root, batches, s_start, nb, g, L, st);
        }
      }
      if constexpr (LAYOUT != kLayoutRef) {
        // min/max slab test when every ray of the wave allows it (wave-uniform choice)
        const bool fast = __ballot(active && !L.regular) == 0ull;
        if (active) {
          bool more;
          if (KIND == kKindClosest)
            more = fast ? round_pair<false, PRIM, S, BLOCK, true, LAYOUT>(sc, L, st)
                        : round_pair<false, PRIM, S, BLOCK, false, LAYOUT>(sc, L, st);
          else if (KIND == kKindAny)
            more = fast ? round_pair<true, PRIM, S, BLOCK, true, LAYOUT>(sc, L, st)
                        : round_pair<true, PRIM, S, BLOCK, false, LAYOUT>(sc, L, st);
          else if (L.any)
            more = fast ? round_pair<true, PRIM, S, BLOCK, true, LAYOUT>(sc, L, st)
                        : round_pair<true, PRIM, S, BLOCK, false, LAYOUT>(sc, L, st);
          else
            more = fast ? round_pair<false, PRIM, S, BLOCK, true, LAYOUT>(sc, L, st)
                        : round_pair<false, PRIM, S, BLOCK, false, LAYOUT>(sc, L, st);
          if (!more) {
            active = false;
            pending = true;
          }
        }
      } else if (active) {
        if (!step_ref<PRIM, S, BLOCK>(sc, L, st)) {
          active = false;
          pending = true;
        }
      }
    }
  }
  if (MODE != kModeWaveStatic && lane == 0) {
    // The last wave out clears the counters for the next launch.
    const unsigned long long old = atomicAdd(ctr + 1, 1ull);
    if (old == (unsigned long long)n_waves - 1) {
      for (int k = 0; k < kCounters; ++k) atomicExch(ctr + k, 0ull);
    }
  }
}

template <int PRIM, int LAYOUT, int BLOCK, int MODE, int KIND>
int launch_w(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  dim3 g(bl.blocks), b(BLOCK);
  const size_t lds = sizeof(int64_t) * (size_t)bl.nb;  // batch start offsets
  int2* sp = reinterpret_cast<int2*>(bl.spill);
  switch (bl.stack_lds) {
    case 8:
      hipLaunchKernelGGL((k_trace_batch<PRIM, LAYOUT, 8, BLOCK, MODE, KIND>), g, b, lds, s, sc,
                         bl.table, bl.counters, sp, bl.spill_stride, bl.nb, bl.refill, bl.chunk,
                         bl.xcd_remap, bl.reverse);
      break;
    case 32:
      hipLaunchKernelGGL((k_trace_batch<PRIM, LAYOUT, 32, BLOCK, MODE, KIND>), g, b, lds, s, sc,
                         bl.table, bl.counters, sp, bl.spill_stride, bl.nb, bl.refill, bl.chunk,
                         bl.xcd_remap, bl.reverse);
      break;
    default:
      hipLaunchKernelGGL((k_trace_batch<PRIM, LAYOUT, 16, BLOCK, MODE, KIND>), g, b, lds, s, sc,
                         bl.table, bl.counters, sp, bl.spill_stride, bl.nb, bl.refill, bl.chunk,
                         bl.xcd_remap, bl.reverse);
      break;
  }
  return (int)hipGetLastError();
}

template <int PRIM, int LAYOUT, int BLOCK, int MODE>
int launch_k(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  switch (bl.kinds) {
    case kKindClosest:
      return launch_w<PRIM, LAYOUT, BLOCK, MODE, kKindClosest>(sc, bl, s);
    case kKindAny:
      return launch_w<PRIM, LAYOUT, BLOCK, MODE, kKindAny>(sc, bl, s);
    default:
      return launch_w<PRIM, LAYOUT, BLOCK, MODE, kKindMixed>(sc, bl, s);
  }
}

// refill 64: whole-wave chunks (chunk 0: static stride); otherwise per-lane refill.
template <int PRIM, int LAYOUT, int BLOCK>
int launch_s(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  if (bl.refill < 64) return launch_k<PRIM, LAYOUT, BLOCK, kModeLane>(sc, bl, s);
  return bl.chunk == 0 ? launch_k<PRIM, LAYOUT, BLOCK, kModeWaveStatic>(sc, bl, s)
                       : launch_k<PRIM, LAYOUT, BLOCK, kModeWave>(sc, bl, s);
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

// One launcher per triangle test and node layout (batch_<triangle>_<layout>.hip).
int launch_batch_ks_pair(const DevScene& sc, const BatchLaunch& bl, hipStream_t s);
int launch_batch_ks_ref(const DevScene& sc, const BatchLaunch& bl, hipStream_t s);
int launch_batch_wald_pair(const DevScene& sc, const BatchLaunch& bl, hipStream_t s);
int launch_batch_wald_ref(const DevScene& sc, const BatchLaunch& bl, hipStream_t s);
int launch_batch_ks_soa(const DevScene& sc, const BatchLaunch& bl, hipStream_t s);
int launch_batch_wald_soa(const DevScene& sc, const BatchLaunch& bl, hipStream_t s);

}  // namespace mh
