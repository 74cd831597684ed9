// Traversal kernels for gfx950 (wave64).
//
// k_trace_ray: one ray per lane, DynBVH::intersect on a one-ray packet
//   (Model/Groups/DynBVH.cc:101-404) over 64-byte PairNodes with an LDS ring stack; the
//   traversal itself is traverse.h's, shared with the batch kernel.
//
// k_trace_packet: one RayPacket of up to 64 consecutive rays per wave, the reference's
//   packet recursion (intersectNode/firstIntersects/lastIntersects, DynBVH.cc:274-905,
//   scalar branch). Node records are read with uniform (scalar) loads, each lane tests
//   its own ray, and ballots give firstActive/lastActive. The interval-arithmetic
//   frustum test (DynBVH.cc:682-721) runs on wave-reduced ray bounds. The wave-uniform
//   stack of (node, firstActive) lives in LDS.
//
// Arithmetic follows the reference's scalar build operation for operation: no
// contraction (the library is compiled with -ffp-contract=off), IEEE division for
// 1/direction and the triangle reciprocal, and the reference's select-based min/max.

#include <hip/hip_runtime.h>

#include <cfloat>

#include "mh_device.h"
#include "trace_common.h"
#include "traverse.h"

namespace mh {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kPacketStack = 128;

// ---------------------------------------------------------------------------
// Single-ray traversal (traverse.h), grid-stride over the rays. Eight waves per SIMD (the
// scalar registers otherwise limit it to seven): more node fetches in flight per SIMD.
// ---------------------------------------------------------------------------
template <bool ANY, int PRIM, int LAYOUT, int S, bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_trace_ray(
    DevScene sc, DevRays rays, DevHits hits, int64_t n, int2* spill, int64_t spill_stride,
    unsigned long long* counters) {
  __shared__ unsigned long long s_ring[S * kBlock];
  const int tid = threadIdx.x;
  const int64_t gtid = (int64_t)blockIdx.x * kBlock + tid;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  unsigned long long nbox = 0, ntri = 0;
  const RefNode root = sc.refs[0];
  LaneStack<S, kBlock> st(s_ring, reinterpret_cast<unsigned long long*>(spill), spill_stride,
                          gtid, tid);
  const GRays gr = global_view(rays);
  const GHits gh = global_view(hits);
  for (int64_t ri = gtid; ri < n; ri += stride)
    trace_store<ANY, PRIM, LAYOUT, S, kBlock, STATS>(sc, root, gr, gh, ri, st, nbox, ntri);
  if (STATS && counters) {
    atomicAdd(counters + 0, nbox);
    atomicAdd(counters + 1, ntri);
  }
}

// ---------------------------------------------------------------------------
// Packet traversal (one RayPacket per wave)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int lowest(unsigned long long m) { return __builtin_ctzll(m); }
__device__ __forceinline__ int highest(unsigned long long m) { return 63 - __builtin_clzll(m); }

// The packet's IAData loop folds its rays in order with selects (DynBVH.cc:246-266):
//   min: acc = (acc < x) ? acc : x from FLT_MAX;  max: acc = (acc < x) ? x : acc from -FLT_MAX.
// The min fold takes a NaN (and restarts from the next ray after it) and keeps the later
// of two equal values (+0/-0); the max fold skips NaNs and keeps the earlier. A butterfly
// over (value, lane) pairs with those tie rules gives the fold's exact bits in any order.
__device__ __forceinline__ float fold_min(float x, bool contrib, int lane) {
  const unsigned long long nan = __ballot(contrib && x != x);
  const int j = nan ? highest(nan) : -1;  // the last NaN: the fold restarts after it
  float v = (contrib && lane > j) ? x : __builtin_inff();
  int p = (contrib && lane > j) ? lane : -1;
  for (int m = 32; m >= 1; m >>= 1) {
    const float y = __shfl_xor(v, m, 64);
    const int q = __shfl_xor(p, m, 64);
    const bool take = (y < v) || (!(v < y) && q > p);
    v = take ? y : v;
    p = take ? q : p;
  }
  if (j < 0) return (FLT_MAX < v) ? FLT_MAX : v;  // from the initial FLT_MAX
  if (p < 0) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
  return v;
}
__device__ __forceinline__ float fold_max(float x, bool contrib, int lane) {
  const bool use = contrib && x == x;
  float v = use ? x : -__builtin_inff();
  int p = use ? lane : 64;
  for (int m = 32; m >= 1; m >>= 1) {
    const float y = __shfl_xor(v, m, 64);
    const int q = __shfl_xor(p, m, 64);
    const bool take = (v < y) || (!(y < v) && q < p);
    v = take ? y : v;
    p = take ? q : p;
  }
  return (-FLT_MAX < v) ? v : -FLT_MAX;  // from the initial -FLT_MAX
}
__device__ __forceinline__ unsigned long long above(int b) {
  return b >= 63 ? 0ull : (~0ull << (b + 1));
}

struct IA {
  float min_rcp[3], max_rcp[3], min_org[3], max_org[3];
};

// Interval-arithmetic frustum exit of the scalar firstIntersects (DynBVH.cc:682-721).
__device__ __forceinline__ bool frustum_miss(const RefNode& nd, const IA& ia) {
  float tmin_f = 1e-5f, tmax_f = FLT_MAX;
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    float a = nd.box.lo[axis] - ia.max_org[axis];
    float b = nd.box.hi[axis] - ia.min_org[axis];
    const float ar0 = a * ia.min_rcp[axis];
    const float ar1 = a * ia.max_rcp[axis];
    const float br0 = b * ia.min_rcp[axis];
    const float br1 = b * ia.max_rcp[axis];
    a = (br0 < br1) ? br0 : br1;
    a = (a < ar1) ? a : ar1;
    a = (a < ar0) ? a : ar0;
    b = (br0 < br1) ? br1 : br0;
    b = (b < ar1) ? ar1 : b;
    b = (b < ar0) ? ar0 : b;
    tmin_f = (tmin_f < a) ? a : tmin_f;
    tmax_f = (tmax_f > b) ? b : tmax_f;
  }
  return tmin_f > tmax_f;
}

template <bool ANY, int PRIM, bool STATS>
__global__ __launch_bounds__(kBlock) void k_trace_packet(DevScene sc, DevRays rays,
                                                         DevHits hits, int64_t n, int psize,
                                                         unsigned long long* counters) {
  __shared__ int2 stk[kWaves][kPacketStack];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t pk = (int64_t)blockIdx.x * kWaves + wv;
  const int64_t first = pk * psize;
  if (first >= n) return;
  const int cnt = (int)((n - first) < psize ? (n - first) : psize);
  const bool in = lane < cnt;
  const int64_t ri = first + lane;

  Ray r;
  float minT;
  int st;  // 0 not hit, 1 hit, -1 masked, -3 outside the packet
  if (in) {
    load_ray(rays, ri, r);
    const bool m = rays.masked && rays.masked[ri];
    st = m ? -1 : 0;
    minT = m ? -kMaxT : rays.tmax[ri];
  } else {
    r.o[0] = r.o[1] = r.o[2] = 0.f;
    r.d[0] = r.d[1] = r.d[2] = 1.f;
    r.inv[0] = r.inv[1] = r.inv[2] = 1.f;
    st = -3;
    minT = -kMaxT;
  }
  // DynBVH::intersect IAData over the packet's unmasked rays (DynBVH.cc:246-266)
  IA ia;
  const bool contrib = st == 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ia.min_rcp[a] = fold_min(r.inv[a], contrib, lane);
    ia.max_rcp[a] = fold_max(r.inv[a], contrib, lane);
    ia.min_org[a] = fold_min(r.o[a], contrib, lane);
    ia.max_org[a] = fold_max(r.o[a], contrib, lane);
  }

  int pos = -1;
  float hu = 0.f, hv = 0.f;
  unsigned long long nbox = 0, ntri = 0;
  int sp = 0;
  stk[wv][sp++] = make_int2(0, 0);
  while (sp > 0) {
    --sp;
    const int2 e = stk[wv][sp];
    const int nid = uni(e.x);
    const int b = uni(e.y);
    const RefNode nd = sc.refs[nid];
    const bool elig = in && lane >= b && st != -1 && !(ANY && st == 1);
    const unsigned long long E = __ballot(elig);
    if (E == 0) continue;
    const int rb = lowest(E);
    float tn;
    const bool hb = elig && slab(nd.box.lo[0], nd.box.hi[0], nd.box.lo[1], nd.box.hi[1],
                                 nd.box.lo[2], nd.box.hi[2], r, minT, tn);
    const unsigned long long H = __ballot(hb);
    if (!((H >> rb) & 1ull)) {
      if (frustum_miss(nd, ia)) {
        if (STATS) nbox += 1;
        continue;
      }
    }
    if (H == 0) {
      if (STATS) nbox += __popcll(E);
      continue;
    }
    const int fa = lowest(H);
    if (STATS) nbox += __popcll(E & (fa >= 63 ? ~0ull : ((2ull << fa) - 1)));
    if (nd.children != 0) {
      const unsigned long long Lm = H & above(b);
      const int la = Lm ? highest(Lm) : b;
      int sb = fa, se = la + 1;
      for (int k = 0; k < nd.children; ++k) {
        const int slot = nd.child + k;
        if (STATS) ntri += (unsigned long long)(se - sb);
        const bool act = lane >= sb && lane < se;
        if (act) {
          float t, u, v;
          if (prim_test<PRIM>(sc, slot, r, minT, t, u, v)) {
            minT = t;
            st = 1;
            pos = slot;
            hu = u;
            hv = v;
          }
        }
        if (ANY) {
          // DynBVH.cc:329-374: opaque hits end the ray (minT = T_EPSILON); the
          // subpacket shrinks to the rays still unoccluded.
          const bool hit_here = act && st == 1;
          if (hit_here) minT = kEps;
          if (__ballot(hit_here)) {
            const unsigned long long alive = __ballot(act && st == 0);
            if (!alive) break;
            sb = lowest(alive);
            se = highest(alive) + 1;
          }
        }
      }
    } else {
      int front;
      if (ANY) {
        front = (nd.bits & 1) ? 0 : 1;
      } else {
        const float da = pick3(nd.axis, r.d[0], r.d[1], r.d[2]);
        front = (__shfl(da, fa, 64) > 0.f) ? 0 : 1;
      }
      stk[wv][sp++] = make_int2(nd.child + 1 - front, fa);
      stk[wv][sp++] = make_int2(nd.child + front, fa);
    }
  }
  if (in) {
    const int32_t prim = st == 1 ? sc.obj_ids[pos] : (st == -1 ? -2 : -1);
    store_hit(hits, ri, prim, minT, hu, hv, st != 0);
  }
  if (STATS && counters && lane == 0) {
    atomicAdd(counters + 0, nbox);
    atomicAdd(counters + 1, ntri);
  }
}

template <bool ANY, int PRIM, int LAYOUT, bool STATS>
int launch_ray_s(const DevScene& sc, const DevRays& rays, const DevHits& hits, int64_t n,
                 const LaunchCfg& cfg, hipStream_t s) {
  unsigned long long* c = reinterpret_cast<unsigned long long*>(cfg.counters);
  int2* sp = reinterpret_cast<int2*>(cfg.spill);
  dim3 g(cfg.blocks), b(kBlock);
  switch (cfg.stack_lds) {
    case 8:
      hipLaunchKernelGGL((k_trace_ray<ANY, PRIM, LAYOUT, 8, STATS>), g, b, 0, s, sc, rays, hits, n, sp,
                         cfg.spill_stride, c);
      break;
    case 32:
      hipLaunchKernelGGL((k_trace_ray<ANY, PRIM, LAYOUT, 32, STATS>), g, b, 0, s, sc, rays, hits, n, sp,
                         cfg.spill_stride, c);
      break;
    default:
      hipLaunchKernelGGL((k_trace_ray<ANY, PRIM, LAYOUT, 16, STATS>), g, b, 0, s, sc, rays, hits, n, sp,
                         cfg.spill_stride, c);
      break;
  }
  return (int)hipGetLastError();
}

template <bool ANY, int PRIM>
int launch_any_prim(const DevScene& sc, const DevRays& rays, const DevHits& hits, int64_t n,
                    const LaunchCfg& cfg, hipStream_t s) {
  const bool stats = cfg.counters != nullptr;
  if (cfg.packet_size > 1) {
    const int64_t packets = (n + cfg.packet_size - 1) / cfg.packet_size;
    const int64_t blocks = (packets + kWaves - 1) / kWaves;
    unsigned long long* c = reinterpret_cast<unsigned long long*>(cfg.counters);
    if (stats)
      hipLaunchKernelGGL((k_trace_packet<ANY, PRIM, true>), dim3((unsigned)blocks), dim3(kBlock),
                         0, s, sc, rays, hits, n, cfg.packet_size, c);
    else
      hipLaunchKernelGGL((k_trace_packet<ANY, PRIM, false>), dim3((unsigned)blocks),
                         dim3(kBlock), 0, s, sc, rays, hits, n, cfg.packet_size, c);
    return (int)hipGetLastError();
  }
  return stats ? launch_ray_s<ANY, PRIM, kLayoutPair, true>(sc, rays, hits, n, cfg, s)
               : launch_ray_s<ANY, PRIM, kLayoutPair, false>(sc, rays, hits, n, cfg, s);
}

}  // namespace

int launch_trace(const DevScene& sc, const DevRays& rays, const DevHits& hits, int64_t n,
                 const LaunchCfg& cfg, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (cfg.any_hit) {
    return sc.prim_kind == 1 ? launch_any_prim<true, 1>(sc, rays, hits, n, cfg, s)
                             : launch_any_prim<true, 0>(sc, rays, hits, n, cfg, s);
  }
  return sc.prim_kind == 1 ? launch_any_prim<false, 1>(sc, rays, hits, n, cfg, s)
                           : launch_any_prim<false, 0>(sc, rays, hits, n, cfg, s);
}

}  // namespace mh
