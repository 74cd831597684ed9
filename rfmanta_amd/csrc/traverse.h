// Single-ray traversal shared by the plain kernel (trace.hip) and the batch kernel
// (batch.hip): DynBVH::intersect on a one-ray packet (Model/Groups/DynBVH.cc:101-404).
//
// PAIR layout: inner nodes are 64-byte PairNodes holding both child boxes, so one fetch
//   tests both children. The child the reference visits first (direction sign on the
//   split axis for closest-hit, isLeftCheaper for any-hit, DynBVH.cc:386-398) is entered
//   and the other is pushed with its entry distance; a popped entry is entered only if
//   that distance is still <= minT, which is exactly the reference's later slab test of
//   that child with the updated minT.
// REF layout: the reference's 32-byte records; each node's own box is tested when it is
//   visited, as firstIntersects does.
// The stack is a per-lane ring of S entries in LDS ([slot][lane], conflict free) whose
// oldest entries spill to a per-thread global area when the ring is full.
#pragma once

#include "mh_device.h"
#include "trace_common.h"

namespace mh {

// Stack entry: node reference in the low 32 bits, entry distance bits in the high 32.
__device__ __forceinline__ unsigned long long stack_entry(int ref, float tn) {
  return (unsigned long long)(unsigned)ref | ((unsigned long long)__float_as_uint(tn) << 32);
}

template <int S, int BLOCK>
struct LaneStack {
  static_assert((S & (S - 1)) == 0, "ring size is a power of two");
  MH_LDS unsigned long long* ring;     // LDS [S][BLOCK], this lane's column
  MH_GLOBAL unsigned long long* spill; // global [depth][spill_stride], this thread's column
  int64_t spill_stride;
  int sp = 0, base = 0;  // entries [base, sp) are in the ring, [0, base) spilled

  __device__ __forceinline__ LaneStack(unsigned long long* lds, unsigned long long* sp_area,
                                       int64_t stride, int64_t gtid, int tid)
      : ring((MH_LDS unsigned long long*)lds + tid),
        spill((MH_GLOBAL unsigned long long*)sp_area + gtid), spill_stride(stride) {}

  __device__ __forceinline__ void reset() { sp = base = 0; }
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ void push(int ref, float tn) {
    if (sp - base == S) {
      spill[(int64_t)base * spill_stride] = ring[(base & (S - 1)) * BLOCK];
      ++base;
    }
    ring[(sp & (S - 1)) * BLOCK] = stack_entry(ref, tn);
    ++sp;
  }
  __device__ __forceinline__ void pop(int& ref, float& tn) {
    --sp;
    unsigned long long e;
    if (sp < base) {
      base = sp;
      e = spill[(int64_t)sp * spill_stride];
    } else {
      e = ring[(sp & (S - 1)) * BLOCK];
    }
    ref = (int)(unsigned)e;
    tn = __uint_as_float((unsigned)(e >> 32));
  }
};

// Result of one ray: the triangle slot of the closest (or first, any-hit) accepted hit,
// or -1, with minT and the barycentrics of that hit.
struct RayHit {
  float minT;
  int best;
  float u, v;
};

// Leaf: the triangles of slots [first, first+cnt) in order; returns true when an any-hit
// ray is done.
template <bool ANY, int PRIM, bool STATS>
__device__ __forceinline__ bool leaf(const DevScene& sc, int first, int cnt, const Ray& r,
                                     RayHit& h, unsigned long long& ntri) {
  for (int k = 0; k < cnt; ++k) {
    if (STATS) ++ntri;
    float t, u, v;
    if (prim_test<PRIM>(sc, first + k, r, h.minT, t, u, v)) {
      h.minT = t;
      h.best = first + k;
      h.u = u;
      h.v = v;
      if (ANY) return true;
    }
  }
  return false;
}

// The four 16-byte words of PairNode `cur`: one record of the pair table (kLayoutPair), or
// one entry of each of the four field arrays (kLayoutSoa).
template <int LAYOUT>
__device__ __forceinline__ void fetch_pair(const DevScene& sc, int cur, float4& A, float4& B,
                                           float4& C, int4& M) {
  float4 Mf;
  if (LAYOUT == kLayoutSoa) {
    const unsigned n = sc.n_pairs;
    const MhBufRsrc t = record_table(sc.pairs_soa, 16, 4 * n);
    A = fetch_record(t, cur, 0);
    B = fetch_record(t, (int)(cur + n), 0);
    C = fetch_record(t, (int)(cur + 2 * n), 0);
    Mf = fetch_record(t, (int)(cur + 3 * n), 0);
  } else {
    const MhBufRsrc pt = record_table(sc.pairs, sizeof(PairNode), sc.n_pairs);
    A = fetch_record(pt, cur, 0);
    B = fetch_record(pt, cur, 16);
    C = fetch_record(pt, cur, 32);
    Mf = fetch_record(pt, cur, 48);
  }
  M = make_int4(__float_as_int(Mf.x), __float_as_int(Mf.y), __float_as_int(Mf.z),
                __float_as_int(Mf.w));
}

template <bool ANY, int PRIM, int S, int BLOCK, bool STATS, bool FAST, int LAYOUT = kLayoutPair>
__device__ __forceinline__ void trace_pair(const DevScene& sc, const RefNode& root, const Ray& r,
                                           RayHit& h, LaneStack<S, BLOCK>& st,
                                           unsigned long long& nbox, unsigned long long& ntri) {
  float tn;
  bool go = slab_v<FAST>(root.box.lo[0], root.box.hi[0], root.box.lo[1], root.box.hi[1],
                         root.box.lo[2], root.box.hi[2], r, h.minT, tn);
  if (STATS) ++nbox;
  int cur = sc.root_ref;
  st.reset();
  while (go) {
    if (cur >= 0) {
      float4 A, B, C;
      int4 M;
      fetch_pair<LAYOUT>(sc, cur, A, B, C, M);
      float tn0, tn1;
      const bool h0 = slab_v<FAST>(A.x, A.y, A.z, A.w, B.x, B.y, r, h.minT, tn0);
      const bool h1 = slab_v<FAST>(B.z, B.w, C.x, C.y, C.z, C.w, r, h.minT, tn1);
      if (STATS) nbox += 2;
      // Front child: isLeftCheaper (any-hit) or the direction's sign on the split axis.
      const bool one_first = ANY ? ((M.z >> 2) & 1) == 0 : __builtin_amdgcn_ubfe(r.dpos, M.z, 1) == 0;
      // Both hit: enter the front child, push the other. One hits: enter it. Written on
      // the hit masks, so the choice stays in scalar mask operations.
      if (h0 && h1) st.push(one_first ? M.x : M.y, one_first ? tn0 : tn1);
      if (h0 || h1) {
        cur = (h1 && (one_first || !h0)) ? M.y : M.x;
        continue;
      }
    } else {
      const int code = ~cur;
      const int first = code >> kLeafCountBits;
      int cnt = code & kLeafCountInline;
      if (cnt == kLeafCountInline) cnt = sc.big_count[first];
      if (leaf<ANY, PRIM, STATS>(sc, first, cnt, r, h, ntri)) return;
    }
    go = false;
    while (!st.empty()) {
      int ref;
      float t;
      st.pop(ref, t);
      if (t <= h.minT) {
        cur = ref;
        go = true;
        break;
      }
    }
  }
}

template <bool ANY, int PRIM, int S, int BLOCK, bool STATS, bool FAST>
__device__ __forceinline__ void trace_ref(const DevScene& sc, const Ray& r, RayHit& h,
                                          LaneStack<S, BLOCK>& st, unsigned long long& nbox,
                                          unsigned long long& ntri) {
  int cur = 0;
  st.reset();
  for (;;) {
    const RefNode nd = sc.refs[cur];
    float tn;
    if (STATS) ++nbox;
    if (slab_v<FAST>(nd.box.lo[0], nd.box.hi[0], nd.box.lo[1], nd.box.hi[1], nd.box.lo[2],
                     nd.box.hi[2], r, h.minT, tn)) {
      if (nd.children != 0) {
        if (leaf<ANY, PRIM, STATS>(sc, nd.child, nd.children, r, h, ntri)) return;
      } else {
        const int front = ANY ? ((nd.bits & 1) ? 0 : 1) : (((r.dpos >> (nd.axis & 3)) & 1) ? 0 : 1);
        st.push(nd.child + 1 - front, 0.f);
        cur = nd.child + front;
        continue;
      }
    }
    if (st.empty()) return;
    float t;
    st.pop(cur, t);
  }
}

// One ray of `rays`/`hits` at index i, traced and stored (masked rays are stored as the
// reference leaves them: prim MH_MASKED, t = -MAXT, occluded).
template <bool ANY, int PRIM, int LAYOUT, int S, int BLOCK, bool STATS>
__device__ __forceinline__ void trace_store(const DevScene& sc, const RefNode& root,
                                            const GRays& rays, const GHits& hits, int64_t i,
                                            LaneStack<S, BLOCK>& st, unsigned long long& nbox,
                                            unsigned long long& ntri) {
  if (rays.masked && rays.masked[i]) {
    store_hit(hits, i, -2, -kMaxT, 0.f, 0.f, true);
    return;
  }
  Ray r;
  load_ray(rays, i, r);
  RayHit h{rays.tmax[i], -1, 0.f, 0.f};
  // The min/max slab test when every ray of the wave allows it (wave-uniform choice).
  if (__ballot(!regular_ray(r, h.minT)) == 0ull) {
    if (LAYOUT == kLayoutRef)
      trace_ref<ANY, PRIM, S, BLOCK, STATS, true>(sc, r, h, st, nbox, ntri);
    else
      trace_pair<ANY, PRIM, S, BLOCK, STATS, true, LAYOUT>(sc, root, r, h, st, nbox, ntri);
  } else {
    if (LAYOUT == kLayoutRef)
      trace_ref<ANY, PRIM, S, BLOCK, STATS, false>(sc, r, h, st, nbox, ntri);
    else
      trace_pair<ANY, PRIM, S, BLOCK, STATS, false, LAYOUT>(sc, root, r, h, st, nbox, ntri);
  }
  const bool hit = h.best >= 0;
  store_hit(hits, i, hit ? sc.obj_ids[h.best] : -1, (ANY && hit) ? kEps : h.minT, h.u, h.v, hit);
}

}  // namespace mh
