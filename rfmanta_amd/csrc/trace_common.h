// Device helpers shared by the traversal kernels: the ray record, the reference's slab
// test (DynBVH::firstIntersects scalar loop, Model/Groups/DynBVH.cc:666-680), the
// Kensler-Shirley and Wald tests (KenslerShirleyTriangle.cc:336-377,
// WaldTriangle.cc:489-555) followed by RayPacket::hit (Interface/RayPacket.h:531-542),
// and ray/hit SoA access. Everything is binary32 without contraction (the library is
// compiled with -ffp-contract=off) and IEEE division.
#pragma once

#include <hip/hip_runtime.h>

#include "mh_device.h"

// Structured-buffer load (buffer_load_dwordx4 ... idxen): the record index goes to the
// address unit as it is, so fetching part of record i of a table costs no vector address
// arithmetic. Indices at or past the descriptor's record count read zeros.
typedef int MhBufRsrc __attribute__((ext_vector_type(4)));
typedef float MhBufF4 __attribute__((ext_vector_type(4)));
__device__ MhBufF4 mh_struct_buffer_load_v4f32(MhBufRsrc rsrc, int index, int offset, int soffset,
                                               int aux) __asm("llvm.amdgcn.struct.buffer.load.v4f32");

namespace mh {

constexpr float kEps = 1.e-4f;   // T_EPSILON

// Descriptor of a table of `count` records of `stride` bytes (stride < 16384) at `base`:
// 48-bit base, stride in word 1 bits 16-29, record count, gfx9 word 3 (32-bit data
// format, no swizzle, no thread-id indexing).
__device__ __forceinline__ MhBufRsrc record_table(const void* base, unsigned stride,
                                                  unsigned count) {
  const unsigned long long a = (unsigned long long)base;
  MhBufRsrc r;
  r.x = (int)(unsigned)a;
  r.y = (int)(((unsigned)(a >> 32) & 0xFFFFu) | (stride << 16));
  r.z = (int)count;
  r.w = 0x00020000;
  return r;
}

// 16 bytes at byte `offset` (a constant) of record `index`.
__device__ __forceinline__ float4 fetch_record(const MhBufRsrc& t, int index, int offset) {
  const MhBufF4 v = mh_struct_buffer_load_v4f32(t, index, offset, 0, 0);
  return make_float4(v.x, v.y, v.z, v.w);
}
constexpr float kMaxT = 1.e19f;  // MAXT

// Explicit address spaces. Pointers read from memory (batch tables) are otherwise
// generic, and the compiler then uses flat instructions; an LDS and a global access on
// the two sides of a branch can also be merged into one flat access. Typed pointers keep
// every access a global_* or ds_* instruction.
#define MH_GLOBAL __attribute__((address_space(1)))
#define MH_LDS __attribute__((address_space(3)))

// Ray and hit arrays as global-memory views.
struct GRays {
  const MH_GLOBAL float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  const MH_GLOBAL uint8_t* masked;
};
struct GHits {
  MH_GLOBAL int32_t* prim;
  MH_GLOBAL float *t, *u, *v;
  MH_GLOBAL uint8_t* occluded;
};
static __device__ __forceinline__ GRays global_view(const DevRays& r) {
  using F = const MH_GLOBAL float*;
  return GRays{(F)r.ox, (F)r.oy, (F)r.oz, (F)r.dx, (F)r.dy, (F)r.dz, (F)r.tmax,
               (const MH_GLOBAL uint8_t*)r.masked};
}
static __device__ __forceinline__ GHits global_view(const DevHits& h) {
  using F = MH_GLOBAL float*;
  return GHits{(MH_GLOBAL int32_t*)h.prim, (F)h.t, (F)h.u, (F)h.v, (MH_GLOBAL uint8_t*)h.occluded};
}

// a, b or c for k = 0, 1, 2, chosen with bit masks so that the compiler cannot turn the
// choice into an indexed (scratch) access.
static __device__ __forceinline__ float pick3(int k, float a, float b, float c) {
  const unsigned m0 = 0u - (unsigned)(k == 0), m1 = 0u - (unsigned)(k == 1),
                 m2 = 0u - (unsigned)(k == 2);
  return __uint_as_float((__float_as_uint(a) & m0) | (__float_as_uint(b) & m1) |
                         (__float_as_uint(c) & m2));
}

struct Ray {
  float o[3], d[3], inv[3];
  // bit k and bit k + 4 set iff d[k] > 0 (front child choice on split axis k); the copy
  // lets a PairNode's meta (axis | isLeftCheaper << 2) index the bit directly.
  unsigned dpos;
};

// Slab test of DynBVH::firstIntersects (scalar loop, DynBVH.cc:666-680).
static __device__ __forceinline__ bool slab(float lx, float hx, float ly, float hy, float lz, float hz,
                                     const Ray& r, float minT, float& tnear) {
  float tmin = kEps, tmax = minT;
  float t0 = (lx - r.o[0]) * r.inv[0];
  float t1 = (hx - r.o[0]) * r.inv[0];
  float nr = (t0 < t1) ? t0 : t1;
  float fr = (t0 < t1) ? t1 : t0;
  tmin = (tmin < nr) ? nr : tmin;
  tmax = (fr < tmax) ? fr : tmax;
  t0 = (ly - r.o[1]) * r.inv[1];
  t1 = (hy - r.o[1]) * r.inv[1];
  nr = (t0 < t1) ? t0 : t1;
  fr = (t0 < t1) ? t1 : t0;
  tmin = (tmin < nr) ? nr : tmin;
  tmax = (fr < tmax) ? fr : tmax;
  t0 = (lz - r.o[2]) * r.inv[2];
  t1 = (hz - r.o[2]) * r.inv[2];
  nr = (t0 < t1) ? t0 : t1;
  fr = (t0 < t1) ? t1 : t0;
  tmin = (tmin < nr) ? nr : tmin;
  tmax = (fr < tmax) ? fr : tmax;
  tnear = tmin;
  return tmin <= tmax;
}

// The same test with min/max instructions. When the origin and the reciprocal direction
// are finite (and the box coordinates are), no t is NaN, and then each select of the
// reference's form equals IEEE min/max: both return one of their inputs, and zero signs
// cannot matter because tmin starts at T_EPSILON > 0. Results are identical bit for bit.
static __device__ __forceinline__ bool slab_fast(float lx, float hx, float ly, float hy, float lz,
                                          float hz, const Ray& r, float minT, float& tnear) {
  const float ax = (lx - r.o[0]) * r.inv[0], bx = (hx - r.o[0]) * r.inv[0];
  const float ay = (ly - r.o[1]) * r.inv[1], by = (hy - r.o[1]) * r.inv[1];
  const float az = (lz - r.o[2]) * r.inv[2], bz = (hz - r.o[2]) * r.inv[2];
  const float tmin = fmaxf(fmaxf(fmaxf(kEps, fminf(ax, bx)), fminf(ay, by)), fminf(az, bz));
  const float tfar = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
  tnear = tmin;
  // tmin <= min(minT, tfar), written as two comparisons (no NaN here), so minT, which
  // the loop carries, needs no canonicalising min.
  return (tmin <= tfar) & (tmin <= minT);
}

template <bool FAST>
static __device__ __forceinline__ bool slab_v(float lx, float hx, float ly, float hy, float lz,
                                       float hz, const Ray& r, float minT, float& tnear) {
  return FAST ? slab_fast(lx, hx, ly, hy, lz, hz, r, minT, tnear)
              : slab(lx, hx, ly, hy, lz, hz, r, minT, tnear);
}

// A ray for which slab_fast is exact: finite origin and reciprocal direction, minT not NaN.
static __device__ __forceinline__ bool regular_ray(const Ray& r, float minT) {
  bool ok = !(minT != minT);
#pragma unroll
  for (int k = 0; k < 3; ++k) ok = ok && fabsf(r.o[k]) <= 3.4028235e38f && fabsf(r.inv[k]) <= 3.4028235e38f;
  return ok;
}

// KenslerShirleyTriangle::intersect, scalar branch (KenslerShirleyTriangle.cc:336-377),
// then RayPacket::hit (Interface/RayPacket.h:531-542). Slot data: p0, e0, e1, n.
static __device__ __forceinline__ bool ks_test(const float4 A, const float4 B, const float4 C,
                                        const Ray& r, float minT, float& t, float& u, float& v) {
  const float n0 = C.y, n1 = C.z, n2 = C.w;
  const float det = n0 * r.d[0] + n1 * r.d[1] + n2 * r.d[2];
  const float rcp = 1.0f / det;
  const float g0 = A.x - r.o[0], g1 = A.y - r.o[1], g2 = A.z - r.o[2];
  const float tt = (n0 * g0 + n1 * g1 + n2 * g2) * rcp;
  if (tt > minT - kEps || tt < kEps) return false;
  const float i0 = g1 * r.d[2] - g2 * r.d[1];
  const float i1 = g2 * r.d[0] - g0 * r.d[2];
  const float i2 = g0 * r.d[1] - g1 * r.d[0];
  const float uu = (i0 * B.z + i1 * B.w + i2 * C.x) * rcp;
  if (uu < 0.0f) return false;
  const float vv = (i0 * A.w + i1 * B.x + i2 * B.y) * rcp;
  if (uu + vv > 1.0f || vv < 0.0f) return false;
  if (!(tt > kEps && tt < minT)) return false;
  t = tt;
  u = uu;
  v = vv;
  return true;
}

// WaldTriangle::intersect, scalar branch (WaldTriangle.cc:489-555), then RayPacket::hit.
static __device__ __forceinline__ bool wald_test(const float4 A, const float4 B, const float4 C,
                                          const Ray& r, float minT, float& t, float& u,
                                          float& v) {
  // Projection axis k and its successors ku = (k+1)%3, kv = (k+2)%3.
  const int k = __float_as_int(A.w);
  const float dk = pick3(k, r.d[0], r.d[1], r.d[2]);
  const float dku = pick3(k, r.d[1], r.d[2], r.d[0]);
  const float dkv = pick3(k, r.d[2], r.d[0], r.d[1]);
  const float ok = pick3(k, r.o[0], r.o[1], r.o[2]);
  const float oku = pick3(k, r.o[1], r.o[2], r.o[0]);
  const float okv = pick3(k, r.o[2], r.o[0], r.o[1]);
  const float nd0 = A.x * dku + A.y * dkv + dk;
  const float nd = 1.f / nd0;
  const float f0 = A.z - (ok + A.x * oku + A.y * okv);
  const float f = f0 * nd;
  if (f < kEps || f > minT) return false;
  const float hu = oku + f * dku;
  const float hv = okv + f * dkv;
  const float lambda = B.z + hu * B.x + hv * B.y;
  if (lambda < 0.f) return false;
  const float mue = C.z + hu * C.x + hv * C.y;
  if (mue < 0.f || mue + lambda > 1.f) return false;
  if (!(f > kEps && f < minT)) return false;
  t = f;
  u = lambda;
  v = mue;
  return true;
}

template <int PRIM>
static __device__ __forceinline__ bool prim_test(const DevScene& sc, int slot, const Ray& r,
                                                 float minT, float& t, float& u, float& v) {
  const MhBufRsrc tt = record_table(sc.tris, 12 * sizeof(float), sc.n_slots);
  float4 A = fetch_record(tt, slot, 0), B = fetch_record(tt, slot, 16),
         C = fetch_record(tt, slot, 32);
  // All 48 bytes are fetched before the first rejection test (one round trip) instead of
  // the compiler deferring the edge loads past the distance test.
  // (Input operands only: the values stay in the registers they were loaded into.)
  asm volatile("" ::"v"(A.x), "v"(A.y), "v"(A.z), "v"(A.w), "v"(B.x), "v"(B.y), "v"(B.z),
               "v"(B.w), "v"(C.x), "v"(C.y), "v"(C.z), "v"(C.w));
  if (PRIM == 0) return ks_test(A, B, C, r, minT, t, u, v);
  return wald_test(A, B, C, r, minT, t, u, v);
}

static __device__ __forceinline__ void load_ray(const GRays& rays, int64_t i, Ray& r) {
  r.o[0] = rays.ox[i];
  r.o[1] = rays.oy[i];
  r.o[2] = rays.oz[i];
  r.d[0] = rays.dx[i];
  r.d[1] = rays.dy[i];
  r.d[2] = rays.dz[i];
  // RayPacket::actualComputeInverseDirections, scalar branch (RayPacket.cc:105-109)
  r.inv[0] = 1.0f / r.d[0];
  r.inv[1] = 1.0f / r.d[1];
  r.inv[2] = 1.0f / r.d[2];
  r.dpos = (unsigned)(r.d[0] > 0.f) | ((unsigned)(r.d[1] > 0.f) << 1) |
           ((unsigned)(r.d[2] > 0.f) << 2);
  r.dpos |= r.dpos << 4;
}

static __device__ __forceinline__ void store_hit(const GHits& h, int64_t i, int32_t prim, float t,
                                          float u, float v, bool occ) {
  if (h.prim) h.prim[i] = prim;
  if (h.t) h.t[i] = t;
  if (h.u) h.u[i] = u;
  if (h.v) h.v[i] = v;
  if (h.occluded) h.occluded[i] = occ ? 1 : 0;
}
static __device__ __forceinline__ void load_ray(const DevRays& rays, int64_t i, Ray& r) {
  load_ray(global_view(rays), i, r);
}
static __device__ __forceinline__ void store_hit(const DevHits& h, int64_t i, int32_t prim, float t,
                                          float u, float v, bool occ) {
  store_hit(global_view(h), i, prim, t, u, v, occ);
}


}  // namespace mh
