// Per-triangle records computed from the current vertices, shared by the refit
// (triangle records in leaf order) and the shading pass (Wald face normals).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mh {
namespace {

struct Tri {
  float p[3][3];
};

__device__ __forceinline__ Tri load_tri(const float* verts, const uint32_t* idx, int tri) {
  Tri t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t vi = idx[3 * (size_t)tri + k];
    t.p[k][0] = verts[3 * (size_t)vi + 0];
    t.p[k][1] = verts[3 * (size_t)vi + 1];
    t.p[k][2] = verts[3 * (size_t)vi + 2];
  }
  return t;
}

// Kensler-Shirley record: p0, e0 = p1 - p0, e1 = p0 - p2, n = Cross(e0, e1)
// (KenslerShirleyTriangle.cc:341-345).
__device__ __forceinline__ void ks_record(const Tri& t, float* o) {
  const float e0x = t.p[1][0] - t.p[0][0], e0y = t.p[1][1] - t.p[0][1],
              e0z = t.p[1][2] - t.p[0][2];
  const float e1x = t.p[0][0] - t.p[2][0], e1y = t.p[0][1] - t.p[2][1],
              e1z = t.p[0][2] - t.p[2][2];
  o[0] = t.p[0][0];
  o[1] = t.p[0][1];
  o[2] = t.p[0][2];
  o[3] = e0x;
  o[4] = e0y;
  o[5] = e0z;
  o[6] = e1x;
  o[7] = e1y;
  o[8] = e1z;
  o[9] = e0y * e1z - e0z * e1y;
  o[10] = e0z * e1x - e0x * e1z;
  o[11] = e0x * e1y - e0y * e1x;
}

// WaldTriangle::setPoints(v0, v1, v2) (WaldTriangle.cc:64-127).
__device__ __forceinline__ void wald_record(const Tri& t, float* o) {
  const float* p1 = t.p[0];
  const float* p2 = t.p[1];
  const float* p3 = t.p[2];
  const float ax = p2[0] - p1[0], ay = p2[1] - p1[1], az = p2[2] - p1[2];
  const float bx = p3[0] - p1[0], by = p3[1] - p1[1], bz = p3[2] - p1[2];
  float nrm[3] = {ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx};
  const float l = sqrtf(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  const float scale = 1 / l;
  nrm[0] *= scale;
  nrm[1] *= scale;
  nrm[2] *= scale;
  int n = 0;
  if (fabsf(nrm[1]) > fabsf(nrm[n])) n = 1;
  if (fabsf(nrm[2]) > fabsf(nrm[n])) n = 2;
  const int u = n == 0 ? 1 : (n == 1 ? 2 : 0);
  const int v = n == 0 ? 2 : (n == 1 ? 0 : 1);
  const float inv_nk = 1 / nrm[n];
  const float n_u = nrm[u] * inv_nk;
  const float n_v = nrm[v] * inv_nk;
  const float n_d = p1[u] * n_u + p1[v] * n_v + p1[n];
  float c_nu = +p2[v] - p1[v];
  float c_nv = -p2[u] + p1[u];
  float c_d = -(p1[u] * c_nu + p1[v] * c_nv);
  float s = 1.f / (p3[u] * c_nu + p3[v] * c_nv + c_d);
  c_nu *= s;
  c_nv *= s;
  c_d *= s;
  float b_nu = +p3[v] - p1[v];
  float b_nv = -p3[u] + p1[u];
  float b_d = -(p1[u] * b_nu + p1[v] * b_nv);
  s = 1.f / (p2[u] * b_nu + p2[v] * b_nv + b_d);
  b_nu *= s;
  b_nv *= s;
  b_d *= s;
  o[0] = n_u;
  o[1] = n_v;
  o[2] = n_d;
  o[3] = __int_as_float(n);
  o[4] = b_nu;
  o[5] = b_nv;
  o[6] = b_d;
  o[7] = nrm[n];
  o[8] = c_nu;
  o[9] = c_nv;
  o[10] = c_d;
  o[11] = 0.f;
}

// The 48-byte record of triangle t at slot `slot` of the triangle array (prim_kind 1 =
// Wald, else Kensler-Shirley), as three 16-byte stores.
__device__ __forceinline__ void store_tri_record(float* tris, int prim_kind, int64_t slot,
                                                 const Tri& t) {
  float o[12];
  if (prim_kind == 1)
    wald_record(t, o);
  else
    ks_record(t, o);
  float4* dst = reinterpret_cast<float4*>(tris + 12 * slot);
  dst[0] = make_float4(o[0], o[1], o[2], o[3]);
  dst[1] = make_float4(o[4], o[5], o[6], o[7]);
  dst[2] = make_float4(o[8], o[9], o[10], o[11]);
}

}  // namespace
}  // namespace mh
