// Shading inputs of hit rays on gfx950, one lane per ray: what Manta's materials ask a
// hit primitive for after intersect.
//   hit position     RayPacket::actualComputeHitPositions (Interface/RayPacket.cc:114-146)
//   normal           KenslerShirleyTriangle::computeNormal (KenslerShirleyTriangle.cc:381-405),
//                    WaldTriangle::computeNormal (WaldTriangle.cc:558-590)
//   geometric normal KenslerShirleyTriangle::computeGeometricNormal (:407-420),
//                    WaldTriangle::computeGeometricNormal (WaldTriangle.cc:592-609)
//   texture coords   computeTexCoords2 (KenslerShirleyTriangle.cc:85-123, WaldTriangle.cc:131-169)
// u and v are the hit's scratchpad weights of vertices 1 and 2; the third weight is
// Real c = (1.0 - a - b), evaluated in double as the reference's expression is. Wald
// normals come from the Wald set-up of the current vertices (the same arithmetic as the
// triangle records). Every access is per ray and coalesced except the gathers of the hit
// triangle's indices and vertices.

#include <hip/hip_runtime.h>

#include "mh_device.h"
#include "prim_setup.h"

namespace mh {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_shade(DevShade s) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= s.n) return;
  float out[12];
  for (int k = 0; k < 12; ++k) out[k] = 0.f;
  const int id = s.prim[r];
  if (id >= 0 && id < s.n_tris) {
    const float t = s.t[r];
    const float o[3] = {s.rays.ox[r], s.rays.oy[r], s.rays.oz[r]};
    const float d[3] = {s.rays.dx[r], s.rays.dy[r], s.rays.dz[r]};
    for (int j = 0; j < 3; ++j) out[j] = o[j] + d[j] * t;
    const float a = s.u[r], b = s.v[r];
    const float c = (float)(1.0 - (double)a - (double)b);
    const Tri tri = load_tri(s.verts, s.idx, id);
    float face[3];
    if (s.prim_kind == 1) {
      float w[12];
      wald_record(tri, w);  // n_u, n_v, n_d, k | b_* , n_k | c_*
      const int k = __float_as_int(w[3]);
      const int ku = (k == 2) ? 0 : k + 1, kv = (k == 0) ? 2 : k - 1;
      const float nk = w[7];
      face[0] = face[1] = face[2] = 0.f;
      for (int j = 0; j < 3; ++j) {
        face[j] = j == k ? nk : (j == ku ? nk * w[0] : (j == kv ? nk * w[1] : face[j]));
      }
    } else {
      const float e0[3] = {tri.p[1][0] - tri.p[0][0], tri.p[1][1] - tri.p[0][1],
                           tri.p[1][2] - tri.p[0][2]};
      const float e1[3] = {tri.p[2][0] - tri.p[0][0], tri.p[2][1] - tri.p[0][1],
                           tri.p[2][2] - tri.p[0][2]};
      face[0] = e0[1] * e1[2] - e0[2] * e1[1];
      face[1] = e0[2] * e1[0] - e0[0] * e1[2];
      face[2] = e0[0] * e1[1] - e0[1] * e1[0];
    }
    for (int j = 0; j < 3; ++j) out[6 + j] = face[j];
    if (s.normals) {
      const uint32_t i0 = s.normal_idx[3 * (size_t)id], i1 = s.normal_idx[3 * (size_t)id + 1],
                     i2 = s.normal_idx[3 * (size_t)id + 2];
      for (int j = 0; j < 3; ++j)
        out[3 + j] = (s.normals[3 * (size_t)i1 + j] * a + s.normals[3 * (size_t)i2 + j] * b) +
                     s.normals[3 * (size_t)i0 + j] * c;
    } else {
      for (int j = 0; j < 3; ++j) out[3 + j] = face[j];
    }
    const uint32_t x0 = s.tex ? s.tex_idx[3 * (size_t)id] : 0xFFFFFFFFu;
    if (x0 == 0xFFFFFFFFu) {  // Mesh::kNoTextureIndex (Model/Groups/Mesh.cc:14)
      out[9] = a;
      out[10] = b;
      out[11] = c;
    } else {
      const uint32_t x1 = s.tex_idx[3 * (size_t)id + 1], x2 = s.tex_idx[3 * (size_t)id + 2];
      for (int j = 0; j < 3; ++j)
        out[9 + j] = (s.tex[3 * (size_t)x1 + j] * a + s.tex[3 * (size_t)x2 + j] * b) +
                     s.tex[3 * (size_t)x0 + j] * c;
    }
  }
  for (int k = 0; k < 12; ++k)
    if (s.out[k]) s.out[k][r] = out[k];
}

}  // namespace

int launch_shade(const DevShade& s, void* stream) {
  if (s.n <= 0) return 0;
  hipLaunchKernelGGL(k_shade, dim3((unsigned)((s.n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), s);
  return (int)hipGetLastError();
}

}  // namespace mh
