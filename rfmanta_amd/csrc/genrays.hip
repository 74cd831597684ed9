// Ray generation on the device, so that frames and requests need no host round trip:
//
// k_camera: pinhole camera rays after PinholeCamera::makeRays
//   (Model/Cameras/PinholeCamera.cc:167-230): dir = V*x + U*y + W, normalised, with
//   x, y in [-1, 1] across the image; pixels in 8x8 tiles (TiledImageTraverser square
//   fragments, Engine/ImageTraversers/TiledImageTraverser.cc:255-296), one tile-ordered
//   image per sample; sample 0 of a 1-spp image is the pixel centre, otherwise the
//   offset inside the pixel comes from PCG32(seed, stream = ray index).
// k_shadow: one shadow ray per primary hit after PointLight::computeLight and
//   HardShadows (Model/Lights/PointLight.cc:31-75, Engine/Shadows/HardShadows.cc:87-124):
//   origin P = o + t*d, direction L - P (unnormalised), minT = 1 - T_EPSILON; rays of
//   primary misses and back-facing rays (w.r.t. the face-forward geometric normal) are
//   masked.
// k_bounce: one cosine-weighted diffuse ray per primary hit about the face-forward normal,
//   directions from PCG32(seed, stream = ray index); misses are masked.

#include <hip/hip_runtime.h>

#include "mh_device.h"

namespace mh {
namespace {

constexpr int kBlock = 256;

struct Pcg {
  unsigned long long state, inc;
  __device__ Pcg(unsigned long long seed, unsigned long long stream) {
    inc = (stream << 1) | 1ull;
    state = 0ull;
    next();
    state += seed;
    next();
  }
  __device__ unsigned next() {
    const unsigned long long old = state;
    state = old * 6364136223846793005ull + inc;
    const unsigned xs = (unsigned)(((old >> 18) ^ old) >> 27);
    const unsigned rot = (unsigned)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  __device__ float u01() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

__global__ __launch_bounds__(kBlock) void k_camera(CameraDev c, RayOut out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t per_image = (int64_t)c.xres * c.yres;
  if (i >= per_image * c.spp) return;
  const int64_t p = i % per_image;
  const int tile = (int)(p >> 6), in = (int)(p & 63);
  const int tiles_x = c.xres >> 3;
  const int px = (tile % tiles_x) * 8 + (in & 7);
  const int py = (tile / tiles_x) * 8 + (in >> 3);
  float jx = 0.5f, jy = 0.5f;
  if (c.spp > 1) {
    Pcg g(c.seed, (unsigned long long)i);
    jx = g.u01();
    jy = g.u01();
  }
  const float x = -1.0f + ((float)px + jx) * (2.0f / (float)c.xres);
  const float y = -1.0f + ((float)py + jy) * (2.0f / (float)c.yres);
  float dx = c.V[0] * x + c.U[0] * y + c.W[0];
  float dy = c.V[1] * x + c.U[1] * y + c.W[1];
  float dz = c.V[2] * x + c.U[2] * y + c.W[2];
  const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
  dx *= inv;
  dy *= inv;
  dz *= inv;
  out.ox[i] = c.eye[0];
  out.oy[i] = c.eye[1];
  out.oz[i] = c.eye[2];
  out.dx[i] = dx;
  out.dy[i] = dy;
  out.dz[i] = dz;
  out.tmax[i] = 1.e19f;
  if (out.masked) out.masked[i] = 0;
}

struct Surface {
  float px, py, pz;  // hit point
  float nx, ny, nz;  // face-forward geometric normal (unnormalised)
};

__device__ __forceinline__ Surface surface(const float* verts, const uint32_t* idx, int prim,
                                           const float o[3], const float d[3], float t) {
  Surface s;
  s.px = o[0] + d[0] * t;
  s.py = o[1] + d[1] * t;
  s.pz = o[2] + d[2] * t;
  const uint32_t i0 = idx[3 * (size_t)prim], i1 = idx[3 * (size_t)prim + 1],
                 i2 = idx[3 * (size_t)prim + 2];
  const float ax = verts[3 * (size_t)i1] - verts[3 * (size_t)i0];
  const float ay = verts[3 * (size_t)i1 + 1] - verts[3 * (size_t)i0 + 1];
  const float az = verts[3 * (size_t)i1 + 2] - verts[3 * (size_t)i0 + 2];
  const float bx = verts[3 * (size_t)i2] - verts[3 * (size_t)i0];
  const float by = verts[3 * (size_t)i2 + 1] - verts[3 * (size_t)i0 + 1];
  const float bz = verts[3 * (size_t)i2 + 2] - verts[3 * (size_t)i0 + 2];
  float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  if (nx * d[0] + ny * d[1] + nz * d[2] > 0.f) {
    nx = -nx;
    ny = -ny;
    nz = -nz;
  }
  s.nx = nx;
  s.ny = ny;
  s.nz = nz;
  return s;
}

__global__ __launch_bounds__(kBlock) void k_shadow(SecondaryArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int prim = a.hit_prim[i];
  const float o[3] = {a.in.ox[i], a.in.oy[i], a.in.oz[i]};
  const float d[3] = {a.in.dx[i], a.in.dy[i], a.in.dz[i]};
  uint8_t masked = 1;
  float px = o[0], py = o[1], pz = o[2], lx = 1.f, ly = 0.f, lz = 0.f;
  if (prim >= 0) {
    const Surface s = surface(a.verts, a.idx, prim, o, d, a.hit_t[i]);
    px = s.px;
    py = s.py;
    pz = s.pz;
    lx = a.light[0] - px;
    ly = a.light[1] - py;
    lz = a.light[2] - pz;
    masked = (lx * s.nx + ly * s.ny + lz * s.nz > 0.f) ? 0 : 1;
  }
  a.out.ox[i] = px;
  a.out.oy[i] = py;
  a.out.oz[i] = pz;
  a.out.dx[i] = lx;
  a.out.dy[i] = ly;
  a.out.dz[i] = lz;
  a.out.tmax[i] = 1.0f - 1.e-4f;
  a.out.masked[i] = masked;
}

__global__ __launch_bounds__(kBlock) void k_bounce(SecondaryArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int prim = a.hit_prim[i];
  const float o[3] = {a.in.ox[i], a.in.oy[i], a.in.oz[i]};
  const float d[3] = {a.in.dx[i], a.in.dy[i], a.in.dz[i]};
  float px = o[0], py = o[1], pz = o[2], rx = 0.f, ry = 1.f, rz = 0.f;
  uint8_t masked = 1;
  if (prim >= 0) {
    const Surface s = surface(a.verts, a.idx, prim, o, d, a.hit_t[i]);
    px = s.px;
    py = s.py;
    pz = s.pz;
    const float il = 1.0f / sqrtf(s.nx * s.nx + s.ny * s.ny + s.nz * s.nz);
    const float nx = s.nx * il, ny = s.ny * il, nz = s.nz * il;
    // orthonormal frame about n
    float ax = 1.f, ay = 0.f, az = 0.f;
    if (fabsf(nx) > 0.9f) {
      ax = 0.f;
      ay = 1.f;
    }
    float t1x = ny * az - nz * ay, t1y = nz * ax - nx * az, t1z = nx * ay - ny * ax;
    const float it = 1.0f / sqrtf(t1x * t1x + t1y * t1y + t1z * t1z);
    t1x *= it;
    t1y *= it;
    t1z *= it;
    const float t2x = ny * t1z - nz * t1y, t2y = nz * t1x - nx * t1z, t2z = nx * t1y - ny * t1x;
    Pcg g(a.seed, (unsigned long long)i);
    const float u1 = g.u01(), u2 = g.u01();
    const float r = sqrtf(u1);
    float sn, cs;
    sincosf(6.28318530717958647692f * u2, &sn, &cs);
    const float lx = r * cs, ly = r * sn, lz = sqrtf(1.0f - u1);
    rx = t1x * lx + t2x * ly + nx * lz;
    ry = t1y * lx + t2y * ly + ny * lz;
    rz = t1z * lx + t2z * ly + nz * lz;
    masked = 0;
  }
  a.out.ox[i] = px;
  a.out.oy[i] = py;
  a.out.oz[i] = pz;
  a.out.dx[i] = rx;
  a.out.dy[i] = ry;
  a.out.dz[i] = rz;
  a.out.tmax[i] = 1.e19f;
  a.out.masked[i] = masked;
}

}  // namespace

int launch_camera(const CameraDev& c, const RayOut& out, void* stream) {
  const int64_t n = (int64_t)c.xres * c.yres * c.spp;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_camera, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), c, out);
  return (int)hipGetLastError();
}

int launch_secondary(const SecondaryArgs& a, int kind, void* stream) {
  if (a.n <= 0) return 0;
  const dim3 g((unsigned)((a.n + kBlock - 1) / kBlock));
  if (kind == 0)
    hipLaunchKernelGGL(k_shadow, g, dim3(kBlock), 0, reinterpret_cast<hipStream_t>(stream), a);
  else
    hipLaunchKernelGGL(k_bounce, g, dim3(kBlock), 0, reinterpret_cast<hipStream_t>(stream), a);
  return (int)hipGetLastError();
}

}  // namespace mh
