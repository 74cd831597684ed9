// C ABI of device ray generation (mh_gen_*, include/mh.h).

#include <hip/hip_runtime_api.h>

#include <cmath>

#include "internal.h"
#include "mh.h"
#include "mh_device.h"

using mh_internal::ensure_init;
using mh_internal::fail;

namespace mh_internal {

// PinholeCamera::setup (Model/Cameras/PinholeCamera.cc:167-198), in double, then binary32.
int camera_setup(const mh_camera& c, mh::CameraDev& d) {
  if (c.xres <= 0 || c.yres <= 0 || (c.xres & 7) || (c.yres & 7) || c.spp < 1)
    return fail(MH_E_INVALID, "camera %dx%d spp %d: sizes must be positive multiples of 8",
                c.xres, c.yres, c.spp);
  if (!(c.vfov_deg > 0.f && c.vfov_deg < 180.f))
    return fail(MH_E_INVALID, "vfov %g", (double)c.vfov_deg);
  double w[3], up[3], v[3], u[3];
  for (int k = 0; k < 3; ++k) {
    w[k] = (double)c.lookat[k] - (double)c.eye[k];
    up[k] = c.up[k];
  }
  const double nearz = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto normalize = [](double* a) {
    const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (l > 0) {
      a[0] /= l;
      a[1] /= l;
      a[2] /= l;
    }
    return l;
  };
  cross(w, up, v);
  if (nearz == 0 || normalize(v) == 0) return fail(MH_E_INVALID, "degenerate camera frame");
  cross(v, w, u);
  normalize(u);
  const double pi = 3.14159265358979323846;
  const double tv = std::tan(0.5 * c.vfov_deg * pi / 180.0);
  const double th = c.hfov_deg > 0.f ? std::tan(0.5 * c.hfov_deg * pi / 180.0)
                                      : tv * (double)c.xres / (double)c.yres;
  for (int k = 0; k < 3; ++k) {
    d.eye[k] = c.eye[k];
    d.U[k] = (float)(u[k] * nearz * tv);
    d.V[k] = (float)(v[k] * nearz * th);
    d.W[k] = (float)w[k];
  }
  d.xres = c.xres;
  d.yres = c.yres;
  d.spp = c.spp;
  d.seed = c.seed;
  return MH_OK;
}

}  // namespace mh_internal

namespace {

int check_out(const mh_ray_buffers* o, bool need_mask) {
  if (!o || !o->ox || !o->oy || !o->oz || !o->dx || !o->dy || !o->dz || !o->tmax ||
      (need_mask && !o->masked))
    return fail(MH_E_INVALID, "output ray arrays must be non-null");
  return MH_OK;
}

mh::RayOut ray_out(const mh_ray_buffers* o) {
  return mh::RayOut{o->ox, o->oy, o->oz, o->dx, o->dy, o->dz, o->tmax, o->masked};
}

int secondary(const mh_scene* s, const mh_rays* p, const mh_hits* h, int64_t n,
              const float* light, uint64_t seed, const mh_ray_buffers* out, int kind,
              void* stream) {
  if (!s || !p || !h || !h->prim || !h->t) return fail(MH_E_INVALID, "null argument");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree");
  if (n < 0) return fail(MH_E_INVALID, "negative count");
  int rc = check_out(out, true);
  if (rc) return rc;
  if ((rc = ensure_init())) return rc;
  mh::SecondaryArgs a{};
  a.in = mh::DevRays{p->ox, p->oy, p->oz, p->dx, p->dy, p->dz, p->tmax, p->masked};
  a.hit_prim = h->prim;
  a.hit_t = h->t;
  a.verts = s->d_verts.as<const float>();
  a.idx = s->d_idx.as<const uint32_t>();
  if (light)
    for (int k = 0; k < 3; ++k) a.light[k] = light[k];
  a.seed = seed;
  a.n = n;
  a.out = ray_out(out);
  HIP_OK((hipError_t)mh::launch_secondary(a, kind, stream));
  return MH_OK;
}

}  // namespace

extern "C" {

int mh_gen_camera_rays(const mh_camera* cam, const mh_ray_buffers* out, void* stream) {
  if (!cam) return fail(MH_E_INVALID, "null camera");
  int rc = check_out(out, false);
  if (rc) return rc;
  if ((rc = ensure_init())) return rc;
  mh::CameraDev d;
  if ((rc = mh_internal::camera_setup(*cam, d))) return rc;
  HIP_OK((hipError_t)mh::launch_camera(d, ray_out(out), stream));
  return MH_OK;
}

int mh_gen_shadow_rays(const mh_scene* s, const mh_rays* p, const mh_hits* h, int64_t n,
                       const float light[3], const mh_ray_buffers* out, void* stream) {
  if (!light) return fail(MH_E_INVALID, "null light");
  return secondary(s, p, h, n, light, 0, out, 0, stream);
}

int mh_gen_bounce_rays(const mh_scene* s, const mh_rays* p, const mh_hits* h, int64_t n,
                       uint64_t seed, const mh_ray_buffers* out, void* stream) {
  return secondary(s, p, h, n, nullptr, seed, out, 1, stream);
}

}  // extern "C"
