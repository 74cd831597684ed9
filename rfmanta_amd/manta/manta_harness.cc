// A Manta-shaped driver for HipDynBVH: a Mesh of Kensler-Shirley or Wald triangles, a
// HipDynBVH set on it, preprocess on every worker thread (RTRT::doPreprocess,
// Engine/Control/RTRT.cc:568-590), then worker threads that each fill their own
// RayPacketData with 64-ray packets and call intersect (Raytracer::traceRays,
// Engine/Renderers/Raytracer.cc:49-53; HardShadows, Engine/Shadows/HardShadows.cc:352).
// Exposed through one C entry point so the tests can compare it with the oracle;
// trace_seconds (may be null) receives the wall time of the intersect phase.

#include <Core/Exceptions/InternalError.h>
#include <Interface/Context.h>
#include <Interface/Material.h>
#include <Interface/RayPacket.h>
#include <Model/Groups/Mesh.h>
#include <Model/Primitives/KenslerShirleyTriangle.h>
#include <Model/Primitives/WaldTriangle.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "HipDynBVH.h"

using namespace Manta;

namespace {
class Opaque : public Material {};

struct Io {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  const uint8_t* masked;
  int32_t* prim;
  float *t, *u, *v;
  uint8_t* occluded;
};
}  // namespace

extern "C" int mh_manta_render(const float* verts, uint32_t n_verts, const uint32_t* idx,
                               uint32_t n_tris, int prim_kind, int threads, const Io* io,
                               int64_t n, int any_hit, int frames_refit, char* err,
                               int err_len, double* trace_seconds) {
  try {
    Mesh mesh;
    mesh.vertices.resize(n_verts);
    for (uint32_t i = 0; i < n_verts; ++i)
      mesh.vertices[i] = Vector(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    mesh.vertex_indices.assign(idx, idx + 3 * (size_t)n_tris);
    mesh.face_material.assign(n_tris, 0);
    Opaque matl;
    mesh.materials.push_back(&matl);
    std::vector<std::unique_ptr<MeshTriangle>> tris;
    for (uint32_t t = 0; t < n_tris; ++t) {
      if (prim_kind == MH_PRIM_WALD)
        tris.emplace_back(new WaldTriangle());
      else
        tris.emplace_back(new KenslerShirleyTriangle());
      mesh.addTriangle(tris.back().get());
    }
    HipDynBVH bvh;
    bvh.setGroup(&mesh);
    int dummy;
    MantaInterface* mi = reinterpret_cast<MantaInterface*>(&dummy);
    auto preprocess = [&]() {
      std::vector<std::thread> th;
      for (int p = 0; p < threads; ++p)
        th.emplace_back([&, p]() { bvh.preprocess(PreprocessContext(mi, p, threads)); });
      for (auto& x : th) x.join();
    };
    preprocess();
    // Optional extra frames through the update path (same group, same vertices).
    for (int f = 0; f < frames_refit; ++f) {
      std::vector<std::thread> th;
      for (int p = 0; p < threads; ++p) th.emplace_back([&, p]() { bvh.update(p, threads); });
      for (auto& x : th) x.join();
    }
    BBox bounds;
    bvh.computeBounds(PreprocessContext(mi, 0, 1), bounds);

    const int64_t packets = (n + 63) / 64;
    std::atomic<int64_t> next(0);
    const auto t_start = std::chrono::steady_clock::now();
    std::vector<std::string> errors(threads);
    std::vector<std::thread> th;
    for (int p = 0; p < threads; ++p)
      th.emplace_back([&, p]() {
        try {
          RenderContext ctx(p, threads);
          RayPacketData data;
          for (;;) {
            const int64_t k = next.fetch_add(1);
            if (k >= packets) break;
            const int64_t first = k * 64;
            const int cnt = (int)std::min<int64_t>(64, n - first);
            RayPacket rays(data, RayPacket::UnknownShape, 0, cnt, 0,
                           any_hit ? RayPacket::AnyHit : 0);
            for (int i = 0; i < cnt; ++i) {
              const int64_t r = first + i;
              rays.setRay(i, Vector(io->ox[r], io->oy[r], io->oz[r]),
                          Vector(io->dx[r], io->dy[r], io->dz[r]));
              if (io->masked && io->masked[r]) {
                rays.maskRay(i);
              } else {
                rays.resetHit(i);
                rays.overrideMinT(i, io->tmax[r]);
              }
              rays.getScratchpad<float>(0)[i] = 0.f;
              rays.getScratchpad<float>(1)[i] = 0.f;
            }
            bvh.intersect(ctx, rays);
            for (int i = 0; i < cnt; ++i) {
              const int64_t r = first + i;
              const bool msk = rays.rayIsMasked(i);
              const bool hit = rays.wasHit(i) && !msk;
              const MeshTriangle* tri = static_cast<const MeshTriangle*>(rays.getHitPrimitive(i));
              if (io->prim) io->prim[r] = msk ? -2 : (hit ? (int32_t)tri->getTriangleID() : -1);
              if (io->t) io->t[r] = rays.getMinT(i);
              if (io->u) io->u[r] = rays.getScratchpad<float>(0)[i];
              if (io->v) io->v[r] = rays.getScratchpad<float>(1)[i];
              if (io->occluded) io->occluded[r] = rays.wasHit(i) ? 1 : 0;
            }
          }
        } catch (const std::exception& ex) {
          errors[p] = ex.what();
        }
      });
    for (auto& x : th) x.join();
    if (trace_seconds)
      *trace_seconds =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    for (auto& e : errors)
      if (!e.empty()) throw InternalError(e);
    return 0;
  } catch (const std::exception& ex) {
    if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", ex.what());
    return -1;
  }
}
