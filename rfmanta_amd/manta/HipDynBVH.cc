// HipDynBVH implementation (see HipDynBVH.h).

#include "HipDynBVH.h"

#include <Core/Exceptions/InternalError.h>
#include <Interface/Context.h>
#include <Model/Groups/Mesh.h>
#include <Model/Primitives/KenslerShirleyTriangle.h>
#include <Model/Primitives/WaldTriangle.h>

#include <chrono>
#include <cstring>
#include <vector>

namespace Manta {

// Gathers the 64-ray packets of concurrent intersect() calls into one launch of the
// packet traversal. Slot k of a batch holds one caller's packet in rays [64k, 64k+64);
// rays beyond the caller's count are masked, which the traversal skips.
class HipDynBVH::PacketBatcher {
 public:
  struct In {
    float soa[7][64];  // ox oy oz dx dy dz tmax
    uint8_t masked[64];
    int count;
  };
  struct Out {
    int32_t prim[64];
    float t[64], u[64], v[64];
    uint8_t occluded[64];
  };

  PacketBatcher(const mh_scene* scene, int max_packets, int wait_us)
      : scene_(scene), max_packets_(max_packets), wait_(wait_us) {}

  int trace(const In& in, Out& out, bool any_hit, int expected) {
    std::unique_lock<std::mutex> lk(mu_);
    std::shared_ptr<Batch>& open = open_[any_hit ? 1 : 0];
    if (!open) open = std::make_shared<Batch>(max_packets_);
    std::shared_ptr<Batch> b = open;
    const int slot = b->count++;
    b->put(slot, in);
    const int target = expected < max_packets_ ? (expected > 0 ? expected : 1) : max_packets_;
    bool leader = false;
    if (b->count >= target) {
      open.reset();
      leader = true;
    } else {
      const auto deadline = std::chrono::steady_clock::now() + wait_;
      while (!b->closed) {
        if (cv_.wait_until(lk, deadline) == std::cv_status::timeout) break;
      }
      if (!b->closed) {
        if (open == b) open.reset();
        leader = true;
      }
    }
    if (leader) {
      b->closed = true;
      cv_.notify_all();
      lk.unlock();
      const int rc = b->launch(scene_, any_hit);
      lk.lock();
      b->status = rc;
      b->done = true;
      cv_.notify_all();
    } else {
      while (!b->done) cv_.wait(lk);
    }
    b->get(slot, out);
    return b->status;
  }

 private:
  struct Batch {
    explicit Batch(int cap) : cap(cap) {
      for (auto& a : soa) a.assign(64 * (size_t)cap, 0.f);
      masked.assign(64 * (size_t)cap, 1);
      prim.resize(64 * (size_t)cap);
      t.resize(64 * (size_t)cap);
      u.resize(64 * (size_t)cap);
      v.resize(64 * (size_t)cap);
      occ.resize(64 * (size_t)cap);
    }
    void put(int slot, const In& in) {
      const size_t o = 64 * (size_t)slot;
      for (int f = 0; f < 7; ++f) std::memcpy(&soa[f][o], in.soa[f], 64 * sizeof(float));
      std::memcpy(&masked[o], in.masked, 64);
    }
    void get(int slot, Out& out) const {
      const size_t o = 64 * (size_t)slot;
      std::memcpy(out.prim, &prim[o], 64 * sizeof(int32_t));
      std::memcpy(out.t, &t[o], 64 * sizeof(float));
      std::memcpy(out.u, &u[o], 64 * sizeof(float));
      std::memcpy(out.v, &v[o], 64 * sizeof(float));
      std::memcpy(out.occluded, &occ[o], 64);
    }
    int launch(const mh_scene* scene, bool any_hit) {
      mh_rays r{soa[0].data(), soa[1].data(), soa[2].data(), soa[3].data(),
                soa[4].data(), soa[5].data(), soa[6].data(), masked.data()};
      mh_hits h{prim.data(), t.data(), u.data(), v.data(), occ.data()};
      mh_trace_opts o{};
      o.packet_size = 64;
      return mh_trace_host(scene, &r, &h, 64 * (int64_t)count, any_hit ? MH_ANY_HIT : MH_CLOSEST_HIT,
                           &o);
    }
    int cap;
    int count = 0;
    bool closed = false, done = false;
    int status = 0;
    std::vector<float> soa[7];
    std::vector<uint8_t> masked;
    std::vector<int32_t> prim;
    std::vector<float> t, u, v;
    std::vector<uint8_t> occ;
  };

  const mh_scene* scene_;
  int max_packets_;
  std::chrono::microseconds wait_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::shared_ptr<Batch> open_[2];
};

HipDynBVH::HipDynBVH()
    : group_(0), mesh_(0), group_changed_(false), scene_(0), prim_kind_(MH_PRIM_KENSLER_SHIRLEY),
      workers_(0), max_packets_(256), wait_us_(200), barrier_count_(0),
      barrier_generation_(0) {}

HipDynBVH::~HipDynBVH() {
  batcher_.reset();
  if (scene_) mh_scene_destroy(scene_);
}

void HipDynBVH::fail(const char* what) const {
  throw InternalError(std::string("HipDynBVH::") + what + ": " + mh_last_error());
}

void HipDynBVH::setGroup(Group* new_group) {
  if (new_group != group_) group_changed_ = true;
  group_ = new_group;
  mesh_ = dynamic_cast<Mesh*>(new_group);
}

Group* HipDynBVH::getGroup() const { return group_; }

void HipDynBVH::groupDirty() { group_changed_ = true; }

void HipDynBVH::addToUpdateGraph(ObjectUpdateGraph*, ObjectUpdateGraphNode*) {}

void HipDynBVH::setBatching(int max_packets, int wait_us) {
  max_packets_ = max_packets > 0 ? max_packets : 1;
  wait_us_ = wait_us >= 0 ? wait_us : 0;
  if (scene_) batcher_.reset(new PacketBatcher(scene_, max_packets_, wait_us_));
}

// DynBVH::preprocess (Model/Groups/DynBVH.cc:930-970): preprocess the group, then build.
void HipDynBVH::preprocess(const PreprocessContext& context) {
  if (!group_) return;
  group_->preprocess(context);
  if (context.isInitialized()) rebuild(context.proc, context.numProcs);
}

// DynBVH::rebuild / update (DynBVH.cc:1574-1581, 1699-1793): every render thread calls
// in; thread 0 does the work and the others wait for it.
void HipDynBVH::rebuild(int proc, int numProcs) {
  if (proc == 0) rebuildSerial(numProcs);
  rendezvous(numProcs);
}

void HipDynBVH::update(int proc, int numProcs) { rebuild(proc, numProcs); }

void HipDynBVH::rendezvous(int numProcs) {
  if (numProcs <= 1) return;
  std::unique_lock<std::mutex> lk(barrier_mu_);
  const unsigned gen = barrier_generation_;
  if (++barrier_count_ == numProcs) {
    barrier_count_ = 0;
    ++barrier_generation_;
    barrier_cv_.notify_all();
  } else {
    barrier_cv_.wait(lk, [&] { return barrier_generation_ != gen; });
  }
}

void HipDynBVH::ensureScene() {
  if (!mesh_) throw InternalError("HipDynBVH: the group is not a Mesh");
  const size_t T = mesh_->vertex_indices.size() / 3;
  if (T == 0) throw InternalError("Tried building BVH over invalid range");
  if (scene_ && !group_changed_) return;
  batcher_.reset();
  if (scene_) mh_scene_destroy(scene_);
  scene_ = 0;
  const float* verts = &mesh_->vertices[0].data[0];
  if (mh_scene_create(verts, (uint32_t)mesh_->vertices.size(), &mesh_->vertex_indices[0],
                      (uint32_t)T, &scene_) != MH_OK)
    fail("ensureScene");
  prim_kind_ = dynamic_cast<const WaldTriangle*>(mesh_->get(0)) ? MH_PRIM_WALD
                                                                 : MH_PRIM_KENSLER_SHIRLEY;
}

void HipDynBVH::rebuildSerial(int numProcs) {
  if (group_changed_ || !scene_) {
    ensureScene();
    // rebuild(proc, numProcs): largeSubtreeSize = nodes.size() / (20 * numProcs)
    // (DynBVH.cc:1776) before the build
    if (mh_scene_set_workers(scene_, numProcs) != MH_OK) fail("rebuild");
    workers_ = numProcs;
    mh_build_opts o{prim_kind_, 1, 0, 0};
    if (mh_scene_build(scene_, &o) != MH_OK) fail("rebuild");
    group_changed_ = false;
    batcher_.reset(new PacketBatcher(scene_, max_packets_, wait_us_));
    if (needToSaveFile) saveToFile(saveFileName);
    return;
  }
  // A loaded tree, or a new thread count: computeSubTreeSizes with this numProcs, as
  // preprocess does before an update (DynBVH.cc:941-950).
  if (workers_ != numProcs) {
    if (mh_scene_set_workers(scene_, numProcs) != MH_OK) fail("update");
    workers_ = numProcs;
  }
  // Same group, moved vertices: DynBVH::update on the device, i.e. the bounds refit with
  // tree rotations and leaf splits (DynBVH.cc:1574-1581, 2012-2066; the tree does not
  // depend on numProcs, only the isLargeSubtree bits do).
  const float* verts = &mesh_->vertices[0].data[0];
  if (mh_scene_set_vertices(scene_, verts, 0, 0) != MH_OK) fail("update");
  if (mh_scene_update(scene_, 0) != MH_OK) fail("update");
}

// DynBVH::buildFromFile (DynBVH.cc:2442-2472).
bool HipDynBVH::buildFromFile(const std::string& file) {
  if (!group_) return false;
  ensureScene();
  if (mh_scene_load_dynbvh_file(scene_, file.c_str(), prim_kind_) != MH_OK) return false;
  group_changed_ = false;
  workers_ = 0;  // preprocess sets the bits for its numProcs
  batcher_.reset(new PacketBatcher(scene_, max_packets_, wait_us_));
  return true;
}

// DynBVH::saveToFile (DynBVH.cc:2475-2501): deferred until the tree exists.
bool HipDynBVH::saveToFile(const std::string& file) {
  if (!scene_ || !batcher_) {
    needToSaveFile = true;
    saveFileName = file;
    return true;
  }
  needToSaveFile = false;
  return mh_scene_save_dynbvh_file(scene_, file.c_str()) == MH_OK;
}

void HipDynBVH::computeBounds(const PreprocessContext& context, BBox& bbox) const {
  if (scene_ && batcher_) {
    float lo[3], hi[3];
    if (mh_scene_bounds(scene_, lo, hi) != MH_OK) fail("computeBounds");
    bbox.extendByPoint(Vector(lo[0], lo[1], lo[2]));
    bbox.extendByPoint(Vector(hi[0], hi[1], hi[2]));
    return;
  }
  if (group_) group_->computeBounds(context, bbox);
}

// DynBVH::intersect (DynBVH.cc:101-404) for rays [begin, end) of the packet.
void HipDynBVH::intersect(const RenderContext& context, RayPacket& rays) const {
  if (!batcher_) throw InternalError("HipDynBVH::intersect before rebuild");
  const int b = rays.begin(), e = rays.end(), n = e - b;
  if (n <= 0) return;
  const bool any = rays.getFlag(RayPacket::AnyHit);
  PacketBatcher::In in;
  in.count = n;
  RayPacketData* d = rays.data;
  for (int k = 0; k < 64; ++k) {
    const int i = b + k;
    if (k >= n) {
      for (int f = 0; f < 7; ++f) in.soa[f][k] = f == 3 ? 1.f : 0.f;
      in.masked[k] = 1;
      continue;
    }
    for (int a = 0; a < 3; ++a) {
      in.soa[a][k] = d->origin[a][i];
      in.soa[3 + a][k] = d->direction[a][i];
    }
    in.soa[6][k] = d->minT[i];
    // Masked rays, and any-hit rays already hit, are skipped by the traversal
    // (DynBVH.cc:415-420, 730-741).
    in.masked[k] = (rays.rayIsMasked(i) || (any && rays.wasHit(i))) ? 1 : 0;
  }
  PacketBatcher::Out out;
  if (batcher_->trace(in, out, any, context.numProcs) != MH_OK) fail("intersect");
  for (int k = 0; k < n; ++k) {
    if (in.masked[k]) continue;
    const int i = b + k;
    if (any) {
      if (!out.occluded[k]) continue;
      const int p = out.prim[k];
      const MeshTriangle* tri = mesh_->get(p);
      rays.setHitMaterial(i, mesh_->materials[mesh_->face_material[p]]);
      rays.setHitPrimitive(i, tri);
      rays.setHitTexCoordMapper(i, dynamic_cast<const TexCoordMapper*>(tri));
      rays.overrideMinT(i, T_EPSILON);  // opaque occluder ends the ray (DynBVH.cc:352)
      continue;
    }
    const int p = out.prim[k];
    if (p < 0) continue;
    const MeshTriangle* tri = mesh_->get(p);
    if (rays.hit(i, out.t[k], mesh_->materials[mesh_->face_material[p]], tri,
                 dynamic_cast<const TexCoordMapper*>(tri))) {
      rays.getScratchpad<float>(0)[i] = out.u[k];  // SCRATCH_U: weight of vertex 1
      rays.getScratchpad<float>(1)[i] = out.v[k];  // SCRATCH_V: weight of vertex 2
    }
  }
}

}  // namespace Manta
