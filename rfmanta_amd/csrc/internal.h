// Library-internal declarations shared by the C-ABI translation units (mh_api.cpp,
// frame_api.cpp): error recording, device initialisation, device buffers and the scene.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "mh.h"
#include "mh_device.h"
#include "sah_builder.h"

namespace mh_internal {

// Records a thread-local message for mh_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);
// Selects the library's device (first call queries it); MH_OK or an error code.
int ensure_init();
// Compute units and architecture name of the library's device (valid after ensure_init()).
int cu_count();
const char* device_name();

// The allocation-free trace's fixed scratch area (noalloc.hip). scratch_setup looks up
// its device address on `device` (which must be current) once; ensure_init calls it. A
// failed lookup is kept and reported by mh_trace_noalloc only.
void scratch_setup(int device);
// mh_trace with the single-ray kernel's stack spill in `spill` (spill_bytes long) instead
// of a buffer the scene allocates: the workgroup count is lowered until every lane's
// spill column fits; an error if fewer than one workgroup per compute unit would fit.
int trace_fixed_spill(const mh_scene* s, const mh_rays* rays, const mh_hits* hits, int64_t n,
                      uint32_t flags, const mh_trace_opts* opts, void* stream, void* spill,
                      size_t spill_bytes);

// Brings the scene's SoA copy of its PairNodes up to date on `stream` (MH_LAYOUT_SOA
// launches); allocates it on first use.
int ensure_soa(mh_scene* s, void* stream);

// Validates a batch launch configuration (frame_api.cpp); MH_OK or an error code.
int check_launch_config(const mh_launch_config& c);

// PinholeCamera::setup for the device generator (rays_api.cpp).
int camera_setup(const mh_camera& c, mh::CameraDev& d);

#define HIP_OK(call)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return ::mh_internal::fail(MH_E_HIP, "%s: %s", #call, hipGetErrorString(e_));               \
  } while (0)

// One device allocation, released with the owner.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int reserve(size_t n) {
    if (n <= bytes && p) return MH_OK;
    release();
    if (n == 0) return MH_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(MH_E_NOMEM, "hipMalloc(%zu): %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return MH_OK;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

}  // namespace mh_internal

struct mh_scene {
  using DevBuf = mh_internal::DevBuf;
  std::vector<float> verts;
  std::vector<uint32_t> idx;
  uint32_t nv = 0, nt = 0;
  mh::DynBvh bvh;
  bool built = false;
  int prim_kind = MH_PRIM_KENSLER_SHIRLEY;
  int root_ref = 0;
  uint32_t next_free = 0;  // records in use (the reference's nextFree)
  // The rotating update needs records [0, next_free) to be exactly the tree and the
  // leaves to partition the object slots; trees loaded from elsewhere may not be.
  bool updatable = false;
  uint32_t n_inner = 0, n_leaves = 0, depth = 0;
  uint32_t n_records = 0;  // records held (2T after a build; the file's count after a load)
  // Render threads of the Manta instance (numProcs): largeSubtreeSize = records / (20 *
  // workers) (DynBVH.cc:945, 1776) decides the isLargeSubtree bits.
  int workers = 1;
  uint32_t large_limit() const { return (uint32_t)(n_records / (20 * (uint32_t)workers)); }

  DevBuf refs, pairs, tris, ids, big, d_verts, d_base, d_idx, leaves, levels;
  // rotating update (mh_scene_update): costs, subtree sizes, parents, object boxes and
  // centroids, depth work arrays, the depth histogram and counters
  DevBuf costs, ssize, parent, obox, ocen, anc0, dep0, anc1, dep1, hist, counters;
  DevBuf split_list;  // rotating update: leaves whose parent's rotation may split them
  DevBuf levels_aux;  // rotating update: first child and word beside each level entry
  // rotating update: per object slot, the record of the leaf whose first slot it is, else -1
  // (filled from the leaf list after a build or load, kept by the update's level pass)
  DevBuf leaf_at;
  bool leaf_at_valid = false;
  DevBuf logtab;  // RTSAH table logarithm (device build)
  DevBuf normals, normal_idx, texcoords, tex_idx;  // shading attributes (mh_shade)
  // The PairNodes as four arrays of their 16-byte fields (MH_LAYOUT_SOA launches):
  // allocated on first use and copied from `pairs` again when pairs_gen (bumped by every
  // call that writes the pair records) has moved past soa_gen.
  DevBuf pairs_soa;
  uint64_t pairs_gen = 0, soa_gen = ~0ull;
  std::vector<int64_t> level_off;  // inner records of depth d: levels[level_off[d], level_off[d+1])

  std::mutex spill_mu;  // guards the per-stream spill areas
  std::mutex stage_mu;  // serialises mh_trace_host calls on the staging buffers
  std::map<void*, DevBuf*> spill;
  DevBuf stage;                       // mh_trace_host: device staging
  void* host_stage = nullptr;         // mh_trace_host: pinned host staging
  size_t host_stage_bytes = 0;
  hipStream_t stage_stream = nullptr;  // mh_trace_host: the scene's own stream
  unsigned* host_ctr = nullptr;        // mh_scene_update: pinned copy of the level counters
  ~mh_scene() {
    for (auto& kv : spill) delete kv.second;
    if (host_stage) (void)hipHostFree(host_stage);
    if (host_ctr) (void)hipHostFree(host_ctr);
    if (stage_stream) (void)hipStreamDestroy(stage_stream);
  }
  size_t device_bytes() const {
    return refs.bytes + pairs.bytes + tris.bytes + ids.bytes + big.bytes + d_verts.bytes +
           d_base.bytes + d_idx.bytes + leaves.bytes + levels.bytes + costs.bytes + ssize.bytes +
           parent.bytes + obox.bytes + ocen.bytes + anc0.bytes + dep0.bytes + anc1.bytes +
           dep1.bytes + hist.bytes + counters.bytes + logtab.bytes + normals.bytes +
           normal_idx.bytes + texcoords.bytes + tex_idx.bytes + pairs_soa.bytes +
           split_list.bytes + levels_aux.bytes + leaf_at.bytes;
  }
  // Traversal stack entries a single ray can hold at most (one push per level).
  uint32_t stack_bound(int /*layout*/) const {
    return depth;
  }
  mh::DevScene dev() const {
    mh::DevScene d{};
    d.pairs = pairs.as<const mh::PairNode>();
    d.refs = refs.as<const mh::RefNode>();
    d.tris = tris.as<const float>();
    d.obj_ids = ids.as<const int32_t>();
    d.big_count = big.as<const int32_t>();
    d.root_ref = root_ref;
    d.prim_kind = prim_kind;
    d.n_pairs = (uint32_t)(pairs.bytes / sizeof(mh::PairNode));
    d.n_slots = (uint32_t)(tris.bytes / (12 * sizeof(float)));
    d.pairs_soa = pairs_soa.p;
    return d;
  }
};
