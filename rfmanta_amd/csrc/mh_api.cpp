// C ABI of libmh.so (declared in include/mh.h): scene ownership, device upload, tree
// import/export, refit and trace dispatch. No C++ exception leaves this file; every
// entry point returns a status code and records a thread-local message.

#include "mh.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "mh_device.h"
#include "rtsah.h"
#include "sah_builder.h"

namespace mh_internal {

namespace {
thread_local std::string g_err;

struct Lib {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  int cu_count = 0;
  size_t hbm = 0;
  char name[256] = {0};
};
Lib g_lib;
}  // namespace

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  // The failure is reported here: take it off the thread's HIP error state too, or the
  // check after this thread's next kernel launch (hipGetLastError) would report it again.
  (void)hipGetLastError();
  return code;
}

int ensure_init() {
  std::lock_guard<std::mutex> lk(g_lib.mu);
  if (g_lib.ready) {
    HIP_OK(hipSetDevice(g_lib.device));
    return MH_OK;
  }
  int count = 0;
  HIP_OK(hipGetDeviceCount(&count));
  if (g_lib.device < 0 || g_lib.device >= count)
    return fail(MH_E_INVALID, "device %d out of range (%d devices)", g_lib.device, count);
  HIP_OK(hipSetDevice(g_lib.device));
  hipDeviceProp_t p;
  HIP_OK(hipGetDeviceProperties(&p, g_lib.device));
  g_lib.cu_count = p.multiProcessorCount;
  g_lib.hbm = p.totalGlobalMem;
  std::snprintf(g_lib.name, sizeof(g_lib.name), "%s", p.gcnArchName);
  scratch_setup(g_lib.device);
  g_lib.ready = true;
  return MH_OK;
}

int cu_count() { return g_lib.cu_count; }
const char* device_name() { return g_lib.name; }

}  // namespace mh_internal

using mh_internal::DevBuf;
using mh_internal::ensure_init;
using mh_internal::fail;
using mh_internal::g_err;
using mh_internal::g_lib;

namespace {

// Rotation state after DynBVH::rebuild with TREE_ROT: computeCost<true>(0)
// (DynBVH.cc:3388-3415) and subtree sizes (DynBVH.cc:1971-1996) of the reachable records.
void rotation_state(const std::vector<mh::RefNode>& nodes, std::vector<float>& costs,
                    std::vector<uint32_t>& ssize) {
  costs.assign(nodes.size(), 0.f);
  ssize.assign(nodes.size(), 0u);
  auto area = [](const mh::Aabb& b) {  // BBox::computeArea, float arithmetic
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2 * (dx * dy + dy * dz + dx * dz);
  };
  struct It {
    int32_t rec;
    bool done;
  };
  std::vector<It> st{{0, false}};
  while (!st.empty()) {
    const It it = st.back();
    st.pop_back();
    const mh::RefNode& n = nodes[it.rec];
    if (n.leaf()) {
      costs[it.rec] = n.children * 1.f;
      ssize[it.rec] = 1;
      continue;
    }
    if (!it.done) {
      st.push_back({it.rec, true});
      st.push_back({n.child + 1, false});
      st.push_back({n.child, false});
      continue;
    }
    const int l = n.child, r = n.child + 1;
    costs[it.rec] = 1.f + (area(nodes[l].box) * costs[l] + area(nodes[r].box) * costs[r]) /
                              area(n.box);
    ssize[it.rec] = 1 + ssize[l] + ssize[r];
  }
}

// Lays the reference tree out for the device: PairNodes indexed by sibling pair, leaf
// references, the leaf and per-depth lists for refits, the rotation state, then uploads.
// With rtsah, isLeftCheaper is recomputed once the records are known to form a tree
// (DynBVH::buildFromFile runs computeTraversalCost after reading, DynBVH.cc:2469-2472).
int upload(mh_scene* s, bool rtsah = false) {
  std::vector<mh::RefNode>& nodes = s->bvh.nodes;
  const uint32_t nrec = (uint32_t)nodes.size();
  const uint32_t T = s->nt;
  std::vector<int32_t> leaves;
  std::vector<std::vector<int32_t>> by_depth;  // inner records per depth, preorder
  std::vector<int32_t> big(T, 0);
  struct It {
    int32_t rec;
    uint32_t depth;
  };
  std::vector<It> st;
  st.push_back({0, 0});
  uint32_t max_depth = 0, n_inner = 0, max_rec = 0, n_reached = 0;
  uint64_t leaf_objects = 0;
  std::vector<uint8_t> seen(nrec, 0);
  while (!st.empty()) {
    It it = st.back();
    st.pop_back();
    const mh::RefNode& n = nodes[it.rec];
    if (seen[it.rec]) return fail(MH_E_INVALID, "record %d is reached twice: not a tree", it.rec);
    seen[it.rec] = 1;
    ++n_reached;
    max_depth = std::max(max_depth, it.depth);
    max_rec = std::max(max_rec, (uint32_t)it.rec);
    if (n.leaf()) {
      if (n.children < 0) return fail(MH_E_INVALID, "record %d: negative leaf size", it.rec);
      if ((uint32_t)n.child + (uint32_t)n.children > T || n.child < 0)
        return fail(MH_E_INVALID, "record %d: leaf range outside object_ids", it.rec);
      leaves.push_back(it.rec);
      leaf_objects += (uint32_t)n.children;
      if (n.children >= mh::kLeafCountInline) big[n.child] = n.children;
      continue;
    }
    if (n.child < 1 || (uint32_t)n.child + 1 >= nrec)
      return fail(MH_E_INVALID, "record %d: child %d outside the node array", it.rec, n.child);
    if ((n.child & 1) == 0)
      return fail(MH_E_INVALID, "record %d: child %d is not the first record of a pair "
                  "(DynBVH allocates sibling records in pairs from 1)", it.rec, n.child);
    if (n.axis > 2) return fail(MH_E_INVALID, "record %d: axis %d", it.rec, (int)n.axis);
    if (++n_inner > nrec) return fail(MH_E_INVALID, "node graph is not a tree");
    if (it.depth >= (uint32_t)mh::kMaxDepth - 1)
      return fail(MH_E_UNSUPPORTED, "tree deeper than %d levels", mh::kMaxDepth - 1);
    if (by_depth.size() <= it.depth) by_depth.resize(it.depth + 1);
    by_depth[it.depth].push_back(it.rec);
    st.push_back({n.child + 1, it.depth + 1});
    st.push_back({n.child, it.depth + 1});
  }
  if (rtsah) mh::compute_rtsah(nodes, 0);
  auto encode = [&](int32_t rec) -> int32_t {
    const mh::RefNode& n = nodes[rec];
    if (!n.leaf()) return (n.child - 1) >> 1;
    const int c = std::min<int>(n.children, mh::kLeafCountInline);
    return ~((n.child << mh::kLeafCountBits) | c);
  };
  const size_t n_pairs = std::max<size_t>(1, nrec / 2);
  std::vector<mh::PairNode> pairs(n_pairs);
  for (const auto& lv : by_depth)
    for (int32_t r : lv) {
      const mh::RefNode& n = nodes[r];
      mh::PairNode& P = pairs[(n.child - 1) >> 1];
      for (int side = 0; side < 2; ++side) {
        const mh::RefNode& c = nodes[n.child + side];
        float* b = P.b + 6 * side;
        b[0] = c.box.lo[0];
        b[1] = c.box.hi[0];
        b[2] = c.box.lo[1];
        b[3] = c.box.hi[1];
        b[4] = c.box.lo[2];
        b[5] = c.box.hi[2];
        P.ref[side] = encode(n.child + side);
      }
      P.meta = (int32_t)(n.axis & 3) | ((n.bits & 1) << 2);
      P.node = r;
    }
  s->root_ref = encode(0);
  s->n_inner = n_inner;
  s->n_leaves = (uint32_t)leaves.size();
  s->leaf_at_valid = false;
  s->depth = max_depth;
  s->next_free = max_rec + 1;
  s->n_records = nrec;
  // Leaf ranges lie inside [0, T); they partition it iff their sizes sum to T and no slot
  // is covered twice.
  bool partition = leaf_objects == T;
  if (partition) {
    std::vector<uint8_t> cover(T, 0);
    for (int32_t r : leaves)
      for (int k = 0; k < nodes[r].children && partition; ++k)
        partition = !cover[nodes[r].child + k]++;
  }
  s->updatable = partition && n_reached == max_rec + 1;
  std::vector<int32_t> levels;
  s->level_off.assign(1, 0);
  for (auto& v : by_depth) {
    levels.insert(levels.end(), v.begin(), v.end());
    s->level_off.push_back((int64_t)levels.size());
  }
  std::vector<float> costs;
  std::vector<uint32_t> ssize;
  rotation_state(nodes, costs, ssize);

  int rc;
  if ((rc = s->refs.reserve(nrec * sizeof(mh::RefNode)))) return rc;
  if ((rc = s->pairs.reserve(n_pairs * sizeof(mh::PairNode)))) return rc;
  if ((rc = s->tris.reserve((size_t)T * 12 * sizeof(float)))) return rc;
  if ((rc = s->ids.reserve((size_t)T * sizeof(int32_t)))) return rc;
  if ((rc = s->big.reserve((size_t)T * sizeof(int32_t)))) return rc;
  // Leaf and level lists: room for T entries, which the rotating update needs (a
  // partition of T objects has at most T leaves and T - 1 inner records), and for every
  // record of a loaded tree whose leaves overlap (it can hold more leaves than objects).
  const size_t list_cap = std::max<size_t>({(size_t)std::max(1u, T), leaves.size(), levels.size()});
  if ((rc = s->leaves.reserve(list_cap * sizeof(int32_t)))) return rc;
  if ((rc = s->levels.reserve(list_cap * sizeof(int32_t)))) return rc;
  if ((rc = s->costs.reserve(nrec * sizeof(float)))) return rc;
  if ((rc = s->ssize.reserve(nrec * sizeof(uint32_t)))) return rc;
  HIP_OK(hipMemcpy(s->refs.p, nodes.data(), nrec * sizeof(mh::RefNode), hipMemcpyHostToDevice));
  ++s->pairs_gen;
  HIP_OK(hipMemcpy(s->pairs.p, pairs.data(), n_pairs * sizeof(mh::PairNode),
                   hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->ids.p, s->bvh.ids.data(), (size_t)T * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->big.p, big.data(), (size_t)T * 4, hipMemcpyHostToDevice));
  if (!leaves.empty())
    HIP_OK(hipMemcpy(s->leaves.p, leaves.data(), leaves.size() * 4, hipMemcpyHostToDevice));
  if (!levels.empty())
    HIP_OK(hipMemcpy(s->levels.p, levels.data(), levels.size() * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->costs.p, costs.data(), nrec * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->ssize.p, ssize.data(), nrec * sizeof(uint32_t), hipMemcpyHostToDevice));

  mh::DevRefit r{};
  r.tris = s->tris.as<float>();
  r.obj_ids = s->ids.as<int32_t>();
  r.verts = s->d_verts.as<float>();
  r.idx = s->d_idx.as<uint32_t>();
  r.n_slots = (int32_t)T;
  r.prim_kind = s->prim_kind;
  HIP_OK((hipError_t)mh::launch_prim_setup(r, nullptr));
  HIP_OK(hipDeviceSynchronize());
  s->built = true;
  return MH_OK;
}

int validate_ids(const std::vector<int32_t>& ids, uint32_t T) {
  if (ids.size() != T) return fail(MH_E_INVALID, "object_ids has %zu entries for %u triangles",
                                   ids.size(), T);
  std::vector<uint8_t> seen(T, 0);
  for (int32_t id : ids) {
    if (id < 0 || (uint32_t)id >= T || seen[id])
      return fail(MH_E_INVALID, "object_ids is not a permutation of the triangles");
    seen[id] = 1;
  }
  return MH_OK;
}

mh::DevScene dev_scene(const mh_scene* s) { return s->dev(); }

int choose_stack(const mh_trace_opts* o) {
  const int v = o ? o->stack_lds : 0;
  return (v == 8 || v == 16 || v == 32) ? v : 8;
}

// fixed_spill / fixed_bytes: a stack spill area the caller provides (mh_trace_noalloc);
// null = the scene's per-stream spill buffer, grown here as needed.
int trace_impl(const mh_scene* cs, const mh_rays* rays, const mh_hits* hits, int64_t n,
               uint32_t flags, const mh_trace_opts* opts, void* stream,
               void* fixed_spill = nullptr, size_t fixed_bytes = 0) {
  mh_scene* s = const_cast<mh_scene*>(cs);
  if (!s || !rays || !hits) return fail(MH_E_INVALID, "null scene, rays or hits");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree: call mh_scene_build first");
  if (n < 0) return fail(MH_E_INVALID, "negative ray count");
  if (flags & ~MH_ANY_HIT) return fail(MH_E_INVALID, "unknown trace flags 0x%x", flags);
  if (n == 0) return MH_OK;
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->tmax)
    return fail(MH_E_INVALID, "ray arrays must be non-null");
  const bool any = flags & MH_ANY_HIT;
  if (any ? !hits->occluded : !hits->prim)
    return fail(MH_E_INVALID, any ? "any-hit needs hits.occluded" : "closest-hit needs hits.prim");
  const int psize = opts && opts->packet_size ? opts->packet_size : 1;
  if (psize < 1 || psize > 64) return fail(MH_E_INVALID, "packet_size %d not in 1..64", psize);
  int rc;
  if ((rc = ensure_init())) return rc;

  mh::LaunchCfg cfg{};
  cfg.any_hit = any;
  cfg.packet_size = psize;
  cfg.counters = opts ? opts->counters : nullptr;
  if (psize > 1) {
    if (s->depth + 2 > 128)
      return fail(MH_E_UNSUPPORTED, "tree depth %u exceeds the packet stack (126)", s->depth);
  } else {
    const int S = choose_stack(opts);
    const int bpc = (opts && opts->blocks_per_cu > 0) ? opts->blocks_per_cu : 16;
    const int64_t want = (n + 255) / 256;
    const int64_t cap = (int64_t)g_lib.cu_count * bpc;
    cfg.blocks = (int)std::max<int64_t>(1, std::min(want, cap));
    cfg.stack_lds = S;
    cfg.spill_stride = (int64_t)cfg.blocks * 256;
    const int64_t need = std::max<int64_t>(0, (int64_t)s->stack_bound(MH_LAYOUT_PAIR) - S);
    if (need > 0 && fixed_spill) {
      // Fewer workgroups (the kernel strides over the rays, results are unchanged) until
      // every lane's spill column fits the area; at least one workgroup per compute unit.
      const int64_t per_block = need * 256 * 8;
      const int64_t fit = (int64_t)(fixed_bytes / (size_t)per_block);
      const int64_t floor_blocks = std::min<int64_t>(want, g_lib.cu_count);
      if (fit < floor_blocks)
        return fail(MH_E_UNSUPPORTED,
                    "tree depth %u needs %lld bytes of stack spill for %lld workgroups; the "
                    "fixed area holds %zu (use mh_trace, or a larger stack_lds)",
                    s->depth, (long long)(per_block * floor_blocks), (long long)floor_blocks,
                    fixed_bytes);
      cfg.blocks = (int)std::min<int64_t>(cfg.blocks, fit);
      cfg.spill_stride = (int64_t)cfg.blocks * 256;
      cfg.spill = reinterpret_cast<int32_t*>(fixed_spill);
    } else if (need > 0) {
      std::lock_guard<std::mutex> lk(s->spill_mu);
      DevBuf*& b = s->spill[stream];
      if (!b) b = new (std::nothrow) DevBuf();
      if (!b) return fail(MH_E_NOMEM, "spill descriptor");
      if ((rc = b->reserve((size_t)need * cfg.spill_stride * 8))) return rc;
      cfg.spill = b->as<int32_t>();
    }
  }
  mh::DevRays dr{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax,
                 rays->masked};
  mh::DevHits dh{hits->prim, hits->t, hits->u, hits->v, hits->occluded};
  HIP_OK((hipError_t)mh::launch_trace(dev_scene(s), dr, dh, n, cfg, stream));
  return MH_OK;
}

}  // namespace

int mh_internal::ensure_soa(mh_scene* s, void* stream) {
  const size_t n = s->pairs.bytes / sizeof(mh::PairNode);
  if (n == 0) return fail(MH_E_STATE, "scene has no pair records");
  if (s->pairs_soa.bytes < n * sizeof(mh::PairNode)) {
    if (int rc = s->pairs_soa.reserve(n * sizeof(mh::PairNode))) return rc;
    s->soa_gen = ~0ull;
  }
  if (s->soa_gen == s->pairs_gen) return MH_OK;
  // field k (16 bytes at offset 16k of every 64-byte pair) into array k
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (size_t k = 0; k < 4; ++k)
    HIP_OK(hipMemcpy2DAsync(s->pairs_soa.as<char>() + k * n * 16, 16, s->pairs.as<char>() + 16 * k,
                            sizeof(mh::PairNode), 16, n, hipMemcpyDeviceToDevice, st));
  s->soa_gen = s->pairs_gen;
  return MH_OK;
}

int mh_internal::trace_fixed_spill(const mh_scene* s, const mh_rays* rays, const mh_hits* hits,
                                   int64_t n, uint32_t flags, const mh_trace_opts* opts,
                                   void* stream, void* spill, size_t spill_bytes) {
  if (!spill) return fail(MH_E_STATE, "no spill area");
  return trace_impl(s, rays, hits, n, flags, opts, stream, spill, spill_bytes);
}

extern "C" {

int mh_abi_version(void) { return MH_ABI_VERSION; }

const char* mh_last_error(void) { return g_err.c_str(); }

int mh_init(int device) {
  {
    std::lock_guard<std::mutex> lk(g_lib.mu);
    if (g_lib.ready && g_lib.device != device)
      return fail(MH_E_STATE, "already initialised on device %d", g_lib.device);
    g_lib.device = device;
  }
  return ensure_init();
}

int mh_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", g_lib.name);
  if (cu_count) *cu_count = g_lib.cu_count;
  if (hbm_bytes) *hbm_bytes = g_lib.hbm;
  return MH_OK;
}

int mh_scene_create(const float* verts, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris,
                    mh_scene** out) {
  if (!out) return fail(MH_E_INVALID, "null output pointer");
  *out = nullptr;
  if (!verts || !idx || n_verts == 0 || n_tris == 0)
    return fail(MH_E_INVALID, "empty mesh (%u vertices, %u triangles)", n_verts, n_tris);
  if (n_tris >= mh::kMaxObjects)
    return fail(MH_E_UNSUPPORTED, "%u triangles exceeds the limit %u", n_tris, mh::kMaxObjects);
  for (size_t i = 0; i < 3 * (size_t)n_tris; ++i)
    if (idx[i] >= n_verts) return fail(MH_E_INVALID, "index %zu = %u out of range", i, idx[i]);
  int rc = ensure_init();
  if (rc) return rc;
  mh_scene* s = new (std::nothrow) mh_scene();
  if (!s) return fail(MH_E_NOMEM, "scene");
  try {
    s->verts.assign(verts, verts + 3 * (size_t)n_verts);
    s->idx.assign(idx, idx + 3 * (size_t)n_tris);
  } catch (...) {
    delete s;
    return fail(MH_E_NOMEM, "host copy of the mesh");
  }
  s->nv = n_verts;
  s->nt = n_tris;
  const size_t vb = 3 * (size_t)n_verts * 4, ib = 3 * (size_t)n_tris * 4;
  if ((rc = s->d_verts.reserve(vb)) || (rc = s->d_base.reserve(vb)) ||
      (rc = s->d_idx.reserve(ib))) {
    delete s;
    return rc;
  }
  hipError_t e = hipMemcpy(s->d_verts.p, verts, vb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->d_base.p, verts, vb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->d_idx.p, idx, ib, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete s;
    return fail(MH_E_HIP, "mesh upload: %s", hipGetErrorString(e));
  }
  *out = s;
  return MH_OK;
}

int mh_scene_destroy(mh_scene* scene) {
  if (!scene) return MH_OK;
  (void)ensure_init();
  (void)hipDeviceSynchronize();
  delete scene;
  return MH_OK;
}

// DynBVH::rebuild with USE_APPROXIMATE_BUILD, on the device (build.hip): object boxes,
// the level loop over nodes of many objects (one counter read-back per level), one lane
// per small node, then the reference numbering and the bottom-up passes.
int build_device(mh_scene* s, int rtsah) {
  using mh_internal::DevBuf;
  const uint32_t T = s->nt;
  const size_t nrec = 2 * (size_t)T;
  const int64_t cap = (int64_t)T / (mh::kBuildSerial + 1) + 2;
  const int64_t vcap = (int64_t)T / mh::kBuildVB + cap + 2;
  int rc;
  if ((rc = s->refs.reserve(nrec * sizeof(mh::RefNode))) ||
      (rc = s->pairs.reserve(std::max<size_t>(1, nrec / 2) * sizeof(mh::PairNode))) ||
      (rc = s->tris.reserve((size_t)T * 48)) || (rc = s->ids.reserve((size_t)T * 4)) ||
      (rc = s->big.reserve((size_t)T * 4)) || (rc = s->leaves.reserve((size_t)T * 4)) ||
      (rc = s->levels.reserve((size_t)T * 4)) || (rc = s->costs.reserve(nrec * 4)) ||
      (rc = s->ssize.reserve(nrec * 4)) || (rc = s->parent.reserve(nrec * 4)) ||
      (rc = s->obox.reserve((size_t)T * 24)) || (rc = s->ocen.reserve((size_t)T * 12)) ||
      (rc = s->anc0.reserve(nrec * 4)) || (rc = s->dep0.reserve(nrec * 4)) ||
      (rc = s->anc1.reserve(nrec * 4)) || (rc = s->dep1.reserve(nrec * 4)) ||
      (rc = s->hist.reserve(mh::kDepthHist * 4)) || (rc = s->counters.reserve(16)))
    return rc;
  if (!s->logtab.p) {
    std::vector<float> tab(mh::rtsah::kLogTable);
    mh::rtsah::fill_log_table(tab.data());
    if ((rc = s->logtab.reserve(tab.size() * 4))) return rc;
    HIP_OK(hipMemcpy(s->logtab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
  }
  // Work arrays of this build only.
  DevBuf slots, trefs, seg0, seg1, small, big, plan, acck, accc, ctr, vbseg, vcnt, gscan,
      scratch, rootk, ic, rank, fin, rt;
  if ((rc = slots.reserve((size_t)T * 48)) || (rc = trefs.reserve(nrec * 32)) ||
      (rc = seg0.reserve(cap * 16)) || (rc = seg1.reserve(cap * 16)) ||
      (rc = small.reserve((size_t)T * 16)) || (rc = big.reserve(cap * 16)) ||
      (rc = plan.reserve(cap * 32)) || (rc = acck.reserve(cap * mh::kKeyFields * 8)) ||
      (rc = accc.reserve(cap * mh::kCntFields * 4)) || (rc = ctr.reserve(16)) ||
      (rc = vbseg.reserve(vcap * 4)) || (rc = vcnt.reserve(vcap * 8)) || (rc = gscan.reserve(2 * (vcap + 1) * 4)) ||
      (rc = scratch.reserve(nrec * 4)) || (rc = rootk.reserve(48)) ||
      (rc = ic.reserve(nrec * 4)) || (rc = rank.reserve(nrec * 4)) ||
      (rc = fin.reserve(nrec * 4)) || (rc = rt.reserve(nrec * 8)))
    return rc;
  hipStream_t st = nullptr;
  unsigned* cnt = s->counters.as<unsigned>();  // [0] records in use, [1] error
  HIP_OK(hipMemsetAsync(cnt, 0, 16, st));
  HIP_OK(hipMemsetAsync(ctr.p, 0, 16, st));
  HIP_OK(hipMemsetAsync(rootk.p, 0xFF, 24, st));
  HIP_OK(hipMemsetAsync(rootk.as<char>() + 24, 0, 24, st));

  mh::DevBuild d{};
  d.u.refs = trefs.as<mh::RefNode>();
  d.u.costs = s->costs.as<float>();  // scratch until the bottom-up pass rewrites them
  d.u.ssize = s->ssize.as<uint32_t>();
  d.u.parent = s->parent.as<int32_t>();
  d.u.ids = s->ids.as<int32_t>();
  d.u.obox = s->obox.as<float>();
  d.u.ocen = s->ocen.as<float>();
  d.u.verts = s->d_verts.as<float>();
  d.u.idx = s->d_idx.as<uint32_t>();
  d.u.next_free = cnt;
  d.u.err = reinterpret_cast<int*>(cnt + 1);
  d.u.n_records = (int32_t)nrec;
  d.u.n_objects = (int32_t)T;
  d.u.large = (uint32_t)(nrec / (20 * (size_t)s->workers));
  d.slots = slots.as<float>();
  d.verts = s->d_verts.as<float>();
  d.n_verts = s->nv;
  d.small = small.as<mh::BuildSeg>();
  d.big = big.as<mh::BuildSeg>();
  d.plan = plan.as<mh::BuildPlan>();
  d.acck = acck.as<unsigned long long>();
  d.accc = accc.as<unsigned>();
  d.cap = cap;
  d.ctr = ctr.as<unsigned>();
  d.vbseg = vbseg.as<int32_t>();
  d.vcnt = vcnt.as<unsigned>();
  d.gscan = gscan.as<unsigned>();
  d.vcap = vcap;
  d.lpos = scratch.as<int32_t>();
  d.rpos = scratch.as<int32_t>() + T;
  d.rootk = rootk.as<unsigned long long>();
  HIP_OK((hipError_t)mh::launch_obj_bounds(d.u, st));
  const bool serial_root = T <= (uint32_t)mh::kBuildSerial;
  d.seg = seg0.as<mh::BuildSeg>();
  HIP_OK((hipError_t)mh::build_launch_setup(d, serial_root ? 1 : 0, st));
  mh::BuildSeg* cur = seg0.as<mh::BuildSeg>();
  mh::BuildSeg* nxt = seg1.as<mh::BuildSeg>();
  int64_t nseg = serial_root ? 0 : 1;
  unsigned back[mh::kCtrCount];
  int err = 0;
  while (nseg > 0) {
    d.seg = cur;
    d.seg_next = nxt;
    d.nseg = (int32_t)nseg;
    HIP_OK(hipMemsetAsync(d.ctr + mh::kCtrNext, 0, 4, st));
    const int64_t nvb_bound = (int64_t)T / mh::kBuildVB + nseg + 1;
    HIP_OK((hipError_t)mh::build_launch_level(d, nvb_bound, st));
    HIP_OK(hipMemcpyAsync(back, d.ctr, sizeof(back), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(&err, cnt + 1, 4, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (err) break;
    nseg = back[mh::kCtrNext];
    std::swap(cur, nxt);
  }
  HIP_OK(hipMemcpyAsync(back, d.ctr, sizeof(back), hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(&err, cnt + 1, 4, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (err == 2) return fail(MH_E_INVALID, "non-finite vertex coordinate");
  if (err == 3) return fail(MH_E_UNSUPPORTED, "a leaf of more than 32767 objects");
  if (err) return fail(MH_E_STATE, "device build ran out of node records");
  HIP_OK((hipError_t)mh::build_launch_serial(d, back[mh::kCtrSmall], back[mh::kCtrBig],
                                             scratch.as<int32_t>(), st));
  // Depths of the temporary tree, then its per-depth lists.
  HIP_OK(hipMemsetAsync(s->hist.p, 0, mh::kDepthHist * 4, st));
  HIP_OK((hipError_t)mh::launch_depths(d.u, s->anc0.as<int32_t>(), s->dep0.as<int32_t>(),
                                       s->anc1.as<int32_t>(), s->dep1.as<int32_t>(),
                                       s->hist.as<unsigned>(), 8, st));
  std::vector<unsigned> hist(mh::kDepthHist);
  unsigned used[2];
  HIP_OK(hipMemcpyAsync(hist.data(), s->hist.p, hist.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(used, cnt, 8, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (used[1]) return fail(MH_E_STATE, "device build ran out of node records");
  const unsigned deepest = hist[mh::kMaxDepth + 1];
  if (deepest >= (unsigned)mh::kMaxDepth - 1)
    return fail(MH_E_UNSUPPORTED, "tree deeper than %d levels", mh::kMaxDepth - 1);
  std::vector<unsigned> cursor(mh::kMaxDepth + 1, 0u);
  std::vector<int64_t> level_off(1, 0);
  for (unsigned dd = 0; dd < deepest; ++dd) {
    cursor[dd] = (unsigned)level_off.back();
    level_off.push_back(level_off.back() + hist[dd]);
  }
  const int64_t n_leaves = hist[mh::kMaxDepth];
  HIP_OK(hipMemcpyAsync(s->hist.p, cursor.data(), cursor.size() * 4, hipMemcpyHostToDevice, st));
  HIP_OK((hipError_t)mh::launch_depth_scatter(d.u, s->dep0.as<int32_t>(), s->hist.as<unsigned>(),
                                              s->levels.as<int32_t>(), s->leaves.as<int32_t>(),
                                              st));
  HIP_OK(hipMemsetAsync(fin.p, 0xFF, nrec * 4, st));
  mh::DevFinish f{};
  f.trefs = trefs.as<mh::RefNode>();
  f.refs = s->refs.as<mh::RefNode>();
  f.tcount = cnt;
  f.ic = ic.as<int32_t>();
  f.rank = rank.as<int32_t>();
  f.fin = fin.as<int32_t>();
  f.levels = s->levels.as<int32_t>();
  f.leaves = s->leaves.as<int32_t>();
  f.costs = s->costs.as<float>();
  f.ssize = s->ssize.as<uint32_t>();
  f.rt = rt.as<float>();
  f.logtab = s->logtab.as<float>();
  const uint32_t n_final = (uint32_t)(level_off.back() + n_leaves);
  f.n_records = (int32_t)nrec;
  f.n_final = (int32_t)n_final;
  f.large = (uint32_t)(nrec / (20 * (size_t)s->workers));
  f.rtsah = rtsah;
  const int n_levels = (int)level_off.size() - 1;
  ++s->pairs_gen;
  HIP_OK((hipError_t)mh::build_launch_ranks(f, level_off.data(), n_levels, n_leaves, st));
  HIP_OK(hipMemcpyAsync(cnt, &n_final, 4, hipMemcpyHostToDevice, st));
  mh::DevUpdate u = d.u;
  u.refs = s->refs.as<mh::RefNode>();
  u.pairs = s->pairs.as<mh::PairNode>();
  u.big_count = s->big.as<int32_t>();
  HIP_OK((hipError_t)mh::launch_rebuild_pairs(u, st));
  const bool root_leaf = n_final == 1;
  if (root_leaf && T >= (uint32_t)mh::kLeafCountInline) {
    const int32_t n = (int32_t)T;
    HIP_OK(hipMemcpyAsync(s->big.p, &n, 4, hipMemcpyHostToDevice, st));
  }
  mh::DevRefit r{};
  r.tris = s->tris.as<float>();
  r.obj_ids = s->ids.as<int32_t>();
  r.verts = s->d_verts.as<float>();
  r.idx = s->d_idx.as<uint32_t>();
  r.n_slots = (int32_t)T;
  r.prim_kind = s->prim_kind;
  HIP_OK((hipError_t)mh::launch_prim_setup(r, st));
  HIP_OK(hipStreamSynchronize(st));
  s->root_ref = root_leaf ? ~(int)std::min<uint32_t>(T, mh::kLeafCountInline) : 0;
  s->next_free = n_final;
  s->n_records = (uint32_t)nrec;
  s->n_inner = (uint32_t)level_off.back();
  s->leaf_at_valid = false;
  s->n_leaves = (uint32_t)n_leaves;
  s->depth = deepest;
  s->level_off = std::move(level_off);
  s->bvh = mh::DynBvh();
  s->updatable = true;
  s->built = true;
  return MH_OK;
}

int mh_scene_build(mh_scene* s, const mh_build_opts* opts) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  const int kind = opts ? opts->prim_kind : MH_PRIM_KENSLER_SHIRLEY;
  if (kind != MH_PRIM_KENSLER_SHIRLEY && kind != MH_PRIM_WALD)
    return fail(MH_E_INVALID, "unknown prim_kind %d", kind);
  const int rtsah = opts ? opts->rtsah : 1;
  const int threads = opts ? opts->threads : 0;
  int rc = ensure_init();
  if (rc) return rc;
  if (opts && opts->approximate) {
    s->built = s->updatable = false;
    s->prim_kind = kind;
    return build_device(s, rtsah);
  }
  try {
    s->bvh = mh::build_dynbvh(s->verts.data(), s->nv, s->idx.data(), s->nt, threads);
    if (s->workers != 1) mh::mark_large_subtrees(s->bvh.nodes, s->workers);
    if (rtsah) mh::compute_rtsah(s->bvh.nodes, threads);
  } catch (const std::bad_alloc&) {
    return fail(MH_E_NOMEM, "host memory for the SAH build");
  } catch (const std::exception& ex) {
    return fail(MH_E_INVALID, "build: %s", ex.what());
  }
  s->prim_kind = kind;
  return upload(s);
}

int mh_scene_set_workers(mh_scene* s, int workers) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  if (workers < 1 || workers > 4096) return fail(MH_E_INVALID, "workers %d not in 1..4096", workers);
  s->workers = workers;
  if (!s->built) return MH_OK;  // applies to the next build
  int rc = ensure_init();
  if (rc) return rc;
  // computeSubTreeSizes over the current tree with the new threshold (DynBVH.cc:941-950,
  // 1971-1996): reachable records only, sizes counted in records.
  const uint32_t nn = s->n_records;
  std::vector<mh::RefNode> nodes(nn);
  HIP_OK(hipMemcpy(nodes.data(), s->refs.p, (size_t)nn * sizeof(mh::RefNode),
                   hipMemcpyDeviceToHost));
  const uint32_t limit = s->large_limit();
  std::vector<uint32_t> size(nn, 0);
  std::vector<std::pair<int32_t, bool>> st{{0, false}};
  while (!st.empty()) {
    auto [id, done] = st.back();
    st.pop_back();
    if (id < 0 || (uint32_t)id >= nn) return fail(MH_E_STATE, "record %d outside the tree", id);
    mh::RefNode& n = nodes[id];
    if (n.leaf()) {
      size[id] = 1;
      n.bits &= (uint8_t)~2;
      continue;
    }
    if (!done) {
      if ((uint32_t)n.child + 1 >= nn) return fail(MH_E_STATE, "record %d: child outside", id);
      st.push_back({id, true});
      st.push_back({n.child + 1, false});
      st.push_back({n.child, false});
      continue;
    }
    size[id] = size[n.child] + size[n.child + 1] + 1;
    if (size[id] > limit)
      n.bits |= 2;
    else
      n.bits &= (uint8_t)~2;
  }
  HIP_OK(hipMemcpy(s->refs.p, nodes.data(), (size_t)nn * sizeof(mh::RefNode),
                   hipMemcpyHostToDevice));
  if (s->bvh.nodes.size() == nodes.size()) s->bvh.nodes = nodes;
  return MH_OK;
}

int mh_dynbvh_build_host(const float* verts, uint32_t n_verts, const uint32_t* idx,
                         uint32_t n_tris, const mh_build_opts* opts, int32_t* ids, void* nodes,
                         uint32_t* num_nodes) {
  if (!verts || !idx || !ids || !nodes) return fail(MH_E_INVALID, "null argument");
  if (n_verts == 0 || n_tris == 0) return fail(MH_E_INVALID, "empty mesh");
  if (n_tris >= mh::kMaxObjects)
    return fail(MH_E_UNSUPPORTED, "%u triangles exceeds the limit %u", n_tris, mh::kMaxObjects);
  for (size_t i = 0; i < 3 * (size_t)n_tris; ++i)
    if (idx[i] >= n_verts) return fail(MH_E_INVALID, "index %zu = %u out of range", i, idx[i]);
  const int rtsah = opts ? opts->rtsah : 1;
  const int threads = opts ? opts->threads : 0;
  try {
    mh::DynBvh b = mh::build_dynbvh(verts, n_verts, idx, n_tris, threads);
    if (rtsah) mh::compute_rtsah(b.nodes, threads);
    std::memcpy(ids, b.ids.data(), b.ids.size() * 4);
    std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(mh::RefNode));
    if (num_nodes) *num_nodes = (uint32_t)b.num_nodes;
  } catch (const std::bad_alloc&) {
    return fail(MH_E_NOMEM, "host memory for the SAH build");
  } catch (const std::exception& ex) {
    return fail(MH_E_INVALID, "build: %s", ex.what());
  }
  return MH_OK;
}

namespace {

int load_tree(mh_scene* s, const int32_t* ids, uint32_t n_ids, const void* nodes,
              uint32_t n_nodes, int prim_kind, bool rtsah) {
  if (!s || !ids || !nodes) return fail(MH_E_INVALID, "null argument");
  if (prim_kind != MH_PRIM_KENSLER_SHIRLEY && prim_kind != MH_PRIM_WALD)
    return fail(MH_E_INVALID, "unknown prim_kind %d", prim_kind);
  if (n_nodes == 0) return fail(MH_E_INVALID, "empty node array");
  int rc = ensure_init();
  if (rc) return rc;
  std::vector<int32_t> idv(ids, ids + n_ids);
  if ((rc = validate_ids(idv, s->nt))) return rc;
  s->bvh.ids = std::move(idv);
  const mh::RefNode* nn = static_cast<const mh::RefNode*>(nodes);
  s->bvh.nodes.assign(nn, nn + n_nodes);
  s->bvh.num_nodes = 0;
  s->prim_kind = prim_kind;
  s->built = false;
  return upload(s, rtsah);
}

}  // namespace

int mh_scene_load_dynbvh(mh_scene* s, const int32_t* ids, uint32_t n_ids, const void* nodes,
                         uint32_t n_nodes, int prim_kind) {
  return load_tree(s, ids, n_ids, nodes, n_nodes, prim_kind, false);
}

int mh_scene_load_dynbvh_file(mh_scene* s, const char* path, int prim_kind) {
  if (!s || !path) return fail(MH_E_INVALID, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(MH_E_IO, "cannot open %s", path);
  uint32_t n_ids = 0, n_nodes = 0;
  in.read(reinterpret_cast<char*>(&n_ids), 4);
  if (!in || n_ids != s->nt) return fail(MH_E_IO, "%s: bad object_ids count", path);
  std::vector<int32_t> ids(n_ids);
  in.read(reinterpret_cast<char*>(ids.data()), (std::streamsize)n_ids * 4);
  in.read(reinterpret_cast<char*>(&n_nodes), 4);
  if (!in || n_nodes == 0 || n_nodes > 4u * n_ids + 4u) return fail(MH_E_IO, "%s: bad node count", path);
  std::vector<mh::RefNode> nodes(n_nodes);
  in.read(reinterpret_cast<char*>(nodes.data()), (std::streamsize)n_nodes * 32);
  if (!in) return fail(MH_E_IO, "%s: truncated", path);
  return load_tree(s, ids.data(), n_ids, nodes.data(), n_nodes, prim_kind, true);
}

int mh_scene_export_dynbvh(const mh_scene* s, int32_t* ids, uint32_t* n_ids, void* nodes,
                           uint32_t* n_nodes) {
  if (!s || !n_ids || !n_nodes) return fail(MH_E_INVALID, "null argument");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree");
  const uint32_t ni = s->nt, nn = s->n_records;
  if (!ids && !nodes) {
    *n_ids = ni;
    *n_nodes = nn;
    return MH_OK;
  }
  if (*n_ids < ni || *n_nodes < nn) return fail(MH_E_INVALID, "export arrays too small");
  int rc = ensure_init();
  if (rc) return rc;
  if (ids) HIP_OK(hipMemcpy(ids, s->ids.p, (size_t)ni * 4, hipMemcpyDeviceToHost));
  if (nodes) HIP_OK(hipMemcpy(nodes, s->refs.p, (size_t)nn * 32, hipMemcpyDeviceToHost));
  *n_ids = ni;
  *n_nodes = nn;
  return MH_OK;
}

int mh_scene_save_dynbvh_file(const mh_scene* s, const char* path) {
  if (!s || !path) return fail(MH_E_INVALID, "null argument");
  uint32_t ni = 0, nn = 0;
  int rc = mh_scene_export_dynbvh(s, nullptr, &ni, nullptr, &nn);
  if (rc) return rc;
  std::vector<int32_t> ids(ni);
  std::vector<mh::RefNode> nodes(nn);
  if ((rc = mh_scene_export_dynbvh(s, ids.data(), &ni, nodes.data(), &nn))) return rc;
  std::ofstream out(path, std::ios::binary);
  if (!out) return fail(MH_E_IO, "cannot create %s", path);
  out.write(reinterpret_cast<const char*>(&ni), 4);
  out.write(reinterpret_cast<const char*>(ids.data()), (std::streamsize)ni * 4);
  out.write(reinterpret_cast<const char*>(&nn), 4);
  out.write(reinterpret_cast<const char*>(nodes.data()), (std::streamsize)nn * 32);
  if (!out) return fail(MH_E_IO, "write to %s failed", path);
  return MH_OK;
}

int mh_scene_bounds(const mh_scene* s, float lo[3], float hi[3]) {
  if (!s || !lo || !hi) return fail(MH_E_INVALID, "null argument");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree");
  int rc = ensure_init();
  if (rc) return rc;
  mh::RefNode root;
  HIP_OK(hipMemcpy(&root, s->refs.p, sizeof(root), hipMemcpyDeviceToHost));
  for (int k = 0; k < 3; ++k) {
    lo[k] = root.box.lo[k];
    hi[k] = root.box.hi[k];
  }
  return MH_OK;
}

int mh_scene_stats(const mh_scene* s, uint32_t* n_nodes, uint32_t* n_inner, uint32_t* depth,
                   size_t* device_bytes) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree");
  if (n_nodes) *n_nodes = s->n_inner + s->n_leaves;
  if (n_inner) *n_inner = s->n_inner;
  if (depth) *depth = s->depth;
  if (device_bytes) *device_bytes = s->device_bytes();
  return MH_OK;
}

int mh_scene_set_vertices(mh_scene* s, const float* verts, int on_device, void* stream) {
  if (!s || !verts) return fail(MH_E_INVALID, "null argument");
  int rc = ensure_init();
  if (rc) return rc;
  const size_t vb = 3 * (size_t)s->nv * 4;
  HIP_OK(hipMemcpyAsync(s->d_verts.p, verts, vb,
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        reinterpret_cast<hipStream_t>(stream)));
  if (!on_device) std::memcpy(s->verts.data(), verts, vb);
  return MH_OK;
}

int mh_scene_vertices_device(mh_scene* s, float** verts) {
  if (!s || !verts) return fail(MH_E_INVALID, "null argument");
  *verts = s->d_verts.as<float>();
  return MH_OK;
}

int mh_scene_displace_sine(mh_scene* s, float amp, float freq, float phase, void* stream) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  int rc = ensure_init();
  if (rc) return rc;
  HIP_OK((hipError_t)mh::launch_displace_sine(s->d_verts.as<float>(), s->d_base.as<float>(),
                                              s->nv, amp, freq, phase, stream));
  return MH_OK;
}

int mh_scene_refit(mh_scene* s, void* stream) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree");
  int rc = ensure_init();
  if (rc) return rc;
  mh::DevRefit r{};
  r.refs = s->refs.as<mh::RefNode>();
  r.pairs = s->pairs.as<mh::PairNode>();
  r.tris = s->tris.as<float>();
  r.obj_ids = s->ids.as<int32_t>();
  r.verts = s->d_verts.as<float>();
  r.idx = s->d_idx.as<uint32_t>();
  r.leaves = s->leaves.as<int32_t>();
  r.levels = s->levels.as<int32_t>();
  r.n_leaves = (int32_t)s->n_leaves;
  r.n_slots = (int32_t)s->nt;
  r.prim_kind = s->prim_kind;
  ++s->pairs_gen;
  HIP_OK((hipError_t)mh::launch_refit(r, s->level_off.data(), (int)s->level_off.size() - 1,
                                      stream));
  return MH_OK;
}

int mh_scene_update(mh_scene* s, void* stream) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  if (!s->built) return fail(MH_E_STATE, "scene has no tree");
  if (!s->updatable)
    return fail(MH_E_UNSUPPORTED, "the tree has records outside it or leaves that do not "
                "partition the objects; use mh_scene_refit");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const uint32_t T = s->nt;
  const size_t nrec = s->refs.bytes / sizeof(mh::RefNode);
  // Work arrays, allocated at the first update.
  if ((rc = s->parent.reserve(nrec * 4)) || (rc = s->hist.reserve(mh::kFrontierWords * 4)) ||
      (rc = s->counters.reserve(16)) ||
      (rc = s->split_list.reserve((size_t)std::max<uint32_t>(1u, T) * 4)) ||
      (rc = s->levels_aux.reserve((size_t)std::max<uint32_t>(1u, T) * 8)) ||
      (rc = s->leaf_at.reserve((size_t)std::max<uint32_t>(1u, T) * 4)))
    return rc;
  if (!s->host_ctr) {
    void* h = nullptr;
    if (hipHostMalloc(&h, (mh::kFrontierWords + 2) * 4, hipHostMallocDefault) != hipSuccess)
      return fail(MH_E_NOMEM, "hipHostMalloc of the update's counter copy");
    s->host_ctr = static_cast<unsigned*>(h);
  }
  unsigned* nfree = s->counters.as<unsigned>();
  int* err = reinterpret_cast<int*>(nfree + 1);
  const unsigned init[3] = {s->next_free, 0u, 0u};  // next_free, err, split list entries
  HIP_OK(hipMemcpyAsync(nfree, init, sizeof(init), hipMemcpyHostToDevice, st));
  mh::DevUpdate u{};
  u.refs = s->refs.as<mh::RefNode>();
  u.pairs = s->pairs.as<mh::PairNode>();
  u.big_count = s->big.as<int32_t>();
  u.costs = s->costs.as<float>();
  u.ssize = s->ssize.as<uint32_t>();
  u.parent = s->parent.as<int32_t>();
  u.ids = s->ids.as<int32_t>();
  u.obox = nullptr;  // object boxes and centroids from the vertices, where they are needed
  u.ocen = nullptr;
  u.verts = s->d_verts.as<float>();
  u.idx = s->d_idx.as<uint32_t>();
  u.next_free = nfree;
  u.err = err;
  u.split_list = s->split_list.as<int32_t>();
  u.split_count = nfree + 2;
  u.leaf_at = s->leaf_at.as<int32_t>();
  u.n_records = (int32_t)nrec;
  u.n_objects = (int32_t)T;
  // largeSubtreeSize = nodes.size() / (20 * numProcs) (DynBVH.cc:1776)
  u.large = (uint32_t)(nrec / (20 * (size_t)s->workers));
  ++s->pairs_gen;
  if (!s->leaf_at_valid) {  // after a build or load: the leaf starting at each slot
    HIP_OK(hipMemsetAsync(s->leaf_at.p, 0xFF, (size_t)std::max<uint32_t>(1u, T) * 4, st));
    HIP_OK((hipError_t)mh::launch_leaf_starts(u, s->leaves.as<int32_t>(), s->n_leaves, st));
    s->leaf_at_valid = true;  // the level pass below keeps it for the next frame
  }
  // leaf pass (triangle records and leaf boxes from the current vertices), leaf splits,
  // rotations depth by depth
  HIP_OK((hipError_t)mh::launch_update_tree(u, s->n_leaves, s->levels.as<int32_t>(),
                                            s->level_off.data(),
                                            (int)s->level_off.size() - 1, st));
  mh::DevRefit r{};  // triangle records in slot order from the current vertices
  r.tris = s->tris.as<float>();
  r.obj_ids = s->ids.as<int32_t>();
  r.verts = s->d_verts.as<float>();
  r.idx = s->d_idx.as<uint32_t>();
  r.n_slots = (int32_t)T;
  r.prim_kind = s->prim_kind;
  HIP_OK((hipError_t)mh::launch_prim_setup(r, st));
  // Level lists, leaf list and PairNodes of the new topology, top down, one pass per depth
  // (the small depths at the top in one launch). The first batch of passes covers the
  // previous depth with two levels to spare; a tree
  // that grew deeper than that gets more passes after the counters are read back.
  // MH_UPDATE_FRONTIER_LEVELS (1 or more) sets the first batch's size instead, so that
  // tests can reach the later batches.
  const int prev_levels = (int)s->level_off.size() - 1;
  const unsigned cap = (unsigned)std::max<uint32_t>(1u, T);
  unsigned* fr = s->hist.as<unsigned>();
  unsigned* hc = s->host_ctr;
  const bool inner_root = s->n_inner > 0;  // rotations never turn a leaf root inner
  int d = 0, d_end = 0;
  if (inner_root) {
    d_end = std::min(prev_levels + 2, mh::kMaxDepth - 1);
    if (const char* e = std::getenv("MH_UPDATE_FRONTIER_LEVELS")) {
      const int v = std::atoi(e);
      if (v >= 1) d_end = std::min(v, mh::kMaxDepth - 1);
    }
    // the depths that were small in the previous frame run in one workgroup
    mh::TopLevels t;
    const int n_top = std::max(1, std::min(mh::top_levels(s->level_off.data(), prev_levels, t),
                                           d_end));
    HIP_OK((hipError_t)mh::launch_frontier_top(u, fr, s->levels.as<int32_t>(),
                                               s->levels_aux.as<int2>(),
                                               s->leaves.as<int32_t>(), cap, n_top, st));
    d = n_top;
  }
  for (;;) {
    for (; d < d_end; ++d) {
      // workgroups for the depth's size in the previous frame (a grid-stride loop covers
      // any other size)
      const int64_t prev = d < prev_levels ? s->level_off[d + 1] - s->level_off[d] : 256;
      const unsigned blocks = (unsigned)std::min<int64_t>(2048, (prev + prev / 4 + 255) / 256);
      HIP_OK((hipError_t)mh::launch_frontier(u, fr, s->levels.as<int32_t>(),
                                             s->levels_aux.as<int2>(),
                                             s->leaves.as<int32_t>(), cap, d, blocks, st));
    }
    if (inner_root)
      HIP_OK(hipMemcpyAsync(hc, fr, mh::kFrontierWords * 4, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hc + mh::kFrontierWords, nfree, 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (!inner_root || hc[d_end] == 0 || d_end >= mh::kMaxDepth - 1) break;
    d_end = std::min(d_end + 16, mh::kMaxDepth - 1);
  }
  // Either failure leaves a device tree that must not be traced: the scene needs a build.
  if (hc[mh::kFrontierWords + 1]) {
    s->built = s->updatable = false;
    return fail(MH_E_STATE, "leaf splits ran out of node records; build the scene again");
  }
  s->next_free = hc[mh::kFrontierWords];
  if (!inner_root) {
    s->n_leaves = 1;
    return MH_OK;
  }
  unsigned deepest = 0;
  while ((int)deepest < d_end && hc[deepest]) ++deepest;
  if (hc[d_end] || deepest >= (unsigned)mh::kMaxDepth - 1) {
    s->built = s->updatable = false;
    return fail(MH_E_UNSUPPORTED, "tree deeper than %d levels after rotations; build the "
                "scene again", mh::kMaxDepth - 1);
  }
  s->level_off.assign(1, 0);
  for (unsigned k = 0; k < deepest; ++k) s->level_off.push_back(s->level_off.back() + hc[k]);
  const uint64_t n_inner = (uint64_t)s->level_off.back(), n_leaves = hc[mh::kMaxDepth];
  if (hc[mh::kMaxDepth + 1] || n_leaves != n_inner + 1 || n_leaves > cap) {
    s->built = s->updatable = false;
    return fail(MH_E_STATE, "the updated records are not a binary tree (%llu inner, %llu "
                "leaves); build the scene again", (unsigned long long)n_inner,
                (unsigned long long)n_leaves);
  }
  s->n_inner = (uint32_t)n_inner;
  s->n_leaves = (uint32_t)n_leaves;
  s->depth = deepest;
  return MH_OK;
}

int mh_scene_set_attributes(mh_scene* s, const float* normals, uint32_t n_normals,
                            const uint32_t* normal_idx, const float* texcoords,
                            uint32_t n_texcoords, const uint32_t* tex_idx) {
  if (!s) return fail(MH_E_INVALID, "null scene");
  if (!normals != !normal_idx || !texcoords != !tex_idx)
    return fail(MH_E_INVALID, "an attribute needs both its values and its indices");
  const size_t n3 = 3 * (size_t)s->nt;
  if (normals) {
    for (size_t k = 0; k < n3; ++k)
      if (normal_idx[k] >= n_normals)
        return fail(MH_E_INVALID, "normal index %zu = %u out of range", k, normal_idx[k]);
  }
  if (texcoords) {
    for (size_t t = 0; t < s->nt; ++t) {
      if (tex_idx[3 * t] == MH_NO_TEXTURE_INDEX) continue;
      for (int k = 0; k < 3; ++k)
        if (tex_idx[3 * t + k] >= n_texcoords)
          return fail(MH_E_INVALID, "texture index %zu = %u out of range", 3 * t + k,
                      tex_idx[3 * t + k]);
    }
  }
  int rc = ensure_init();
  if (rc) return rc;
  s->normals.release();
  s->normal_idx.release();
  s->texcoords.release();
  s->tex_idx.release();
  if (normals) {
    if ((rc = s->normals.reserve(std::max<size_t>(1, n_normals) * 12)) ||
        (rc = s->normal_idx.reserve(n3 * 4)))
      return rc;
    HIP_OK(hipMemcpy(s->normals.p, normals, (size_t)n_normals * 12, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->normal_idx.p, normal_idx, n3 * 4, hipMemcpyHostToDevice));
  }
  if (texcoords) {
    if ((rc = s->texcoords.reserve(std::max<size_t>(1, n_texcoords) * 12)) ||
        (rc = s->tex_idx.reserve(n3 * 4)))
      return rc;
    HIP_OK(hipMemcpy(s->texcoords.p, texcoords, (size_t)n_texcoords * 12,
                     hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->tex_idx.p, tex_idx, n3 * 4, hipMemcpyHostToDevice));
  }
  return MH_OK;
}

int mh_shade(const mh_scene* s, const mh_rays* rays, const mh_hits* hits, int64_t n,
             const mh_shading* out, void* stream) {
  if (!s || !rays || !hits || !out) return fail(MH_E_INVALID, "null argument");
  if (n < 0) return fail(MH_E_INVALID, "negative ray count");
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
    return fail(MH_E_INVALID, "ray origin and direction arrays are required");
  if (!hits->prim || !hits->t || !hits->u || !hits->v)
    return fail(MH_E_INVALID, "hits need prim, t, u and v");
  if (n == 0) return MH_OK;
  int rc = ensure_init();
  if (rc) return rc;
  mh::DevShade d{};
  d.rays.ox = rays->ox;
  d.rays.oy = rays->oy;
  d.rays.oz = rays->oz;
  d.rays.dx = rays->dx;
  d.rays.dy = rays->dy;
  d.rays.dz = rays->dz;
  d.prim = hits->prim;
  d.t = hits->t;
  d.u = hits->u;
  d.v = hits->v;
  d.verts = s->d_verts.as<float>();
  d.idx = s->d_idx.as<uint32_t>();
  d.normals = s->normals.as<float>();
  d.normal_idx = s->normal_idx.as<uint32_t>();
  d.tex = s->texcoords.as<float>();
  d.tex_idx = s->tex_idx.as<uint32_t>();
  float* o[12] = {out->px, out->py, out->pz, out->nx, out->ny, out->nz,
                  out->gx, out->gy, out->gz, out->tu, out->tv, out->tw};
  for (int k = 0; k < 12; ++k) d.out[k] = o[k];
  d.n = n;
  d.n_tris = (int32_t)s->nt;
  d.prim_kind = s->prim_kind;
  HIP_OK((hipError_t)mh::launch_shade(d, stream));
  return MH_OK;
}

int mh_trace(const mh_scene* s, const mh_rays* rays, const mh_hits* hits, int64_t n,
             uint32_t flags, const mh_trace_opts* opts, void* stream) {
  return trace_impl(s, rays, hits, n, flags, opts, stream);
}

// Ray count up to which mh_trace_host packs rays and results into pinned host memory.
constexpr int64_t kPinnedStageRays = int64_t(1) << 16;

int mh_trace_host(const mh_scene* cs, const mh_rays* rays, const mh_hits* hits, int64_t n,
                  uint32_t flags, const mh_trace_opts* opts) {
  mh_scene* s = const_cast<mh_scene*>(cs);
  if (!s || !rays || !hits) return fail(MH_E_INVALID, "null argument");
  if (n <= 0) return n == 0 ? MH_OK : fail(MH_E_INVALID, "negative ray count");
  int rc = ensure_init();
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(s->stage_mu);
  // One staging layout in device memory (and, for small sets, in pinned host memory): 7
  // float arrays and the mask bytes in, 4 result arrays and the occluded bytes out, on the
  // scene's own stream.
  const size_t f = (size_t)n * 4, al = 256;
  auto up = [&](size_t x) { return (x + al - 1) / al * al; };
  const size_t in_bytes = 7 * up(f) + up((size_t)n);
  const size_t out_bytes = 4 * up(f) + up((size_t)n);
  if ((rc = s->stage.reserve(in_bytes + out_bytes))) return rc;
  // A blocking stream: work the caller queued on the null stream (a refit, an upload)
  // comes first, as it did for the null-stream copies of this entry point.
  if (!s->stage_stream) HIP_OK(hipStreamCreateWithFlags(&s->stage_stream, hipStreamDefault));
  hipStream_t st = s->stage_stream;
  char* base = s->stage.as<char>();
  if (n > kPinnedStageRays) {
    // Large sets: the runtime's own pipelined copies from pageable memory, array by array.
    const float* src[7] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
    float* d[7];
    for (int i = 0; i < 7; ++i) {
      if (!src[i]) return fail(MH_E_INVALID, "ray array %d is null", i);
      d[i] = reinterpret_cast<float*>(base + i * up(f));
      HIP_OK(hipMemcpyAsync(d[i], src[i], f, hipMemcpyHostToDevice, st));
    }
    uint8_t* dm = nullptr;
    if (rays->masked) {
      dm = reinterpret_cast<uint8_t*>(base + 7 * up(f));
      HIP_OK(hipMemcpyAsync(dm, rays->masked, (size_t)n, hipMemcpyHostToDevice, st));
    }
    char* ob = base + in_bytes;
    mh_hits dh{};
    dh.prim = hits->prim ? reinterpret_cast<int32_t*>(ob) : nullptr;
    dh.t = hits->t ? reinterpret_cast<float*>(ob + up(f)) : nullptr;
    dh.u = hits->u ? reinterpret_cast<float*>(ob + 2 * up(f)) : nullptr;
    dh.v = hits->v ? reinterpret_cast<float*>(ob + 3 * up(f)) : nullptr;
    dh.occluded = hits->occluded ? reinterpret_cast<uint8_t*>(ob + 4 * up(f)) : nullptr;
    mh_rays dr{d[0], d[1], d[2], d[3], d[4], d[5], d[6], dm};
    if ((rc = trace_impl(s, &dr, &dh, n, flags, opts, st))) return rc;
    if (hits->prim) HIP_OK(hipMemcpyAsync(hits->prim, dh.prim, f, hipMemcpyDeviceToHost, st));
    if (hits->t) HIP_OK(hipMemcpyAsync(hits->t, dh.t, f, hipMemcpyDeviceToHost, st));
    if (hits->u) HIP_OK(hipMemcpyAsync(hits->u, dh.u, f, hipMemcpyDeviceToHost, st));
    if (hits->v) HIP_OK(hipMemcpyAsync(hits->v, dh.v, f, hipMemcpyDeviceToHost, st));
    if (hits->occluded)
      HIP_OK(hipMemcpyAsync(hits->occluded, dh.occluded, (size_t)n, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return MH_OK;
  }
  // Small sets (the packets of a render thread gathered by HipDynBVH): packed into pinned
  // host memory, one copy each way.
  if (s->host_stage_bytes < in_bytes + out_bytes) {
    if (s->host_stage) (void)hipHostFree(s->host_stage);
    s->host_stage = nullptr;
    s->host_stage_bytes = 0;
    HIP_OK(hipHostMalloc(&s->host_stage, in_bytes + out_bytes, hipHostMallocDefault));
    s->host_stage_bytes = in_bytes + out_bytes;
  }
  char* hbase = static_cast<char*>(s->host_stage);
  const float* src[7] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
  float* d[7];
  for (int i = 0; i < 7; ++i) {
    if (!src[i]) return fail(MH_E_INVALID, "ray array %d is null", i);
    std::memcpy(hbase + i * up(f), src[i], f);
    d[i] = reinterpret_cast<float*>(base + i * up(f));
  }
  uint8_t* dm = nullptr;
  if (rays->masked) {
    std::memcpy(hbase + 7 * up(f), rays->masked, (size_t)n);
    dm = reinterpret_cast<uint8_t*>(base + 7 * up(f));
  }
  HIP_OK(hipMemcpyAsync(base, hbase, rays->masked ? in_bytes : 7 * up(f), hipMemcpyHostToDevice,
                        st));
  char* ob = base + in_bytes;
  mh_hits dh{};
  dh.prim = hits->prim ? reinterpret_cast<int32_t*>(ob) : nullptr;
  dh.t = hits->t ? reinterpret_cast<float*>(ob + up(f)) : nullptr;
  dh.u = hits->u ? reinterpret_cast<float*>(ob + 2 * up(f)) : nullptr;
  dh.v = hits->v ? reinterpret_cast<float*>(ob + 3 * up(f)) : nullptr;
  dh.occluded = hits->occluded ? reinterpret_cast<uint8_t*>(ob + 4 * up(f)) : nullptr;
  mh_rays dr{d[0], d[1], d[2], d[3], d[4], d[5], d[6], dm};
  if ((rc = trace_impl(s, &dr, &dh, n, flags, opts, st))) return rc;
  HIP_OK(hipMemcpyAsync(hbase + in_bytes, ob, out_bytes, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  const char* ho = hbase + in_bytes;
  if (hits->prim) std::memcpy(hits->prim, ho, f);
  if (hits->t) std::memcpy(hits->t, ho + up(f), f);
  if (hits->u) std::memcpy(hits->u, ho + 2 * up(f), f);
  if (hits->v) std::memcpy(hits->v, ho + 3 * up(f), f);
  if (hits->occluded) std::memcpy(hits->occluded, ho + 4 * up(f), (size_t)n);
  return MH_OK;
}

int mh_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(MH_E_INVALID, "null argument");
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail(MH_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return MH_OK;
}

int mh_free(void* ptr) {
  if (!ptr) return MH_OK;
  int rc = ensure_init();
  if (rc) return rc;
  HIP_OK(hipFree(ptr));
  return MH_OK;
}

int mh_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return MH_OK;
}

int mh_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return MH_OK;
}

int mh_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if ((!dst || !src) && bytes) return fail(MH_E_INVALID, "null argument");
  int rc = ensure_init();
  if (rc) return rc;
  HIP_OK((hipError_t)mh::launch_copy(dst, src, bytes, g_lib.cu_count, stream));
  return MH_OK;
}

int mh_stream_create(void** stream) {
  if (!stream) return fail(MH_E_INVALID, "null argument");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t s;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return MH_OK;
}

int mh_stream_destroy(void* stream) {
  if (!stream) return MH_OK;
  HIP_OK(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)));
  return MH_OK;
}

int mh_stream_sync(void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  HIP_OK(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
  return MH_OK;
}

int mh_time_trace(const mh_scene* s, const mh_rays* rays, const mh_hits* hits, int64_t n,
                  uint32_t flags, const mh_trace_opts* opts, void* stream, int iters, float* ms) {
  if (!ms || iters < 1) return fail(MH_E_INVALID, "bad timing arguments");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  HIP_OK(hipEventRecord(a, st));
  for (int i = 0; i < iters; ++i)
    if ((rc = trace_impl(s, rays, hits, n, flags, opts, stream))) {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      return rc;
    }
  HIP_OK(hipEventRecord(b, st));
  HIP_OK(hipEventSynchronize(b));
  HIP_OK(hipEventElapsedTime(ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return MH_OK;
}

}  // extern "C"
