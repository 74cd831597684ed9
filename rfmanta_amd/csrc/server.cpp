// Serving loop (mh_server_*, include/mh.h): resident scenes, HBM-budget admission and
// continuous batching over the persistent batch kernel.
//
// Threads: callers submit and wait; one scheduler thread owns the device work. Each
// scheduler round
//   1. admits queued requests in arrival order while their buffers fit the arena (the
//      HBM budget) and fewer than max_active are in flight, generating their camera rays
//      on the device;
//   2. builds, per scene, a batch table with one chunk of every admitted request's
//      remaining rays (primary closest-hit first, then its shadow rays, generated on the
//      device once the primary rays are done) and launches one persistent trace per
//      scene, each scene on its own stream so that the launches overlap;
//   3. waits for the launches, and completes every request whose rays are all traced:
//      its hits are copied to the host and its arena range is freed.
// A request therefore joins the device work at the first round after it fits, and large
// requests never hold small ones back for more than one round.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "internal.h"
#include "mh.h"
#include "mh_device.h"

using mh_internal::DevBuf;
using mh_internal::ensure_init;
using mh_internal::fail;

namespace {

// First-fit range allocator over [0, size) with coalescing free list.
class Arena {
 public:
  explicit Arena(size_t size) { free_[0] = size; }
  bool alloc(size_t n, size_t& off) {
    n = (n + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        off = it->first;
        const size_t rest = it->second - n;
        const size_t start = it->first + n;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += n;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~size_t(255);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0;
};

struct Request {
  uint64_t id = 0;
  int scene = 0;
  mh_camera cam{};
  int64_t n = 0;  // primary rays
  int status = MH_OK;
  double t_submit = 0, t_admit = 0, t_done = 0;
  // device arena range and the arrays inside it
  size_t off = 0, bytes = 0;
  mh_rays prim{};
  mh_hits hits{};
  mh_rays shadow{};
  uint8_t* occluded = nullptr;
  int32_t* shadow_prim = nullptr;
  int64_t next = 0, sh_next = 0, sh_done_rounds = 0;
  bool primary_done = false, shadow_generated = false, done = false;
  // host results
  std::vector<int32_t> h_prim;
  std::vector<float> h_t;
  std::vector<uint8_t> h_occ;
};

size_t request_bytes(int64_t n, bool shadows) {
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t b = 7 * up(4 * (size_t)n) + 4 * up(4 * (size_t)n);  // primary rays + hits
  if (shadows) b += 7 * up(4 * (size_t)n) + 2 * up((size_t)n) + up(4 * (size_t)n);
  return b;
}

}  // namespace

struct mh_server {
  std::vector<const mh_scene*> scenes;
  mh_server_opts opts{};
  mh_launch_config launch{};
  DevBuf arena_mem;
  std::unique_ptr<Arena> arena;
  // per scene: stream, batch table (device + pinned host), counters, spill
  struct PerScene {
    hipStream_t stream = nullptr;
    DevBuf table, counters, spill;
    void* host_table = nullptr;
    int64_t spill_stride = 0;
    int blocks = 0;
  };
  std::vector<std::unique_ptr<PerScene>> per;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::vector<std::shared_ptr<Request>> active;
  std::map<uint64_t, std::shared_ptr<Request>> finished;
  uint64_t next_id = 1;
  bool stop = false;
  mh_server_stats st{};
  std::thread worker;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  ~mh_server() {
    for (auto& p : per) {
      if (p->stream) (void)hipStreamDestroy(p->stream);
      if (p->host_table) (void)hipHostFree(p->host_table);
    }
  }
  void run();
  int round();
  int admit_locked(std::vector<std::shared_ptr<Request>>& fresh);
};

namespace {

constexpr int kTableCap = MH_MAX_BATCHES;

size_t table_bytes() { return offsetof(mh::BatchTable, batches) + sizeof(mh::BatchDev) * kTableCap; }

mh::BatchDev chunk(const mh_rays& r, const mh_hits& h, int64_t first, int64_t count, bool any) {
  auto off = [&](const float* p) { return p ? p + first : nullptr; };
  mh::BatchDev d{};
  d.rays = mh::DevRays{off(r.ox), off(r.oy), off(r.oz), off(r.dx), off(r.dy), off(r.dz),
                       off(r.tmax), r.masked ? r.masked + first : nullptr};
  auto offw = [&](float* p) { return p ? p + first : nullptr; };
  d.hits = mh::DevHits{h.prim ? h.prim + first : nullptr, offw(h.t), offw(h.u), offw(h.v),
                       h.occluded ? h.occluded + first : nullptr};
  d.count = count;
  d.flags = any ? 1u : 0u;
  return d;
}

}  // namespace

int mh_server::admit_locked(std::vector<std::shared_ptr<Request>>& fresh) {
  while (!pending.empty() && (int)active.size() < opts.max_active) {
    std::shared_ptr<Request> r = pending.front();
    const size_t need = request_bytes(r->n, opts.shadows != 0);
    size_t off;
    if (!arena->alloc(need, off)) break;  // wait for memory to free up (arrival order)
    pending.pop_front();
    r->off = off;
    r->bytes = need;
    r->t_admit = now();
    char* base = arena_mem.as<char>() + off;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t f = up(4 * (size_t)r->n), b1 = up((size_t)r->n);
    float* a[7];
    for (int i = 0; i < 7; ++i) a[i] = reinterpret_cast<float*>(base + i * f);
    r->prim = mh_rays{a[0], a[1], a[2], a[3], a[4], a[5], a[6], nullptr};
    char* hb = base + 7 * f;
    r->hits = mh_hits{reinterpret_cast<int32_t*>(hb), reinterpret_cast<float*>(hb + f),
                      reinterpret_cast<float*>(hb + 2 * f), reinterpret_cast<float*>(hb + 3 * f),
                      nullptr};
    if (opts.shadows) {
      char* sb = hb + 4 * f;
      float* s[7];
      for (int i = 0; i < 7; ++i) s[i] = reinterpret_cast<float*>(sb + i * f);
      uint8_t* smask = reinterpret_cast<uint8_t*>(sb + 7 * f);
      r->occluded = reinterpret_cast<uint8_t*>(sb + 7 * f + b1);
      r->shadow_prim = reinterpret_cast<int32_t*>(sb + 7 * f + 2 * b1);
      r->shadow = mh_rays{s[0], s[1], s[2], s[3], s[4], s[5], s[6], smask};
    }
    active.push_back(r);
    fresh.push_back(r);
  }
  return MH_OK;
}

int mh_server::round() {
  std::vector<std::shared_ptr<Request>> fresh, work;
  {
    std::lock_guard<std::mutex> lk(mu);
    admit_locked(fresh);
    work = active;
  }
  // Camera rays of newly admitted requests, on their scene's stream.
  for (auto& r : fresh) {
    mh::CameraDev cd;
    int rc = mh_internal::camera_setup(r->cam, cd);
    if (rc) {
      r->status = rc;
      continue;
    }
    mh::RayOut o{const_cast<float*>(r->prim.ox), const_cast<float*>(r->prim.oy),
                 const_cast<float*>(r->prim.oz), const_cast<float*>(r->prim.dx),
                 const_cast<float*>(r->prim.dy), const_cast<float*>(r->prim.dz),
                 const_cast<float*>(r->prim.tmax), nullptr};
    HIP_OK((hipError_t)mh::launch_camera(cd, o, per[r->scene]->stream));
  }
  // One batch table per scene: a chunk of each request's remaining rays.
  const int ns = (int)scenes.size();
  std::vector<int> nb(ns, 0);
  std::vector<int64_t> total(ns, 0);
  std::vector<std::vector<std::pair<Request*, int64_t>>> plan(ns);  // (request, rays scheduled)
  for (auto& r : work) {
    if (r->status != MH_OK || r->done) continue;
    const int k = r->scene;
    mh::BatchTable* t = reinterpret_cast<mh::BatchTable*>(per[k]->host_table);
    if (nb[k] >= kTableCap) continue;
    if (!r->primary_done) {
      const int64_t c = std::min<int64_t>(opts.chunk_rays, r->n - r->next);
      mh::BatchDev d = chunk(r->prim, r->hits, r->next, c, false);
      d.start = total[k];
      t->batches[nb[k]++] = d;
      total[k] += c;
      r->next += c;
    } else if (opts.shadows && r->shadow_generated) {
      const int64_t c = std::min<int64_t>(opts.chunk_rays, r->n - r->sh_next);
      mh_hits sh{r->shadow_prim, nullptr, nullptr, nullptr, r->occluded};
      mh::BatchDev d = chunk(r->shadow, sh, r->sh_next, c, true);
      d.start = total[k];
      t->batches[nb[k]++] = d;
      total[k] += c;
      r->sh_next += c;
    }
  }
  for (int k = 0; k < ns; ++k) {
    if (nb[k] == 0) continue;
    mh::BatchTable* t = reinterpret_cast<mh::BatchTable*>(per[k]->host_table);
    t->nb = nb[k];
    t->total = total[k];
    const size_t bytes = offsetof(mh::BatchTable, batches) + sizeof(mh::BatchDev) * nb[k];
    HIP_OK(hipMemcpyAsync(per[k]->table.p, t, bytes, hipMemcpyHostToDevice, per[k]->stream));
    mh::BatchLaunch bl{};
    bl.table = per[k]->table.as<const mh::BatchTable>();
    bl.nb = nb[k];
    bl.counters = per[k]->counters.as<unsigned long long>();
    bl.spill = per[k]->spill.as<int32_t>();
    bl.spill_stride = per[k]->spill_stride;
    bl.blocks = per[k]->blocks;
    bl.block = launch.block;
    bl.stack_lds = launch.stack_lds;
    bl.layout = launch.layout;
    bl.refill = launch.refill;
    bl.chunk = launch.chunk;
    bl.xcd_remap = launch.xcd_remap;
    HIP_OK((hipError_t)mh::launch_trace_batch(scenes[k]->dev(), bl, per[k]->stream));
  }
  for (int k = 0; k < ns; ++k) HIP_OK(hipStreamSynchronize(per[k]->stream));
  int64_t traced = 0;
  for (int k = 0; k < ns; ++k) traced += total[k];
  // Progress: primary done -> shadow generation; everything done -> results to host.
  std::vector<std::shared_ptr<Request>> completed;
  for (auto& r : work) {
    if (r->done) continue;
    if (r->status != MH_OK) {
      completed.push_back(r);
      continue;
    }
    if (!r->primary_done && r->next >= r->n) {
      r->primary_done = true;
      if (opts.shadows) {
        mh::SecondaryArgs a{};
        a.in = mh::DevRays{r->prim.ox, r->prim.oy, r->prim.oz, r->prim.dx, r->prim.dy, r->prim.dz,
                           r->prim.tmax, nullptr};
        a.hit_prim = r->hits.prim;
        a.hit_t = r->hits.t;
        a.verts = scenes[r->scene]->d_verts.as<const float>();
        a.idx = scenes[r->scene]->d_idx.as<const uint32_t>();
        for (int i = 0; i < 3; ++i) a.light[i] = opts.light[i];
        a.n = r->n;
        a.out = mh::RayOut{const_cast<float*>(r->shadow.ox), const_cast<float*>(r->shadow.oy),
                           const_cast<float*>(r->shadow.oz), const_cast<float*>(r->shadow.dx),
                           const_cast<float*>(r->shadow.dy), const_cast<float*>(r->shadow.dz),
                           const_cast<float*>(r->shadow.tmax),
                           const_cast<uint8_t*>(r->shadow.masked)};
        HIP_OK((hipError_t)mh::launch_secondary(a, 0, per[r->scene]->stream));
        r->shadow_generated = true;
      }
    }
    const bool all = r->primary_done && (!opts.shadows || (r->shadow_generated && r->sh_next >= r->n));
    if (all) completed.push_back(r);
  }
  for (auto& r : completed) {
    if (r->status == MH_OK) {
      hipStream_t s = per[r->scene]->stream;
      r->h_prim.resize(r->n);
      r->h_t.resize(r->n);
      HIP_OK(hipMemcpyAsync(r->h_prim.data(), r->hits.prim, 4 * (size_t)r->n,
                            hipMemcpyDeviceToHost, s));
      HIP_OK(hipMemcpyAsync(r->h_t.data(), r->hits.t, 4 * (size_t)r->n, hipMemcpyDeviceToHost, s));
      if (opts.shadows) {
        r->h_occ.resize(r->n);
        HIP_OK(hipMemcpyAsync(r->h_occ.data(), r->occluded, (size_t)r->n, hipMemcpyDeviceToHost, s));
      }
    }
  }
  for (int k = 0; k < ns; ++k) HIP_OK(hipStreamSynchronize(per[k]->stream));
  {
    std::lock_guard<std::mutex> lk(mu);
    st.rays += traced;
    for (auto& r : completed) {
      r->t_done = now();
      r->done = true;
      arena->release(r->off, r->bytes);
      active.erase(std::find(active.begin(), active.end(), r));
      finished[r->id] = r;
      st.completed++;
    }
  }
  if (!completed.empty()) cv_done.notify_all();
  return MH_OK;
}

void mh_server::run() {
  (void)hipSetDevice(0);
  (void)ensure_init();
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
      if (stop && active.empty() && pending.empty()) return;  // drain before exiting
    }
    const double t_start = now();
    const int rc = round();
    std::lock_guard<std::mutex> lk(mu);
    st.busy_ms += now() - t_start;
    if (rc != MH_OK) {
      // A device error ends every request in flight with that status.
      for (auto& r : active) {
        r->status = rc;
        r->done = true;
        r->t_done = now();
        arena->release(r->off, r->bytes);
        finished[r->id] = r;
      }
      active.clear();
      cv_done.notify_all();
    }
  }
}

extern "C" {

int mh_server_create(const mh_scene* const* scenes, int n_scenes, const mh_server_opts* o,
                     mh_server** out) {
  if (!scenes || n_scenes < 1 || !out) return fail(MH_E_INVALID, "no scenes");
  *out = nullptr;
  for (int i = 0; i < n_scenes; ++i)
    if (!scenes[i] || !scenes[i]->built) return fail(MH_E_STATE, "scene %d has no tree", i);
  int rc = ensure_init();
  if (rc) return rc;
  std::unique_ptr<mh_server> s(new (std::nothrow) mh_server());
  if (!s) return fail(MH_E_NOMEM, "server");
  s->scenes.assign(scenes, scenes + n_scenes);
  mh_server_opts d{};
  d.hbm_budget = size_t(1) << 30;
  d.chunk_rays = int64_t(1) << 20;
  d.max_active = 64;
  d.shadows = 0;
  if (o) {
    d = *o;
    if (!d.hbm_budget) d.hbm_budget = size_t(1) << 30;
    if (d.chunk_rays <= 0) d.chunk_rays = int64_t(1) << 20;
    if (d.max_active <= 0) d.max_active = 64;
  }
  s->opts = d;
  s->launch = (d.launch.block == 0) ? mh_launch_config{MH_LAYOUT_PAIR, 8, 256, 8, 64, 256, 1}
                                    : d.launch;
  if (s->opts.max_active > kTableCap) s->opts.max_active = kTableCap;
  if ((rc = mh_internal::check_launch_config(s->launch))) return rc;
  if ((rc = s->arena_mem.reserve(d.hbm_budget))) return rc;
  s->arena.reset(new Arena(d.hbm_budget));
  for (int k = 0; k < n_scenes; ++k) s->per.emplace_back(new mh_server::PerScene());
  for (int k = 0; k < n_scenes; ++k) {
    auto& p = *s->per[k];
    HIP_OK(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
    if ((rc = p.table.reserve(table_bytes())) || (rc = p.counters.reserve(16))) return rc;
    HIP_OK(hipMemset(p.counters.p, 0, 16));
    HIP_OK(hipHostMalloc(&p.host_table, table_bytes(), hipHostMallocDefault));
    p.blocks = s->launch.blocks_per_cu * mh_internal::cu_count();
    p.spill_stride = (int64_t)p.blocks * s->launch.block;
    const int64_t need = (int64_t)scenes[k]->depth > s->launch.stack_lds
                             ? (int64_t)scenes[k]->depth - s->launch.stack_lds : 0;
    if (need > 0 && (rc = p.spill.reserve((size_t)need * p.spill_stride * 8))) return rc;
  }
  mh_server* raw = s.get();
  s->worker = std::thread([raw] { raw->run(); });
  *out = s.release();
  return MH_OK;
}

int mh_server_submit(mh_server* s, const mh_request* req, uint64_t* id) {
  if (!s || !req) return fail(MH_E_INVALID, "null argument");
  if (req->scene < 0 || req->scene >= (int)s->scenes.size())
    return fail(MH_E_INVALID, "scene %d out of range", req->scene);
  mh::CameraDev cd;
  int rc = mh_internal::camera_setup(req->camera, cd);
  if (rc) return rc;
  auto r = std::make_shared<Request>();
  r->scene = req->scene;
  r->cam = req->camera;
  r->n = (int64_t)req->camera.xres * req->camera.yres * req->camera.spp;
  const size_t need = request_bytes(r->n, s->opts.shadows != 0);
  std::lock_guard<std::mutex> lk(s->mu);
  if (need > s->opts.hbm_budget) {
    s->st.rejected++;
    return fail(MH_E_UNSUPPORTED, "request needs %zu bytes, budget is %zu", need,
                s->opts.hbm_budget);
  }
  r->id = s->next_id++;
  r->t_submit = s->now();
  s->pending.push_back(r);
  s->st.submitted++;
  if (id) *id = r->id;
  s->cv_work.notify_one();
  return MH_OK;
}

int mh_server_wait(mh_server* s, uint64_t id, int32_t* prim, float* t, uint8_t* occluded,
                   mh_request_info* info) {
  if (!s) return fail(MH_E_INVALID, "null server");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto known = [&] {
      if (s->finished.count(id)) return true;
      for (auto& p : s->pending)
        if (p->id == id) return true;
      for (auto& a : s->active)
        if (a->id == id) return true;
      return false;
    };
    // Unknown ids and ids whose results were already collected are errors, not waits.
    if (id == 0 || id >= s->next_id || !known())
      return fail(MH_E_INVALID, "unknown or already collected request %llu",
                  (unsigned long long)id);
    s->cv_done.wait(lk, [&] { return s->finished.count(id) != 0; });
    r = s->finished[id];
    s->finished.erase(id);
  }
  if (info) {
    info->id = r->id;
    info->status = r->status;
    info->rays = r->n * (s->opts.shadows ? 2 : 1);
    info->submit_ms = r->t_submit;
    info->admit_ms = r->t_admit;
    info->done_ms = r->t_done;
  }
  if (r->status != MH_OK) return fail(r->status, "request %llu failed", (unsigned long long)id);
  if (prim) std::memcpy(prim, r->h_prim.data(), 4 * (size_t)r->n);
  if (t) std::memcpy(t, r->h_t.data(), 4 * (size_t)r->n);
  if (occluded && !r->h_occ.empty()) std::memcpy(occluded, r->h_occ.data(), (size_t)r->n);
  return MH_OK;
}

int mh_server_get_stats(mh_server* s, mh_server_stats* out) {
  if (!s || !out) return fail(MH_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->active = (int)s->active.size();
  out->pending = (int)s->pending.size();
  out->hbm_in_use = s->arena->used();
  return MH_OK;
}

double mh_server_clock_ms(const mh_server* s) { return s ? s->now() : 0.0; }

int mh_server_destroy(mh_server* s) {
  if (!s) return MH_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)ensure_init();
  delete s;
  return MH_OK;
}

}  // extern "C"
