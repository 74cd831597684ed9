// Serving loop (mh_server_*, include/mh.h): resident scenes, HBM-budget admission and
// continuous batching over the persistent batch kernel.
//
// Threads: callers submit and wait; one scheduler thread owns the device work. Each
// scheduler round
//   1. admits queued requests in arrival order while their device buffers fit the device
//      arena (the HBM budget) and fewer than max_active are in flight, generating their
//      camera rays on the device;
//   2. builds, per scene, a batch table with one chunk of each remaining ray set of every
//      admitted request (primary closest-hit first; then its shadow (any-hit) and bounce
//      (closest-hit) rays, generated on the device from the primary hits) and launches
//      one persistent trace per scene, each scene on its own stream so that the launches
//      overlap;
//   3. waits for the launches; a request whose ray sets are all traced gets a range of the
//      pinned host arena, its results are copied there and its device range is freed.
// mh_server_wait copies a finished request's results out of the host arena and frees
// that range. A request joins the device work at the first round after it fits, so large
// requests never hold small ones back for more than one round.

#include <hip/hip_runtime_api.h>

#include <algorithm>
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

constexpr size_t kAlign = 256;
size_t up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit range allocator over [0, size) with a coalescing free list.
class Arena {
 public:
  explicit Arena(size_t size) {
    if (size) free_[0] = size;
  }
  bool alloc(size_t n, size_t& off) {
    n = up(n);
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
    n = up(n);
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

// Hands out 256-byte aligned sub-ranges of one allocation.
struct Carver {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(int64_t n) {
    T* p = reinterpret_cast<T*>(base + off);
    off += up(sizeof(T) * (size_t)n);
    return p;
  }
};

// One set of rays of a request and where its hits go.
struct RaySet {
  mh_rays rays{};
  mh_hits hits{};
  int64_t next = 0;  // rays handed to launches so far
  bool any = false;
};

struct Request {
  uint64_t id = 0;
  int scene = 0;
  mh_camera cam{};
  int64_t n = 0;  // primary rays
  int status = MH_OK;
  double t_submit = 0, t_admit = 0, t_done = 0;
  size_t dev_off = 0, dev_bytes = 0;    // device arena range
  size_t host_off = 0, host_bytes = 0;  // host arena range (results)
  bool has_host = false;
  RaySet prim;
  std::vector<RaySet> sec;  // shadow and/or bounce
  bool primary_done = false, sec_generated = false, done = false;
  // results in the host arena
  int32_t* h_prim = nullptr;
  float* h_t = nullptr;
  uint8_t* h_occ = nullptr;
  int32_t* h_bprim = nullptr;
  float* h_bt = nullptr;
};

size_t device_bytes(int64_t n, bool shadows, bool bounces) {
  const size_t f = up(4 * (size_t)n), b = up((size_t)n);
  size_t s = 7 * f + 4 * f;             // primary rays, hits (prim t u v)
  if (shadows) s += 7 * f + 2 * b;      // shadow rays + mask, occluded
  if (bounces) s += 7 * f + b + 2 * f;  // bounce rays + mask, hits (prim t)
  return s;
}

size_t host_bytes(int64_t n, bool shadows, bool bounces) {
  const size_t f = up(4 * (size_t)n), b = up((size_t)n);
  return 2 * f + (shadows ? b : 0) + (bounces ? 2 * f : 0);
}

}  // namespace

struct mh_server {
  std::vector<const mh_scene*> scenes;
  mh_server_opts opts{};
  mh_launch_config launch{};
  DevBuf arena_mem;
  std::unique_ptr<Arena> arena;
  char* host_mem = nullptr;
  std::unique_ptr<Arena> host_arena;
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
    if (host_mem) (void)hipHostFree(host_mem);
  }
  bool progressed = false;  // the last round launched or finished something
  void run();
  int round();
  void admit_locked(std::vector<std::shared_ptr<Request>>& fresh);
};

namespace {

constexpr int kTableCap = MH_MAX_BATCHES;

size_t table_bytes() { return offsetof(mh::BatchTable, batches) + sizeof(mh::BatchDev) * kTableCap; }

mh::BatchDev chunk(const RaySet& s, int64_t count) {
  const int64_t first = s.next;
  auto rd = [&](const float* p) { return p ? p + first : nullptr; };
  auto wr = [&](float* p) { return p ? p + first : nullptr; };
  const mh_rays& r = s.rays;
  const mh_hits& h = s.hits;
  mh::BatchDev d{};
  d.rays = mh::DevRays{rd(r.ox), rd(r.oy), rd(r.oz), rd(r.dx), rd(r.dy), rd(r.dz), rd(r.tmax),
                       r.masked ? r.masked + first : nullptr};
  d.hits = mh::DevHits{h.prim ? h.prim + first : nullptr, wr(h.t), wr(h.u), wr(h.v),
                       h.occluded ? h.occluded + first : nullptr};
  d.count = count;
  d.flags = s.any ? 1u : 0u;
  return d;
}

mh::RayOut ray_out(const mh_rays& r) {
  return mh::RayOut{const_cast<float*>(r.ox), const_cast<float*>(r.oy), const_cast<float*>(r.oz),
                    const_cast<float*>(r.dx), const_cast<float*>(r.dy), const_cast<float*>(r.dz),
                    const_cast<float*>(r.tmax), const_cast<uint8_t*>(r.masked)};
}

}  // namespace

void mh_server::admit_locked(std::vector<std::shared_ptr<Request>>& fresh) {
  const bool sh = opts.shadows != 0, bo = opts.bounces != 0;
  while (!pending.empty() && (int)active.size() < opts.max_active) {
    std::shared_ptr<Request> r = pending.front();
    const size_t need = device_bytes(r->n, sh, bo);
    size_t off;
    if (!arena->alloc(need, off)) break;  // wait for memory to free up (arrival order)
    pending.pop_front();
    r->dev_off = off;
    r->dev_bytes = need;
    r->t_admit = now();
    Carver c{arena_mem.as<char>() + off};
    const int64_t n = r->n;
    auto rays = [&](bool masked) {
      mh_rays x{};
      x.ox = c.take<float>(n);
      x.oy = c.take<float>(n);
      x.oz = c.take<float>(n);
      x.dx = c.take<float>(n);
      x.dy = c.take<float>(n);
      x.dz = c.take<float>(n);
      x.tmax = c.take<float>(n);
      x.masked = masked ? c.take<uint8_t>(n) : nullptr;
      return x;
    };
    r->prim.rays = rays(false);
    r->prim.hits.prim = c.take<int32_t>(n);
    r->prim.hits.t = c.take<float>(n);
    r->prim.hits.u = c.take<float>(n);
    r->prim.hits.v = c.take<float>(n);
    if (sh) {
      RaySet s;
      s.rays = rays(true);
      s.hits.occluded = c.take<uint8_t>(n);
      s.any = true;
      r->sec.push_back(s);
    }
    if (bo) {
      RaySet b;
      b.rays = rays(true);
      b.hits.prim = c.take<int32_t>(n);
      b.hits.t = c.take<float>(n);
      r->sec.push_back(b);
    }
    active.push_back(r);
    fresh.push_back(r);
  }
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
    HIP_OK((hipError_t)mh::launch_camera(cd, ray_out(r->prim.rays), per[r->scene]->stream));
  }
  // One batch table per scene: a chunk of each remaining ray set of each request.
  const int ns = (int)scenes.size();
  std::vector<int> nb(ns, 0), n_any(ns, 0);
  std::vector<int64_t> total(ns, 0);
  auto add = [&](Request& r, RaySet& s) {
    const int k = r.scene;
    if (s.next >= r.n || nb[k] >= kTableCap) return;
    const int64_t c = std::min<int64_t>(opts.chunk_rays, r.n - s.next);
    mh::BatchTable* t = reinterpret_cast<mh::BatchTable*>(per[k]->host_table);
    mh::BatchDev d = chunk(s, c);
    d.start = total[k];
    t->batches[nb[k]++] = d;
    n_any[k] += s.any ? 1 : 0;
    total[k] += c;
    s.next += c;
  };
  for (auto& r : work) {
    if (r->status != MH_OK || r->done) continue;
    if (!r->primary_done)
      add(*r, r->prim);
    else if (r->sec_generated)
      for (auto& s : r->sec) add(*r, s);
  }
  for (int k = 0; k < ns; ++k) {
    if (nb[k] == 0) continue;
    mh::BatchTable* t = reinterpret_cast<mh::BatchTable*>(per[k]->host_table);
    t->nb = nb[k];
    t->total = total[k];
    const size_t bytes = offsetof(mh::BatchTable, batches) + sizeof(mh::BatchDev) * nb[k];
    HIP_OK(hipMemcpyAsync(per[k]->table.p, t, bytes, hipMemcpyHostToDevice, per[k]->stream));
    // A rotating update between rounds can deepen the tree: grow the spill area first.
    const int64_t need = std::max<int64_t>(
        0, (int64_t)scenes[k]->stack_bound(launch.layout) - launch.stack_lds);
    const size_t spill_bytes = (size_t)need * per[k]->spill_stride * 8;
    if (spill_bytes > per[k]->spill.bytes) {
      HIP_OK(hipStreamSynchronize(per[k]->stream));
      if (int rc = per[k]->spill.reserve(spill_bytes)) return rc;
    }
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
    bl.kinds = n_any[k] == 0 ? 1 : (n_any[k] == nb[k] ? 2 : 0);
    HIP_OK((hipError_t)mh::launch_trace_batch(scenes[k]->dev(), bl, per[k]->stream));
  }
  for (int k = 0; k < ns; ++k) HIP_OK(hipStreamSynchronize(per[k]->stream));
  int64_t traced = 0;
  for (int k = 0; k < ns; ++k) traced += total[k];
  // Progress: primary done -> secondary generation; all done -> results to the host.
  std::vector<std::shared_ptr<Request>> completed;
  for (auto& r : work) {
    if (r->done) continue;
    if (r->status != MH_OK) {
      completed.push_back(r);
      continue;
    }
    if (!r->primary_done && r->prim.next >= r->n) {
      r->primary_done = true;
      const mh_scene* sc = scenes[r->scene];
      mh::SecondaryArgs a{};
      const mh_rays& p = r->prim.rays;
      a.in = mh::DevRays{p.ox, p.oy, p.oz, p.dx, p.dy, p.dz, p.tmax, nullptr};
      a.hit_prim = r->prim.hits.prim;
      a.hit_t = r->prim.hits.t;
      a.verts = sc->d_verts.as<const float>();
      a.idx = sc->d_idx.as<const uint32_t>();
      for (int i = 0; i < 3; ++i) a.light[i] = opts.light[i];
      a.seed = r->cam.seed + 1;
      a.n = r->n;
      for (auto& s : r->sec) {
        a.out = ray_out(s.rays);
        HIP_OK((hipError_t)mh::launch_secondary(a, s.any ? 0 : 1, per[r->scene]->stream));
      }
      r->sec_generated = true;
    }
    bool all = r->primary_done;
    for (auto& s : r->sec) all = all && s.next >= r->n;
    if (all) completed.push_back(r);
  }
  // Host ranges for the finished requests (those that do not fit wait for collection).
  std::vector<std::shared_ptr<Request>> copied;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& r : completed) {
      if (r->status != MH_OK) {
        copied.push_back(r);
        continue;
      }
      const size_t hb = host_bytes(r->n, opts.shadows != 0, opts.bounces != 0);
      size_t off;
      if (!host_arena->alloc(hb, off)) continue;
      r->host_off = off;
      r->host_bytes = hb;
      r->has_host = true;
      copied.push_back(r);
    }
  }
  for (auto& r : copied) {
    if (r->status != MH_OK) continue;
    hipStream_t s = per[r->scene]->stream;
    const int64_t n = r->n;
    Carver c{host_mem + r->host_off};
    r->h_prim = c.take<int32_t>(n);
    r->h_t = c.take<float>(n);
    HIP_OK(hipMemcpyAsync(r->h_prim, r->prim.hits.prim, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(r->h_t, r->prim.hits.t, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    for (auto& x : r->sec) {
      if (x.any) {
        r->h_occ = c.take<uint8_t>(n);
        HIP_OK(hipMemcpyAsync(r->h_occ, x.hits.occluded, (size_t)n, hipMemcpyDeviceToHost, s));
      } else {
        r->h_bprim = c.take<int32_t>(n);
        r->h_bt = c.take<float>(n);
        HIP_OK(hipMemcpyAsync(r->h_bprim, x.hits.prim, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(r->h_bt, x.hits.t, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
      }
    }
  }
  for (int k = 0; k < ns; ++k) HIP_OK(hipStreamSynchronize(per[k]->stream));
  {
    std::lock_guard<std::mutex> lk(mu);
    st.rays += traced;
    for (auto& r : copied) {
      r->t_done = now();
      r->done = true;
      arena->release(r->dev_off, r->dev_bytes);
      active.erase(std::find(active.begin(), active.end(), r));
      finished[r->id] = r;
      st.completed++;
    }
  }
  progressed = !fresh.empty() || traced > 0 || !copied.empty();
  if (!copied.empty()) cv_done.notify_all();
  return MH_OK;
}

void mh_server::run() {
  (void)ensure_init();
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
      if (stop && active.empty() && pending.empty()) return;  // drain before exiting
    }
    const double t_start = now();
    const int rc = round();
    std::unique_lock<std::mutex> lk(mu);
    st.busy_ms += now() - t_start;
    if (rc != MH_OK) {
      // A device error ends every request in flight with that status.
      for (auto& r : active) {
        r->status = rc;
        r->done = true;
        r->t_done = now();
        arena->release(r->dev_off, r->dev_bytes);
        if (r->has_host) host_arena->release(r->host_off, r->host_bytes);
        r->has_host = false;
        finished[r->id] = r;
      }
      active.clear();
      cv_done.notify_all();
    } else if (!progressed) {
      // Finished requests that found no host space wait for collections
      // (mh_server_wait notifies) instead of spinning.
      cv_work.wait_for(lk, std::chrono::milliseconds(1));
    }
    if (stop) {
      // Uncollected results are discarded at shutdown so the drain cannot stall on them.
      for (auto& kv : finished) {
        Request& r = *kv.second;
        if (!r.has_host) continue;
        host_arena->release(r.host_off, r.host_bytes);
        r.has_host = false;
        r.h_prim = nullptr;
        r.h_t = nullptr;
        r.h_occ = nullptr;
        r.h_bprim = nullptr;
        r.h_bt = nullptr;
        if (r.status == MH_OK) r.status = MH_E_STATE;
      }
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
  if (o) d = *o;
  if (!d.hbm_budget) d.hbm_budget = size_t(1) << 30;
  if (!d.host_budget) d.host_budget = d.hbm_budget / 4;
  if (d.chunk_rays <= 0) d.chunk_rays = int64_t(1) << 20;
  if (d.max_active <= 0) d.max_active = 64;
  if (d.max_active > kTableCap / 2) d.max_active = kTableCap / 2;  // two table entries each
  s->opts = d;
  s->launch = (d.launch.block == 0) ? mh_launch_config{MH_LAYOUT_PAIR, 8, 256, 8, 64, 256, 0}
                                    : d.launch;
  if ((rc = mh_internal::check_launch_config(s->launch))) return rc;
  if ((rc = s->arena_mem.reserve(d.hbm_budget))) return rc;
  s->arena.reset(new Arena(d.hbm_budget));
  HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&s->host_mem), d.host_budget,
                       hipHostMallocDefault));
  s->host_arena.reset(new Arena(d.host_budget));
  for (int k = 0; k < n_scenes; ++k) s->per.emplace_back(new mh_server::PerScene());
  for (int k = 0; k < n_scenes; ++k) {
    auto& p = *s->per[k];
    HIP_OK(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
    if ((rc = p.table.reserve(table_bytes())) || (rc = p.counters.reserve(mh::kBatchCounterBytes))) return rc;
    HIP_OK(hipMemset(p.counters.p, 0, mh::kBatchCounterBytes));
    HIP_OK(hipHostMalloc(&p.host_table, table_bytes(), hipHostMallocDefault));
    p.blocks = s->launch.blocks_per_cu * mh_internal::cu_count();
    p.spill_stride = (int64_t)p.blocks * s->launch.block;
    const int64_t need = std::max<int64_t>(
        0, (int64_t)scenes[k]->stack_bound(s->launch.layout) - s->launch.stack_lds);
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
  const bool sh = s->opts.shadows != 0, bo = s->opts.bounces != 0;
  const size_t need = device_bytes(r->n, sh, bo), hneed = host_bytes(r->n, sh, bo);
  std::lock_guard<std::mutex> lk(s->mu);
  if (need > s->opts.hbm_budget || hneed > s->opts.host_budget) {
    s->st.rejected++;
    return fail(MH_E_UNSUPPORTED,
                "request needs %zu device / %zu host bytes, budgets are %zu / %zu", need, hneed,
                s->opts.hbm_budget, s->opts.host_budget);
  }
  r->id = s->next_id++;
  r->t_submit = s->now();
  s->pending.push_back(r);
  s->st.submitted++;
  if (id) *id = r->id;
  s->cv_work.notify_one();
  return MH_OK;
}

int mh_server_wait(mh_server* s, uint64_t id, const mh_request_result* out,
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
    info->rays = r->n * (1 + (int64_t)r->sec.size());
    info->submit_ms = r->t_submit;
    info->admit_ms = r->t_admit;
    info->done_ms = r->t_done;
  }
  if (r->status == MH_OK && out) {
    const size_t n = (size_t)r->n;
    if (out->prim && r->h_prim) std::memcpy(out->prim, r->h_prim, 4 * n);
    if (out->t && r->h_t) std::memcpy(out->t, r->h_t, 4 * n);
    if (out->occluded && r->h_occ) std::memcpy(out->occluded, r->h_occ, n);
    if (out->bounce_prim && r->h_bprim) std::memcpy(out->bounce_prim, r->h_bprim, 4 * n);
    if (out->bounce_t && r->h_bt) std::memcpy(out->bounce_t, r->h_bt, 4 * n);
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->has_host) s->host_arena->release(r->host_off, r->host_bytes);
    r->has_host = false;
  }
  s->cv_work.notify_one();
  if (r->status != MH_OK) return fail(r->status, "request %llu failed", (unsigned long long)id);
  return MH_OK;
}

int mh_server_get_stats(mh_server* s, mh_server_stats* out) {
  if (!s || !out) return fail(MH_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->active = (int)s->active.size();
  out->pending = (int)s->pending.size();
  out->hbm_in_use = s->arena->used();
  out->host_in_use = s->host_arena->used();
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
