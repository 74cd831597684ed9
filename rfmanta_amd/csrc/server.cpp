// Serving loop (mh_server_*, include/mh.h): resident scenes, HBM-budget admission and
// continuous batching over the persistent batch kernel.
//
// Threads: callers submit, wait and cancel; one scheduler thread owns the device work and
// never waits for a launch it could overlap. Each step of the scheduler
//   1. admits queued requests in arrival order while their device buffers fit the device
//      arena (the HBM budget) and fewer than max_active are in flight, and queues their
//      camera rays on their scene's stream;
//   2. for every scene with fewer than two launches queued, fills the free one of the
//      scene's two batch tables with one chunk of each remaining ray set of its admitted
//      requests (primary closest-hit first; once every primary chunk of a request is
//      queued, its shadow (any-hit) and bounce (closest-hit) rays are generated on the
//      device by a kernel queued behind them, and their chunks join later tables) and
//      queues the table upload and one persistent trace; scenes have their own streams, so
//      their launches overlap, and the next table is built while a launch runs;
//   3. for a request whose primary rays are all queued and whose secondary rays follow,
//      reserves a range of the pinned host arena and queues the primary results' copies on
//      a copy stream behind an event of its last primary launch, so they overlap the
//      secondary launches; for a request whose rays are all queued, queues the remaining
//      copies behind an event of its last launch, so they overlap the next launches;
//   4. retires launches and copies whose events have completed; a request is finished, and
//      its device range freed, when its copies are done.
// mh_server_wait copies a finished request's results out of the host arena and frees that
// range. A request joins the device work with the next table of its scene after it fits.

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
  bool handed_all = false;       // every chunk of every ray set is in a queued launch
  bool cancelled = false;        // mh_server_cancel: no further chunks are handed out
  hipEvent_t retire = nullptr;   // completes when the request's copies (or work) are done
  hipEvent_t after = nullptr;    // its last launch, which the copies wait for
  hipEvent_t prim_after = nullptr;  // its last primary launch (requests with secondary rays)
  bool prim_copied = false;         // primary results already queued for copying
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
  // per scene: stream, two batch tables (device + pinned host) used in turn, counters,
  // spill, and the launches queued on the stream (with the table each one reads)
  struct PerScene {
    hipStream_t stream = nullptr;
    DevBuf table[2], counters, spill;
    void* host_table[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};  // completion of the launch reading table k
    std::deque<int> queued;                   // tables of launches not yet seen complete
    int64_t spill_stride = 0;
    int blocks = 0;
  };
  std::vector<std::unique_ptr<PerScene>> per;
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> free_events;

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
      for (int k = 0; k < 2; ++k) {
        if (p->host_table[k]) (void)hipHostFree(p->host_table[k]);
        if (p->done[k]) (void)hipEventDestroy(p->done[k]);
      }
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    if (host_mem) (void)hipHostFree(host_mem);
  }
  bool progressed = false;  // the last step queued or retired something
  double busy_since = -1;   // server clock when requests were last seen in flight
  void run();
  int step();
  int queue_scene(int k, const std::vector<std::shared_ptr<Request>>& work, int64_t& traced);
  int event(hipEvent_t& e);
  void admit_locked(std::vector<std::shared_ptr<Request>>& fresh);
  bool gpu_busy() const;
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

// An event from the pool (created on first use; timing disabled: only ordering).
int mh_server::event(hipEvent_t& e) {
  if (!free_events.empty()) {
    e = free_events.back();
    free_events.pop_back();
    return MH_OK;
  }
  HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return MH_OK;
}

bool mh_server::gpu_busy() const {
  for (auto& p : per)
    if (!p->queued.empty()) return true;
  for (auto& r : active)
    if (r->retire) return true;
  return false;
}

// Fills scene k's free table with a chunk of each remaining ray set of its requests and
// queues the upload and the launch, then the secondary-ray generation of requests whose
// primary rays are now all queued.
int mh_server::queue_scene(int k, const std::vector<std::shared_ptr<Request>>& work,
                           int64_t& traced) {
  PerScene& P = *per[k];
  const int slot = P.queued.empty() ? 0 : 1 - P.queued.back();
  mh::BatchTable* t = reinterpret_cast<mh::BatchTable*>(P.host_table[slot]);
  int nb = 0, n_any = 0;
  int64_t total = 0;
  auto add = [&](Request& r, RaySet& s) {
    if (s.next >= r.n || nb >= kTableCap) return;
    const int64_t c = std::min<int64_t>(opts.chunk_rays, r.n - s.next);
    mh::BatchDev d = chunk(s, c);
    d.start = total;
    t->batches[nb++] = d;
    n_any += s.any ? 1 : 0;
    total += c;
    s.next += c;
  };
  for (auto& r : work) {
    if (r->scene != k || r->status != MH_OK || r->cancelled || r->handed_all) continue;
    if (r->prim.next < r->n)
      add(*r, r->prim);
    else if (r->sec_generated)
      for (auto& x : r->sec) add(*r, x);
  }
  if (nb > 0) {
    t->nb = nb;
    t->total = total;
    const size_t bytes = offsetof(mh::BatchTable, batches) + sizeof(mh::BatchDev) * nb;
    HIP_OK(hipMemcpyAsync(P.table[slot].p, t, bytes, hipMemcpyHostToDevice, P.stream));
    // A rotating update between requests can deepen the tree: grow the spill area first
    // (the stream's earlier launches may still use the old one).
    const int64_t need = std::max<int64_t>(
        0, (int64_t)scenes[k]->stack_bound(launch.layout) - launch.stack_lds);
    const size_t spill_bytes = (size_t)need * P.spill_stride * 8;
    if (spill_bytes > P.spill.bytes) {
      HIP_OK(hipStreamSynchronize(P.stream));
      if (int rc = P.spill.reserve(spill_bytes)) return rc;
    }
    if (launch.layout == MH_LAYOUT_SOA)
      if (int rc = mh_internal::ensure_soa(const_cast<mh_scene*>(scenes[k]), P.stream)) return rc;
    mh::BatchLaunch bl{};
    bl.table = P.table[slot].as<const mh::BatchTable>();
    bl.nb = nb;
    bl.counters = P.counters.as<unsigned long long>();
    bl.spill = P.spill.as<int32_t>();
    bl.spill_stride = P.spill_stride;
    bl.blocks = P.blocks;
    bl.block = launch.block;
    bl.stack_lds = launch.stack_lds;
    bl.layout = launch.layout;
    bl.refill = launch.refill;
    bl.chunk = launch.chunk;
    bl.xcd_remap = launch.xcd_remap;
    bl.reverse = launch.reverse;
    bl.kinds = n_any == 0 ? 1 : (n_any == nb ? 2 : 0);
    HIP_OK((hipError_t)mh::launch_trace_batch(scenes[k]->dev(), bl, P.stream));
    HIP_OK(hipEventRecord(P.done[slot], P.stream));
    P.queued.push_back(slot);
    traced += total;
  }
  // Secondary rays of requests whose primary chunks are all queued: generated on the
  // stream behind those launches, traced by later tables.
  for (auto& r : work) {
    if (r->scene != k || r->status != MH_OK || r->cancelled || r->sec_generated) continue;
    if (r->prim.next < r->n) continue;
    r->primary_done = true;
    if (!r->sec.empty()) {
      // the primary results are final once the launches queued so far have run: their
      // copies can overlap the secondary rays' launches
      if (int rc = event(r->prim_after)) return rc;
      HIP_OK(hipEventRecord(r->prim_after, P.stream));
      const mh_scene* sc = scenes[k];
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
      for (auto& x : r->sec) {
        a.out = ray_out(x.rays);
        HIP_OK((hipError_t)mh::launch_secondary(a, x.any ? 0 : 1, P.stream));
      }
    }
    r->sec_generated = true;
  }
  for (auto& r : work) {
    if (r->scene != k || r->status != MH_OK || r->cancelled || r->handed_all) continue;
    bool all = r->sec_generated;
    for (auto& x : r->sec) all = all && x.next >= r->n;
    if (all) r->handed_all = true;
  }
  return MH_OK;
}

int mh_server::step() {
  std::vector<std::shared_ptr<Request>> fresh, work;
  {
    std::lock_guard<std::mutex> lk(mu);
    admit_locked(fresh);
    work = active;
  }
  bool moved = !fresh.empty();
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
  // Launches seen complete free their tables.
  for (auto& P : per) {
    while (!P->queued.empty()) {
      const hipError_t q = hipEventQuery(P->done[P->queued.front()]);
      if (q == hipErrorNotReady) break;
      HIP_OK(q);
      P->queued.pop_front();
      moved = true;
    }
  }
  // Next launch of every scene with a free table.
  int64_t traced = 0;
  for (int k = 0; k < (int)scenes.size(); ++k)
    if (per[k]->queued.size() < 2)
      if (int rc = queue_scene(k, work, traced)) return rc;
  moved = moved || traced > 0;
  // Requests with all their rays queued: host ranges and result copies behind their
  // scene's launches; cancelled requests: an event behind the work already queued.
  // Requests whose primary launches are all queued and whose secondary rays are not: a host
  // range and the primary copies behind those launches, ahead of the rest.
  std::vector<std::shared_ptr<Request>> to_copy, prim_copy;
  {
    std::lock_guard<std::mutex> lk(mu);
    st.rays += traced;
    for (auto& r : work) {
      if (r->retire || r->done) continue;
      if (r->status != MH_OK || r->cancelled) {
        to_copy.push_back(r);
        continue;
      }
      const bool early = r->prim_after && !r->prim_copied && !r->handed_all;
      if (!r->handed_all && !early) continue;
      if (!r->has_host) {
        const size_t hb = host_bytes(r->n, opts.shadows != 0, opts.bounces != 0);
        size_t off;
        if (!host_arena->alloc(hb, off)) continue;  // waits for collections
        r->host_off = off;
        r->host_bytes = hb;
        r->has_host = true;
      }
      (early ? prim_copy : to_copy).push_back(r);
    }
  }
  for (auto& r : prim_copy) {
    HIP_OK(hipStreamWaitEvent(copy_stream, r->prim_after, 0));
    const int64_t n = r->n;
    Carver c{host_mem + r->host_off};
    r->h_prim = c.take<int32_t>(n);
    r->h_t = c.take<float>(n);
    HIP_OK(hipMemcpyAsync(r->h_prim, r->prim.hits.prim, 4 * (size_t)n, hipMemcpyDeviceToHost,
                          copy_stream));
    HIP_OK(hipMemcpyAsync(r->h_t, r->prim.hits.t, 4 * (size_t)n, hipMemcpyDeviceToHost,
                          copy_stream));
    r->prim_copied = true;
    moved = true;
  }
  for (auto& r : to_copy) {
    hipStream_t s = per[r->scene]->stream;
    hipEvent_t after;
    if (int rc = event(after)) return rc;
    HIP_OK(hipEventRecord(after, s));
    if (r->status != MH_OK || r->cancelled) {
      // its buffers are free once the queued work has passed, and its host range once the
      // primary copies already queued for it are done
      if (r->prim_copied) {
        HIP_OK(hipStreamWaitEvent(copy_stream, after, 0));
        hipEvent_t copied;
        if (int rc = event(copied)) return rc;
        HIP_OK(hipEventRecord(copied, copy_stream));
        r->after = after;
        r->retire = copied;
      } else {
        r->retire = after;
      }
      moved = true;
      continue;
    }
    HIP_OK(hipStreamWaitEvent(copy_stream, after, 0));
    r->after = after;
    const int64_t n = r->n;
    Carver c{host_mem + r->host_off};
    r->h_prim = c.take<int32_t>(n);
    r->h_t = c.take<float>(n);
    if (!r->prim_copied) {
      HIP_OK(hipMemcpyAsync(r->h_prim, r->prim.hits.prim, 4 * (size_t)n, hipMemcpyDeviceToHost,
                            copy_stream));
      HIP_OK(hipMemcpyAsync(r->h_t, r->prim.hits.t, 4 * (size_t)n, hipMemcpyDeviceToHost,
                            copy_stream));
    }
    for (auto& x : r->sec) {
      if (x.any) {
        r->h_occ = c.take<uint8_t>(n);
        HIP_OK(hipMemcpyAsync(r->h_occ, x.hits.occluded, (size_t)n, hipMemcpyDeviceToHost,
                              copy_stream));
      } else {
        r->h_bprim = c.take<int32_t>(n);
        r->h_bt = c.take<float>(n);
        HIP_OK(hipMemcpyAsync(r->h_bprim, x.hits.prim, 4 * (size_t)n, hipMemcpyDeviceToHost,
                              copy_stream));
        HIP_OK(hipMemcpyAsync(r->h_bt, x.hits.t, 4 * (size_t)n, hipMemcpyDeviceToHost,
                              copy_stream));
      }
    }
    hipEvent_t copied;
    if (int rc = event(copied)) return rc;
    HIP_OK(hipEventRecord(copied, copy_stream));
    r->retire = copied;
    moved = true;
  }
  // Requests whose copies (or queued work, when cancelled) are done.
  std::vector<std::shared_ptr<Request>> retired;
  for (auto& r : work) {
    if (!r->retire || r->done) continue;
    const hipError_t q = hipEventQuery(r->retire);
    if (q == hipErrorNotReady) continue;
    HIP_OK(q);
    free_events.push_back(r->retire);
    r->retire = nullptr;
    if (r->after) free_events.push_back(r->after);
    r->after = nullptr;
    if (r->prim_after) free_events.push_back(r->prim_after);
    r->prim_after = nullptr;
    retired.push_back(r);
  }
  if (!retired.empty()) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& r : retired) {
      if (r->cancelled && r->status == MH_OK) r->status = MH_E_CANCELED;
      if (r->status != MH_OK && r->has_host) {
        host_arena->release(r->host_off, r->host_bytes);
        r->has_host = false;
      }
      r->t_done = now();
      r->done = true;
      arena->release(r->dev_off, r->dev_bytes);
      active.erase(std::find(active.begin(), active.end(), r));
      finished[r->id] = r;
      if (r->status == MH_OK) st.completed++;
    }
    moved = true;
    cv_done.notify_all();
  }
  progressed = moved;
  return MH_OK;
}

void mh_server::run() {
  (void)ensure_init();
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
      if (stop && active.empty() && pending.empty()) return;  // drain before exiting
      if (busy_since < 0) busy_since = now();
    }
    const int rc = step();
    std::unique_lock<std::mutex> lk(mu);
    if (rc != MH_OK) {
      // A device error ends every request in flight with that status (the device work
      // queued for them is not waited for: after an error the streams are unusable).
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
      // Nothing to queue or retire: poll the device again shortly while it has work,
      // else wait for a new request or a collection that frees host space.
      if (gpu_busy()) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
      } else {
        cv_work.wait_for(lk, std::chrono::milliseconds(1));
      }
    }
    if (active.empty() && pending.empty() && busy_since >= 0) {
      st.busy_ms += now() - busy_since;
      busy_since = -1;
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
  // Default: the static stride, 16 workgroups of 256 lanes per CU. The waves in flight
  // then cover a contiguous part of each batch, which keeps the records they read in the
  // caches; counter chunks of 256 rays spread the first chunks of all waves over a whole
  // 1080p batch, and on an 8M-triangle scene (records larger than the Infinity Cache)
  // traced its primary rays 2.8 times slower.
  s->launch = (d.launch.block == 0) ? mh_launch_config{MH_LAYOUT_PAIR, 8, 256, 16, 64, 0, 0}
                                    : d.launch;
  if ((rc = mh_internal::check_launch_config(s->launch))) return rc;
  if ((rc = s->arena_mem.reserve(d.hbm_budget))) return rc;
  s->arena.reset(new Arena(d.hbm_budget));
  HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&s->host_mem), d.host_budget,
                       hipHostMallocDefault));
  s->host_arena.reset(new Arena(d.host_budget));
  for (int k = 0; k < n_scenes; ++k) s->per.emplace_back(new mh_server::PerScene());
  HIP_OK(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
  for (int k = 0; k < n_scenes; ++k) {
    auto& p = *s->per[k];
    HIP_OK(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
    if ((rc = p.table[0].reserve(table_bytes())) || (rc = p.table[1].reserve(table_bytes())) ||
        (rc = p.counters.reserve(mh::kBatchCounterBytes)))
      return rc;
    HIP_OK(hipMemset(p.counters.p, 0, mh::kBatchCounterBytes));
    for (int t = 0; t < 2; ++t) {
      HIP_OK(hipHostMalloc(&p.host_table[t], table_bytes(), hipHostMallocDefault));
      HIP_OK(hipEventCreateWithFlags(&p.done[t], hipEventDisableTiming));
    }
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

int mh_server_cancel(mh_server* s, uint64_t id) {
  if (!s) return fail(MH_E_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto it = s->pending.begin(); it != s->pending.end(); ++it) {
    if ((*it)->id != id) continue;
    std::shared_ptr<Request> r = *it;
    s->pending.erase(it);
    r->status = MH_E_CANCELED;
    r->done = true;
    r->t_done = s->now();
    s->finished[id] = r;
    s->cv_done.notify_all();
    return MH_OK;
  }
  for (auto& r : s->active) {
    if (r->id != id) continue;
    r->cancelled = true;  // the scheduler stops its chunks and frees it behind its work
    s->cv_work.notify_one();
    return MH_OK;
  }
  if (s->finished.count(id))
    return fail(MH_E_STATE, "request %llu has already finished", (unsigned long long)id);
  return fail(MH_E_INVALID, "unknown or already collected request %llu", (unsigned long long)id);
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
