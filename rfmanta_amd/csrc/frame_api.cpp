// C ABI of persistent multi-batch frames (mh_frame_*, include/mh.h): the device batch
// table, launch configuration, HIP-graph capture and replay, timing and the autotuner
// with its JSON cache.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "internal.h"
#include "mh.h"
#include "mh_device.h"

using mh_internal::DevBuf;
using mh_internal::ensure_init;
using mh_internal::fail;

struct mh_frame {
  const mh_scene* scene = nullptr;
  int nb = 0;
  int64_t total = 0;
  int n_any = 0;
  DevBuf table, counters, spill;
  mh_launch_config cfg{};
  int blocks = 0;
  int64_t spill_stride = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // The scene's device view as the captured launch holds it (kernel arguments are
  // frozen by value): buffer addresses and bounds, root reference, triangle test.
  mh::DevScene captured{};
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~mh_frame() { drop_graph(); }
};

namespace mh_internal {

int check_launch_config(const mh_launch_config& c) {
  if (c.layout != MH_LAYOUT_PAIR && c.layout != MH_LAYOUT_REF && c.layout != MH_LAYOUT_SOA)
    return fail(MH_E_INVALID, "layout %d", c.layout);
  if (c.stack_lds != 8 && c.stack_lds != 16 && c.stack_lds != 32)
    return fail(MH_E_INVALID, "stack_lds %d not in {8,16,32}", c.stack_lds);
  if (c.block != 64 && c.block != 128 && c.block != 256)
    return fail(MH_E_INVALID, "block %d not in {64,128,256}", c.block);
  if (c.blocks_per_cu < 1 || c.blocks_per_cu > 64)
    return fail(MH_E_INVALID, "blocks_per_cu %d", c.blocks_per_cu);
  if (c.refill < 1 || c.refill > 64) return fail(MH_E_INVALID, "refill %d not in 1..64", c.refill);
  if (c.chunk < 0 || c.chunk > (1 << 20) || c.chunk % 64 != 0)
    return fail(MH_E_INVALID, "chunk %d is not a multiple of 64 in 0..2^20", c.chunk);
  if (c.chunk == 0 && c.refill < 64)
    return fail(MH_E_INVALID, "chunk 0 (static stride) needs refill 64");
  if (c.xcd_remap != 0 && c.xcd_remap != 1) return fail(MH_E_INVALID, "xcd_remap %d", c.xcd_remap);
  if (c.reverse != 0 && c.reverse != 1) return fail(MH_E_INVALID, "reverse %d", c.reverse);
  const size_t lds = (size_t)c.stack_lds * c.block * 8 + 256 * 8;
  if (lds > 65536) return fail(MH_E_UNSUPPORTED, "LDS per workgroup %zu > 64 KiB", lds);
  return MH_OK;
}

}  // namespace mh_internal

namespace {

const mh_launch_config kDefault = {MH_LAYOUT_PAIR, 8, 256, 16, 64, 256, 1, 0};


// Kernel specialisation for a table: closest-hit only, any-hit only, or mixed.
int kinds_of(int nb, int n_any) { return n_any == 0 ? 1 : (n_any == nb ? 2 : 0); }

int build_table(const mh_batch* b, int n, std::vector<uint8_t>& out, int64_t& total, int& n_any) {
  if (!b || n < 1 || n > MH_MAX_BATCHES)
    return fail(MH_E_INVALID, "batch count %d not in 1..%d", n, MH_MAX_BATCHES);
  const size_t bytes = offsetof(mh::BatchTable, batches) + sizeof(mh::BatchDev) * (size_t)n;
  out.assign(bytes, 0);
  mh::BatchTable* t = reinterpret_cast<mh::BatchTable*>(out.data());
  total = 0;
  n_any = 0;
  for (int i = 0; i < n; ++i) {
    const mh_batch& x = b[i];
    if (x.count < 0) return fail(MH_E_INVALID, "batch %d: negative count", i);
    if (x.flags & ~MH_ANY_HIT) return fail(MH_E_INVALID, "batch %d: flags 0x%x", i, x.flags);
    const bool any = x.flags & MH_ANY_HIT;
    if (x.count > 0) {
      const mh_rays& r = x.rays;
      if (!r.ox || !r.oy || !r.oz || !r.dx || !r.dy || !r.dz || !r.tmax)
        return fail(MH_E_INVALID, "batch %d: null ray array", i);
      if (any ? !x.hits.occluded : !x.hits.prim)
        return fail(MH_E_INVALID, "batch %d: %s", i, any ? "any-hit needs hits.occluded"
                                                         : "closest-hit needs hits.prim");
    }
    mh::BatchDev& d = t->batches[i];
    d.rays = mh::DevRays{x.rays.ox, x.rays.oy, x.rays.oz, x.rays.dx, x.rays.dy, x.rays.dz,
                         x.rays.tmax, x.rays.masked};
    d.hits = mh::DevHits{x.hits.prim, x.hits.t, x.hits.u, x.hits.v, x.hits.occluded};
    d.count = x.count;
    d.start = total;
    d.flags = x.flags;
    total += x.count;
    n_any += any ? 1 : 0;
  }
  t->total = total;
  t->nb = n;
  return MH_OK;
}

int prepare(mh_frame* f, const mh_launch_config& c) {
  int rc = mh_internal::check_launch_config(c);
  if (rc) return rc;
  const int blocks = c.blocks_per_cu * mh_internal::cu_count();
  const int64_t stride = (int64_t)blocks * c.block;
  const int64_t need = std::max<int64_t>(0, (int64_t)f->scene->stack_bound(c.layout) - c.stack_lds);
  void* old_spill = f->spill.p;
  if (need > 0) {
    if ((rc = f->spill.reserve((size_t)need * stride * 8))) return rc;
  }
  if (f->spill.p != old_spill) f->drop_graph();  // a captured launch holds the old area
  if (c.layout != f->cfg.layout || c.stack_lds != f->cfg.stack_lds || c.block != f->cfg.block ||
      c.blocks_per_cu != f->cfg.blocks_per_cu || c.refill != f->cfg.refill ||
      c.chunk != f->cfg.chunk || c.xcd_remap != f->cfg.xcd_remap || c.reverse != f->cfg.reverse)
    f->drop_graph();
  f->cfg = c;
  f->blocks = blocks;
  f->spill_stride = stride;
  return MH_OK;
}

// True when the spill area still covers the scene's deepest stack (a rotating update can
// deepen the tree after the frame was prepared).
bool spill_ok(const mh_frame* f) {
  const int64_t need =
      std::max<int64_t>(0, (int64_t)f->scene->stack_bound(f->cfg.layout) - f->cfg.stack_lds);
  return (size_t)need * (size_t)f->spill_stride * 8 <= f->spill.bytes;
}

// The SoA copy of the scene's PairNodes, current on `stream`, for SOA-layout launches.
int soa_ready(mh_frame* f, void* stream) {
  if (f->cfg.layout != MH_LAYOUT_SOA) return MH_OK;
  return mh_internal::ensure_soa(const_cast<mh_scene*>(f->scene), stream);
}

int launch(mh_frame* f, void* stream) {
  if (f->total == 0) return MH_OK;
  if (!spill_ok(f)) {
    int rc = prepare(f, f->cfg);
    if (rc) return rc;
  }
  if (int rc = soa_ready(f, stream)) return rc;
  mh::BatchLaunch bl{};
  bl.table = f->table.as<const mh::BatchTable>();
  bl.nb = f->nb;
  bl.counters = f->counters.as<unsigned long long>();
  bl.spill = f->spill.as<int32_t>();
  bl.spill_stride = f->spill_stride;
  bl.blocks = f->blocks;
  bl.block = f->cfg.block;
  bl.stack_lds = f->cfg.stack_lds;
  bl.layout = f->cfg.layout;
  bl.refill = f->cfg.refill;
  bl.chunk = f->cfg.chunk;
  bl.xcd_remap = f->cfg.xcd_remap;
  bl.reverse = f->cfg.reverse;
  bl.kinds = kinds_of(f->nb, f->n_any);
  HIP_OK((hipError_t)mh::launch_trace_batch(f->scene->dev(), bl, stream));
  return MH_OK;
}

// A captured graph replays the launch exactly as it was recorded: refuse it when the
// scene's device view changed since (a rebuild, a load, another triangle test, buffers
// that moved) or when the tree deepened past the spill area it was given.
int graph_current(const mh_frame* f) {
  if (!f->exec) return fail(MH_E_STATE, "no captured graph for the current configuration");
  const mh::DevScene now = f->scene->dev();
  if (std::memcmp(&now, &f->captured, sizeof(now)) != 0)
    return fail(MH_E_STATE, "the scene was rebuilt or reloaded since the capture: capture "
                "the frame again");
  if (!spill_ok(f))
    return fail(MH_E_STATE, "the tree deepened since the capture: capture the frame again");
  return MH_OK;
}

std::string cache_key(const mh_frame* f) {
  std::ostringstream k;
  k << mh_internal::device_name() << ":T=" << f->scene->nt << ":depth=" << f->scene->depth
    << ":kind=" << f->scene->prim_kind << ":nb=" << f->nb << ":rays=" << f->total
    << ":anyhit=" << f->n_any;
  return k.str();
}

// The cache is a flat JSON object {"key": {"layout": i, ..., "ms": x}, ...} written here.
std::map<std::string, std::string> read_cache(const char* path) {
  std::map<std::string, std::string> m;
  std::ifstream in(path);
  if (!in) return m;
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string s = ss.str();
  size_t p = 0;
  for (;;) {
    const size_t k0 = s.find('"', p);
    if (k0 == std::string::npos) break;
    const size_t k1 = s.find('"', k0 + 1);
    const size_t o0 = s.find('{', k1);
    const size_t o1 = s.find('}', o0);
    if (k1 == std::string::npos || o0 == std::string::npos || o1 == std::string::npos) break;
    m[s.substr(k0 + 1, k1 - k0 - 1)] = s.substr(o0, o1 - o0 + 1);
    p = o1 + 1;
  }
  return m;
}

bool field(const std::string& obj, const char* name, double& v) {
  const std::string key = std::string("\"") + name + "\"";
  const size_t p = obj.find(key);
  if (p == std::string::npos) return false;
  const size_t c = obj.find(':', p);
  if (c == std::string::npos) return false;
  return std::sscanf(obj.c_str() + c + 1, "%lf", &v) == 1;
}

void write_cache(const char* path, const std::map<std::string, std::string>& m) {
  std::ofstream out(path);
  out << "{\n";
  size_t i = 0;
  for (const auto& kv : m) {
    out << "  \"" << kv.first << "\": " << kv.second << (++i < m.size() ? ",\n" : "\n");
  }
  out << "}\n";
}

int time_frame(mh_frame* f, int iters, bool graph, void* stream, float* ms) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  int rc = graph ? soa_ready(f, stream) : MH_OK;  // replays read the copy as it stands
  hipError_t e = hipEventRecord(a, st);
  for (int i = 0; i < iters && rc == MH_OK && e == hipSuccess; ++i) {
    if (graph)
      e = hipGraphLaunch(f->exec, st);
    else
      rc = launch(f, stream);
  }
  if (e == hipSuccess) e = hipEventRecord(b, st);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  if (e != hipSuccess) return fail(MH_E_HIP, "frame timing: %s", hipGetErrorString(e));
  *ms = t / (float)iters;
  return MH_OK;
}

}  // namespace

extern "C" {

int mh_frame_create(const mh_scene* scene, const mh_batch* batches, int n_batches,
                    mh_frame** out) {
  if (!scene || !out) return fail(MH_E_INVALID, "null argument");
  *out = nullptr;
  if (!scene->built) return fail(MH_E_STATE, "scene has no tree");
  int rc = ensure_init();
  if (rc) return rc;
  std::vector<uint8_t> host;
  int64_t total;
  int n_any;
  if ((rc = build_table(batches, n_batches, host, total, n_any))) return rc;
  mh_frame* f = new (std::nothrow) mh_frame();
  if (!f) return fail(MH_E_NOMEM, "frame");
  f->scene = scene;
  f->nb = n_batches;
  f->total = total;
  f->n_any = n_any;
  f->cfg = mh_launch_config{-1, 0, 0, 0, 0};
  if ((rc = f->table.reserve(host.size())) || (rc = f->counters.reserve(mh::kBatchCounterBytes)) ||
      (rc = prepare(f, kDefault))) {
    delete f;
    return rc;
  }
  hipError_t e = hipMemcpy(f->table.p, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(f->counters.p, 0, mh::kBatchCounterBytes);
  if (e != hipSuccess) {
    delete f;
    return fail(MH_E_HIP, "frame upload: %s", hipGetErrorString(e));
  }
  *out = f;
  return MH_OK;
}

int mh_frame_set_batches(mh_frame* f, const mh_batch* batches, int n_batches, void* stream) {
  if (!f) return fail(MH_E_INVALID, "null frame");
  if (n_batches != f->nb)
    return fail(MH_E_INVALID, "frame has %d batches, got %d", f->nb, n_batches);
  std::vector<uint8_t> host;
  int64_t total;
  int n_any;
  int rc = build_table(batches, n_batches, host, total, n_any);
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(f->table.p, host.data(), host.size(), hipMemcpyHostToDevice,
                        reinterpret_cast<hipStream_t>(stream)));
  HIP_OK(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
  // A captured graph holds the kernel specialised for the old mix of batch kinds.
  if (kinds_of(f->nb, n_any) != kinds_of(f->nb, f->n_any)) f->drop_graph();
  f->total = total;
  f->n_any = n_any;
  return MH_OK;
}

int mh_frame_set_config(mh_frame* f, const mh_launch_config* cfg) {
  if (!f) return fail(MH_E_INVALID, "null frame");
  int rc = ensure_init();
  if (rc) return rc;
  return prepare(f, cfg ? *cfg : kDefault);
}

int mh_frame_get_config(const mh_frame* f, mh_launch_config* cfg) {
  if (!f || !cfg) return fail(MH_E_INVALID, "null argument");
  *cfg = f->cfg;
  return MH_OK;
}

int mh_frame_launch(mh_frame* f, void* stream) {
  if (!f) return fail(MH_E_INVALID, "null frame");
  int rc = ensure_init();
  if (rc) return rc;
  return launch(f, stream);
}

int mh_frame_capture(mh_frame* f, void* stream) {
  if (!f) return fail(MH_E_INVALID, "null frame");
  if (!stream) return fail(MH_E_INVALID, "graph capture needs a created stream");
  int rc = ensure_init();
  if (rc) return rc;
  f->drop_graph();
  // the SoA copy is brought up to date outside the graph (before each replay as well)
  if ((rc = soa_ready(f, stream))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rc = launch(f, stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(MH_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  f->graph = g;
  HIP_OK(hipGraphInstantiate(&f->exec, g, nullptr, nullptr, 0));
  f->captured = f->scene->dev();
  return MH_OK;
}

int mh_frame_replay(mh_frame* f, void* stream) {
  if (!f) return fail(MH_E_INVALID, "null frame");
  int rc = graph_current(f);
  if (rc) return rc;
  if ((rc = soa_ready(f, stream))) return rc;
  HIP_OK(hipGraphLaunch(f->exec, reinterpret_cast<hipStream_t>(stream)));
  return MH_OK;
}

int mh_frame_time(mh_frame* f, int iters, int use_graph, void* stream, float* ms) {
  if (!f || !ms || iters < 1) return fail(MH_E_INVALID, "bad arguments");
  int rc = ensure_init();
  if (rc) return rc;
  if (use_graph && (rc = graph_current(f))) return rc;
  return time_frame(f, iters, use_graph != 0, stream, ms);
}

int mh_frame_autotune(mh_frame* f, int iters, const char* cache_path, int force,
                      mh_launch_config* best_out, float* best_ms_out) {
  if (!f || iters < 1) return fail(MH_E_INVALID, "bad arguments");
  int rc = ensure_init();
  if (rc) return rc;
  const std::string key = cache_key(f);
  std::map<std::string, std::string> cache;
  if (cache_path) cache = read_cache(cache_path);
  if (cache_path && !force && cache.count(key)) {
    const std::string& obj = cache[key];
    double v[9];
    const char* names[9] = {"layout", "stack_lds", "block", "blocks_per_cu",
                            "refill", "chunk", "xcd_remap", "reverse", "ms"};
    bool ok = true;
    for (int i = 0; i < 9; ++i) ok = ok && field(obj, names[i], v[i]);
    if (ok) {  // entries written before the reverse field was added are tuned again
      mh_launch_config c{(int)v[0], (int)v[1], (int)v[2], (int)v[3],
                         (int)v[4], (int)v[5], (int)v[6], (int)v[7]};
      if (mh_internal::check_launch_config(c) == MH_OK) {
        if ((rc = prepare(f, c))) return rc;
        if (best_out) *best_out = c;
        if (best_ms_out) *best_ms_out = (float)v[8];
        return MH_OK;
      }
    }
  }
  mh_launch_config best = f->cfg;
  float best_ms = 1e30f;
  auto trial = [&](const mh_launch_config& c) -> int {
    if (mh_internal::check_launch_config(c) != MH_OK) return MH_OK;  // outside the limits
    int r = prepare(f, c);
    if (r) return r;
    if ((r = launch(f, nullptr))) return r;  // warm-up
    float ms;
    if ((r = time_frame(f, iters, false, nullptr, &ms))) return r;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
    return MH_OK;
  };
  // Stage 1: launch shape (stack ring x workgroup size x workgroups per CU) jointly with
  // the schedule (static stride, small and large counter chunks), PAIR layout.
  for (int s : {8, 16})
    for (int b : {64, 128, 256})
      for (int p : {4, 8, 12, 16, 24})
        for (int c : {0, 64, 256})
          if ((rc = trial(mh_launch_config{MH_LAYOUT_PAIR, s, b, p, 64, c, 0}))) return rc;
  // Stage 2: chunk size x XCD remap at the best shape.
  mh_launch_config base = best;
  for (int c : {0, 64, 128, 256, 512, 1024})
    for (int x = 0; x < 2; ++x) {
      mh_launch_config t = base;
      t.chunk = c;
      t.xcd_remap = x;
      if ((rc = trial(t))) return rc;
    }
  // Stage 3: the SoA and the reference's 32-byte node layouts and the deepest LDS ring.
  base = best;
  for (int l : {MH_LAYOUT_PAIR, MH_LAYOUT_SOA, MH_LAYOUT_REF})
    for (int s : {8, 16, 32}) {
      mh_launch_config t = base;
      t.layout = l;
      t.stack_lds = s;
      if ((rc = trial(t))) return rc;
    }
  // Stage 4: per-lane refill, with shapes of its own: its waves stay resident and take
  // rays in chunks, so grids of about the resident wave count (five to seven per SIMD)
  // and chunks that leave every wave several grabs; the best ring and layout so far, in
  // both hand-out orders (the order moves the best shape).
  base = best;
  for (int rev = 0; rev < 2; ++rev)
    for (int r : {16, 32, 48})
      for (int c : {64, 128, 256})
        for (const auto& bp : {std::pair<int, int>{256, 5}, {256, 6}, {256, 7}, {128, 10},
                               {128, 12}, {128, 14}}) {
          mh_launch_config t = base;
          t.refill = r;
          t.chunk = c;
          t.block = bp.first;
          t.blocks_per_cu = bp.second;
          t.xcd_remap = 0;
          t.reverse = rev;
          if ((rc = trial(t))) return rc;
        }
  // Stage 5: the hand-out order (a table that lists its costly batches last ends on its
  // short rays when handed out from the end).
  {
    mh_launch_config t = best;
    t.reverse = 1 - t.reverse;
    if ((rc = trial(t))) return rc;
  }
  if ((rc = prepare(f, best))) return rc;
  if (cache_path) {
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "{\"layout\": %d, \"stack_lds\": %d, \"block\": %d, \"blocks_per_cu\": %d, "
                  "\"refill\": %d, \"chunk\": %d, \"xcd_remap\": %d, \"reverse\": %d, "
                  "\"ms\": %.5f}",
                  best.layout, best.stack_lds, best.block, best.blocks_per_cu, best.refill,
                  best.chunk, best.xcd_remap, best.reverse, best_ms);
    cache[key] = buf;
    write_cache(cache_path, cache);
  }
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return MH_OK;
}

int mh_frame_destroy(mh_frame* f) {
  if (!f) return MH_OK;
  (void)ensure_init();
  (void)hipDeviceSynchronize();
  delete f;
  return MH_OK;
}

}  // extern "C"
