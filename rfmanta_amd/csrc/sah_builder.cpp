// Host-side DynBVH construction (see sah_builder.h).
//
// Structure: every object's box and centroid are computed once; a recursive splitter
// builds a temporary tree with explicit child links, running large right subtrees and
// the three per-axis sweeps of large nodes on their own threads; a final pass lays the
// tree out in the reference numbering (children of the k-th inner node in depth-first
// preorder at slots 2k+1, 2k+2, DynBVH.cc:1832-1849).
//
// The split arithmetic is the reference's, operation for operation, because the chosen
// splits must be identical: float box areas (BBox::computeArea, Core/Geometry/BBox.h:
// 115-119) widened to double only where the reference divides by them, the SAH cost
// (n_l*A_l + n_r*A_r)*invA*C_isec + C_trav (DynBVH.cc:2418-2424), the strict '<' scans
// (right-to-left over events, axis 0..2), std::sort on centroid keys with the same
// comparison, and the two-object rule (DynBVH.cc:2277-2308).

#include "sah_builder.h"

#include "rtsah.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace mh {
namespace {

inline float pick_lo(float a, float b) { return a < b ? a : b; }  // Core/Math/MinMax.h:88-91
inline float pick_hi(float a, float b) { return a > b ? a : b; }  // Core/Math/MinMax.h:83-86

inline void box_clear(Aabb& a) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = FLT_MAX;
    a.hi[k] = -FLT_MAX;
  }
}
inline void box_grow(Aabb& a, const Aabb& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = pick_lo(a.lo[k], b.lo[k]);
    a.hi[k] = pick_hi(a.hi[k], b.hi[k]);
  }
}
inline float box_area(const Aabb& a) {
  const float ex = a.hi[0] - a.lo[0], ey = a.hi[1] - a.lo[1], ez = a.hi[2] - a.lo[2];
  return 2 * (ex * ey + ey * ez + ex * ez);
}
inline float inv_area(const Aabb& a) { return (float)(1.0f / (double)box_area(a)); }

struct Key {
  float pos;
  int32_t id;
};

struct TmpNode {
  Aabb box;
  int32_t kid[2];
  int32_t first, count;
  int axis;
};

struct AxisSweep {
  float cost = FLT_MAX;
  int event = -1;
  Aabb right;
  std::vector<Key> keys;
};

class Splitter {
 public:
  Splitter(const float* v, uint32_t nv, const uint32_t* ix, uint32_t nt, int threads,
           std::vector<int32_t>& ids)
      : V_(v), NV_(nv), I_(ix), T_(nt), ids_(ids), max_threads_(threads) {}

  int run(Aabb& root_box) {
    boxes_.resize(T_);
    cen_.resize(3 * (size_t)T_);
    ids_.resize(T_);
    par_for(T_, [&](size_t lo, size_t hi) {
      for (size_t t = lo; t < hi; ++t) {
        const float* p0 = V_ + 3 * (size_t)I_[3 * t + 0];
        const float* p1 = V_ + 3 * (size_t)I_[3 * t + 1];
        const float* p2 = V_ + 3 * (size_t)I_[3 * t + 2];
        Aabb b;
        for (int k = 0; k < 3; ++k) {
          b.lo[k] = pick_lo(pick_lo(p0[k], p1[k]), p2[k]);
          b.hi[k] = pick_hi(pick_hi(p0[k], p1[k]), p2[k]);
        }
        boxes_[t] = b;
        for (int k = 0; k < 3; ++k) cen_[3 * t + k] = b.hi[k] * 0.5f + b.lo[k] * 0.5f;
        ids_[t] = (int32_t)t;
      }
    });
    // Mesh::computeBounds is the box of every vertex (Model/Groups/Mesh.cc:195-225).
    box_clear(root_box);
    for (uint32_t i = 0; i < NV_; ++i) {
      Aabb p;
      for (int k = 0; k < 3; ++k) p.lo[k] = p.hi[k] = V_[3 * (size_t)i + k];
      box_grow(root_box, p);
    }
    tmp_.resize(2 * (size_t)T_);
    next_.store(0);
    live_.store(1);
    return split(root_box, 0, (int)T_);
  }

  const std::vector<TmpNode>& tmp() const { return tmp_; }

 private:
  template <class F>
  void par_for(size_t n, F f) {
    int nt = std::max(1, std::min<int>(max_threads_, (int)(n / 65536) + 1));
    if (nt == 1) {
      f(0, n);
      return;
    }
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i)
      th.emplace_back([&, i]() { f(n * i / nt, n * (i + 1) / nt); });
    for (auto& x : th) x.join();
  }

  void sweep_axis(const Aabb& box, int b, int e, int axis, AxisSweep& out) {
    const int n = e - b;
    out.keys.resize(n);
    for (int k = 0; k < n; ++k) {
      const int32_t id = ids_[b + k];
      out.keys[k].pos = cen_[3 * (size_t)id + axis];
      out.keys[k].id = id;
    }
    std::sort(out.keys.begin(), out.keys.end(),
              [](const Key& x, const Key& y) { return x.pos < y.pos; });
    // Area of the union of the first k sorted objects, k = 0..n-1.
    std::vector<float> left_area(n);
    Aabb acc;
    box_clear(acc);
    for (int k = 0; k < n; ++k) {
      left_area[k] = box_area(acc);
      box_grow(acc, boxes_[out.keys[k].id]);
    }
    const float inv = inv_area(box);
    Aabb right;
    box_clear(right);
    out.cost = FLT_MAX;
    out.event = -1;
    for (int k = n - 1; k >= 0; --k) {
      box_grow(right, boxes_[out.keys[k].id]);
      if (k > 0 && n - k > 0) {
        float c = (float)k * left_area[k] + (float)(n - k) * box_area(right);
        c *= inv;
        c *= 1.f;  // C_isec
        c += 1.f;  // C_trav
        if (c < out.cost) {
          out.cost = c;
          out.event = k;
          out.right = right;
        }
      }
    }
  }

  int make_leaf(int self, int b, int e) {
    std::sort(ids_.begin() + b, ids_.begin() + e);
    TmpNode& t = tmp_[self];
    t.kid[0] = t.kid[1] = -1;
    t.first = b;
    t.count = e - b;
    t.axis = 0;
    return self;
  }

  int split(const Aabb& box, int b, int e) {
    const int self = next_.fetch_add(1);
    tmp_[self].box = box;
    const int n = e - b;
    if (n == 1) return make_leaf(self, b, e);
    int axis = -1, mid = 0;
    Aabb lbox, rbox;
    if (n == 2) {
      const float node_a = box_area(box);
      const float l_a = box_area(boxes_[ids_[b]]);
      const float r_a = box_area(boxes_[ids_[b + 1]]);
      const float c = 1.f + (l_a * 1.f + r_a * 1.f) / node_a;
      if (!(c < 2.f)) return make_leaf(self, b, e);
      const float ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1],
                  ez = box.hi[2] - box.lo[2];
      axis = (ex > ey) ? (ex > ez ? 0 : 2) : (ey > ez ? 1 : 2);  // indexOfMaxComponent
      if (boxes_[ids_[b]].lo[axis] > boxes_[ids_[b + 1]].lo[axis]) std::swap(ids_[b], ids_[b + 1]);
      lbox = boxes_[ids_[b]];
      rbox = boxes_[ids_[b + 1]];
      mid = b + 1;
    } else {
      AxisSweep sw[3];
      if (n >= 32768) {
        std::thread t1([&]() { sweep_axis(box, b, e, 1, sw[1]); });
        std::thread t2([&]() { sweep_axis(box, b, e, 2, sw[2]); });
        sweep_axis(box, b, e, 0, sw[0]);
        t1.join();
        t2.join();
      } else {
        for (int a = 0; a < 3; ++a) sweep_axis(box, b, e, a, sw[a]);
      }
      float best = 1.f * (float)n;  // leaf cost
      for (int a = 0; a < 3; ++a)
        if (sw[a].event != -1 && sw[a].cost < best) {
          best = sw[a].cost;
          axis = a;
        }
      if (axis < 0) return make_leaf(self, b, e);
      const AxisSweep& w = sw[axis];
      mid = b + w.event;
      box_clear(lbox);
      for (int k = 0; k < n; ++k) {
        ids_[b + k] = w.keys[k].id;
        if (k < w.event) box_grow(lbox, boxes_[w.keys[k].id]);
      }
      rbox = w.right;
    }
    tmp_[self].axis = axis;
    tmp_[self].first = b;
    tmp_[self].count = 0;
    int left = -1, right = -1;
    bool spawned = false;
    std::thread th;
    if (e - mid >= 16384 && live_.fetch_add(1) < max_threads_) {
      spawned = true;
      th = std::thread([&]() {
        right = split(rbox, mid, e);
        live_.fetch_sub(1);
      });
    } else if (e - mid >= 16384) {
      live_.fetch_sub(1);
    }
    left = split(lbox, b, mid);
    if (spawned)
      th.join();
    else
      right = split(rbox, mid, e);
    tmp_[self].kid[0] = left;
    tmp_[self].kid[1] = right;
    return self;
  }

  const float* V_;
  uint32_t NV_;
  const uint32_t* I_;
  uint32_t T_;
  std::vector<int32_t>& ids_;
  int max_threads_;
  std::vector<Aabb> boxes_;
  std::vector<float> cen_;
  std::vector<TmpNode> tmp_;
  std::atomic<int> next_{0};
  std::atomic<int> live_{1};
};

// ---------------------------------------------------------------------------
// RTSAH child order (DynBVH.cc:1594-1678, Model/Groups/TreeTraversalProbability.*)
// ---------------------------------------------------------------------------
class RtsahPass {
 public:
  RtsahPass(std::vector<RefNode>& n, int threads)
      : nodes_(n), max_threads_(threads), log_(rtsah::kLogTable) {
    rtsah::fill_log_table(log_.data());
  }
  void run() {
    live_.store(1);
    visit(0, 0);
  }

 private:
  rtsah::CostPair visit(int id, int depth) {
    RefNode& nd = nodes_[id];
    if (nd.leaf()) return rtsah::leaf_cost(nd.children);
    rtsah::CostPair L, R;
    if (depth < 12 && live_.fetch_add(1) < max_threads_) {
      std::thread th([&]() {
        R = visit(nd.child + 1, depth + 1);
        live_.fetch_sub(1);
      });
      L = visit(nd.child, depth + 1);
      th.join();
    } else {
      if (depth < 12) live_.fetch_sub(1);
      L = visit(nd.child, depth + 1);
      R = visit(nd.child + 1, depth + 1);
    }
    bool left_cheaper;
    const rtsah::CostPair c = rtsah::combine(log_.data(), nd.box, nodes_[nd.child].box,
                                             nodes_[nd.child + 1].box, L, R, left_cheaper);
    if (left_cheaper)
      nd.bits |= 1;
    else
      nd.bits &= (uint8_t)~1;
    return c;
  }
  std::vector<RefNode>& nodes_;
  int max_threads_;
  std::atomic<int> live_{1};
  std::vector<float> log_;
};

}  // namespace

DynBvh build_dynbvh(const float* verts, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris,
                    int threads) {
  if (n_tris == 0) throw std::invalid_argument("DynBVH build over an empty mesh");
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  DynBvh out;
  Splitter sp(verts, n_verts, idx, n_tris, threads, out.ids);
  Aabb root_box;
  const int root = sp.run(root_box);
  const std::vector<TmpNode>& tn = sp.tmp();

  RefNode blank;
  box_clear(blank.box);
  blank.child = 0;
  blank.axis = 0;
  blank.bits = 0;
  blank.children = 0;
  out.nodes.assign(2 * (size_t)n_tris, blank);

  struct Item {
    int t, r;
  };
  std::vector<Item> stack;
  stack.push_back({root, 0});
  int next_free = 1, used = 0;
  while (!stack.empty()) {
    const Item it = stack.back();
    stack.pop_back();
    const TmpNode& t = tn[it.t];
    RefNode& r = out.nodes[it.r];
    r.box = (it.r == 0) ? root_box : t.box;
    ++used;
    if (t.kid[0] < 0) {
      r.child = t.first;
      r.children = (int16_t)t.count;
      r.axis = 0;
    } else {
      const int c = next_free;
      next_free += 2;
      r.child = c;
      r.children = 0;
      r.axis = (uint8_t)t.axis;
      stack.push_back({t.kid[1], c + 1});
      stack.push_back({t.kid[0], c});
    }
  }
  out.num_nodes = used;
  mark_large_subtrees(out.nodes);
  return out;
}

void compute_rtsah(std::vector<RefNode>& nodes, int threads) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  RtsahPass pass(nodes, threads);
  pass.run();
}

void mark_large_subtrees(std::vector<RefNode>& nodes, int workers) {
  const unsigned limit = (unsigned)(nodes.size() / (20 * (size_t)(workers > 0 ? workers : 1)));
  // Post-order over an explicit stack: subtree record counts.
  std::vector<unsigned> size(nodes.size(), 0);
  std::vector<std::pair<int, bool>> st;
  st.push_back({0, false});
  while (!st.empty()) {
    auto [id, done] = st.back();
    st.pop_back();
    RefNode& n = nodes[id];
    if (n.leaf()) {
      size[id] = 1;
      n.bits &= (uint8_t)~2;
      continue;
    }
    if (!done) {
      st.push_back({id, true});
      st.push_back({n.child + 1, false});
      st.push_back({n.child, false});
    } else {
      size[id] = size[n.child] + size[n.child + 1] + 1;
      if (size[id] > limit)
        n.bits |= 2;
      else
        n.bits &= (uint8_t)~2;
    }
  }
}

}  // namespace mh
