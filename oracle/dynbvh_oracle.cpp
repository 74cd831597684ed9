// ORACLE — test infrastructure only.
//
// A CPU restatement of Manta's DynBVH hot path, used by the parity tests, by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py as the checker.
// The product library (rfmanta_amd/libmh.so) never links, loads or calls it.
//
// What it restates (reference file:line):
//   * SAH build: DynBVH::rebuild/build/partitionSAH/buildEvents/partition2Objs
//     (Model/Groups/DynBVH.cc:1699-1857, 2277-2440), BBox (Core/Geometry/BBox.h:45-130),
//     Mesh::getBBox (Model/Groups/Mesh.cc:228-237), Mesh::computeBounds (Mesh.cc:195-225).
//   * RTSAH child order for any-hit: DynBVH::computeSubTreeTraversalCost
//     (DynBVH.cc:1594-1678) with TreeTraversalProb::computeProbs
//     (Model/Groups/TreeTraversalProbability.cc:3-95, .h:13-143), fast_atan2f
//     (Core/Math/Trig.h:153-178) and ICSI_log (Core/Math/ExponSSE.h:20-73).
//   * Packet traversal: DynBVH::intersect/intersectNode/firstIntersects/lastIntersects,
//     the scalar (non-MANTA_SSE) branches (DynBVH.cc:101-404, 655-725, 882-905),
//     RayPacket::hit/maskRay/wasHit (Interface/RayPacket.h:407-442,488-502,531-542),
//     RayPacket::actualComputeInverseDirections scalar branch (Interface/RayPacket.cc:105-109).
//   * Kensler-Shirley test, scalar branch (Model/Primitives/KenslerShirleyTriangle.cc:336-377).
//   * Wald test: setPoints (Model/Primitives/WaldTriangle.cc:64-127), scalar intersect
//     (WaldTriangle.cc:489-555).
//   * Refit without tree rotations: DynBVH::update -> parallelUpdateBounds(numProcs==1)
//     -> updateBounds<false> (DynBVH.cc:1574-1581, 2012-2123).
//   * DynBVH binary file layout (DynBVH.cc:2442-2501): u32 n_ids, i32 ids[], u32 n_nodes,
//     32-byte nodes.
//
//   * The SSE build of the packet path (MANTA_SSE, the reference's x86 default), as a
//     second mode for the CPU baseline and for classifying the rcpps hazard:
//     RayPacket::actualComputeInverseDirections SSE branch (Interface/RayPacket.cc:80-104)
//     with oneOver (Core/Math/SSEDefs.h:193-197), the SSE interval data of
//     DynBVH::intersect (DynBVH.cc:130-236) with min4f/max4f (SSEDefs.h:269-299),
//     firstIntersects and lastIntersects SSE branches (DynBVH.cc:423-629, 743-879),
//     KenslerShirleyTriangle::intersect SSE branch (KenslerShirleyTriangle.cc:134-332,
//     cross4 :127-132, dot4 SSEDefs.h:148-154) and RayPacket::hitFinal
//     (Interface/RayPacket.h:545-605). Kensler-Shirley only.
//
// Numerics: every float operation is IEEE binary32 with round-to-nearest and no
// contraction (g++ -O2 on baseline x86-64 emits no FMA). The scalar build of the
// reference is the default mode; the SSE mode's rcpps reciprocals are those of the CPU it
// runs on (Intel and AMD differ in the low bits), so its results are machine specific.
//
// Parity pins: tests/test_oracle_pins.py checks this file against brute force over all
// triangles, double-precision closed forms, hand-computed known-answer rays and the BVH
// invariants; see that file for which function each check pins.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include <xmmintrin.h>

namespace {

// include/CMakeLists.txt:185,188 with include/Parameters.h.CMakeTemplate:33,36.
const float kTEps = (float)1.e-4;
const float kMaxT = (float)1.e19;
// DynBVH.cc:29-30
const float kCIsec = 1.f;
const float kCTrav = 1.f;

struct Vec {
  float v[3];
};
inline Vec vsub(const Vec& a, const Vec& b) {
  Vec r;
  r.v[0] = a.v[0] - b.v[0];
  r.v[1] = a.v[1] - b.v[1];
  r.v[2] = a.v[2] - b.v[2];
  return r;
}
// Core/Geometry/Vector.h:526-531 (left-to-right sum)
inline float vdot(const Vec& a, const Vec& b) {
  return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2];
}
// Core/Geometry/Vector.h:534-539
inline Vec vcross(const Vec& a, const Vec& b) {
  Vec r;
  r.v[0] = a.v[1] * b.v[2] - a.v[2] * b.v[1];
  r.v[1] = a.v[2] * b.v[0] - a.v[0] * b.v[2];
  r.v[2] = a.v[0] * b.v[1] - a.v[1] * b.v[0];
  return r;
}
// Core/Math/MinMax.h:83-91
inline float fmin_(float a, float b) { return a < b ? a : b; }
inline float fmax_(float a, float b) { return a > b ? a : b; }

// Core/Geometry/BBox.h
struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int i = 0; i < 3; ++i) {
      lo[i] = FLT_MAX;
      hi[i] = -FLT_MAX;
    }
  }
  void extendPoint(const float* p) {
    for (int i = 0; i < 3; ++i) {
      lo[i] = fmin_(lo[i], p[i]);
      hi[i] = fmax_(hi[i], p[i]);
    }
  }
  void extendBox(const Box& b) {
    for (int i = 0; i < 3; ++i) {
      lo[i] = fmin_(lo[i], b.lo[i]);
      hi[i] = fmax_(hi[i], b.hi[i]);
    }
  }
  // BBox::computeArea returns double, but the arithmetic is float (BBox.h:115-119).
  double area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    float a = 2 * (dx * dy + dy * dz + dx * dz);
    return (double)a;
  }
};

// Same 32-byte layout as DynBVH::BVHNode (Model/Groups/DynBVH.h:37-77): bounds @0,
// child @24, axis @28, {isLeftCheaper bit0, isLargeSubtree bit1} @29, children @30.
struct Node32 {
  Box bounds;
  int32_t child;
  uint8_t axis;
  uint8_t bits;
  int16_t children;
  bool isLeaf() const { return children != 0; }
  bool leftCheaper() const { return bits & 1; }
};
static_assert(sizeof(Node32) == 32, "BVHNode is 32 bytes");

struct Mesh {
  const float* verts;  // 3 floats per vertex
  int nverts;
  const uint32_t* idx;  // 3 per triangle
  int ntris;
  const float* vtx(int tri, int k) const { return verts + 3 * (size_t)idx[3 * (size_t)tri + k]; }
};

// Mesh::getBBox (Mesh.cc:228-237) into a reset box (DynBVH.cc:1736-1737).
inline Box triBox(const Mesh& m, int t) {
  Box b;
  const float* p1 = m.vtx(t, 0);
  for (int i = 0; i < 3; ++i) {
    b.lo[i] = p1[i];
    b.hi[i] = p1[i];
  }
  b.extendPoint(m.vtx(t, 1));
  b.extendPoint(m.vtx(t, 2));
  return b;
}

// ---------------------------------------------------------------------------
// SAH build (DynBVH.cc:1699-1857, 2277-2440)
// ---------------------------------------------------------------------------
struct Builder {
  const Mesh& mesh;
  std::vector<Node32>& nodes;
  std::vector<int32_t>& ids;
  std::vector<Box> obj_bounds;
  std::vector<Vec> obj_centroids;
  int next_free = 1;
  int num_nodes = 0;

  struct Partition {
    Box bounds_l, bounds_r;
    int axis;
    int split;
  };
  struct Eval {
    Box bounds_r;
    float position;
    float cost;
    int num_left, num_right, event, axis;
  };
  struct Event {
    float position;
    int obj_id;
    float left_area;
  };

  Builder(const Mesh& m, std::vector<Node32>& n, std::vector<int32_t>& i)
      : mesh(m), nodes(n), ids(i) {}

  // DynBVH.cc:2277-2308
  Partition partition2(int nodeID, int objBegin) {
    Partition p;
    p.bounds_l.reset();
    p.bounds_r.reset();
    const float nodeArea = (float)nodes[nodeID].bounds.area();
    const float leftArea = (float)obj_bounds[ids[objBegin]].area();
    const float rightArea = (float)obj_bounds[ids[objBegin + 1]].area();
    const float leafCost = kCIsec * 2;
    const float newCost = kCTrav + (leftArea * kCIsec + rightArea * kCIsec) / nodeArea;
    if (newCost < leafCost) {
      const Box& nb = nodes[nodeID].bounds;
      float d0 = nb.hi[0] - nb.lo[0], d1 = nb.hi[1] - nb.lo[1], d2 = nb.hi[2] - nb.lo[2];
      // Vector::indexOfMaxComponent (Core/Geometry/Vector.h:382-393)
      int axis = (d0 > d1) ? (d0 > d2 ? 0 : 2) : (d1 > d2 ? 1 : 2);
      p.axis = axis;
      p.split = objBegin + 1;
      if (obj_bounds[ids[objBegin]].lo[axis] > obj_bounds[ids[objBegin + 1]].lo[axis])
        std::swap(ids[objBegin], ids[objBegin + 1]);
      p.bounds_l = obj_bounds[ids[objBegin]];
      p.bounds_r = obj_bounds[ids[objBegin + 1]];
      return p;
    }
    p.axis = -1;
    return p;
  }

  void makeEvents(int first, int last, int axis, std::vector<Event>& ev) {
    ev.clear();
    ev.reserve(last - first);
    for (int i = first; i < last; ++i) {
      Event e;
      e.position = obj_centroids[ids[i]].v[axis];
      e.obj_id = ids[i];
      e.left_area = 0.f;
      ev.push_back(e);
    }
    std::sort(ev.begin(), ev.end(),
              [](const Event& x, const Event& y) { return x.position < y.position; });
  }

  // DynBVH.cc:2372-2440
  bool buildEvents(int parentID, int first, int last, int axis, Eval& best) {
    std::vector<Event> ev;
    makeEvents(first, last, axis, ev);
    const int n = (int)ev.size();
    Box left_box;
    left_box.reset();
    std::vector<int> num_left(n), num_right(n);
    int nl = 0, nr = n;
    for (int i = 0; i < n; ++i) {
      num_left[i] = nl;
      num_right[i] = nr;
      ev[i].left_area = (float)left_box.area();
      left_box.extendBox(obj_bounds[ev[i].obj_id]);
      nl++;
      nr--;
    }
    Box right_box;
    right_box.reset();
    best.cost = FLT_MAX;
    best.event = -1;
    const float inv_parent_area = (float)(1.0f / nodes[parentID].bounds.area());
    for (int i = n - 1; i >= 0; --i) {
      right_box.extendBox(obj_bounds[ev[i].obj_id]);
      if (num_left[i] > 0 && num_right[i] > 0) {
        const float right_area = (float)right_box.area();
        float cost = (num_left[i] * ev[i].left_area + num_right[i] * right_area);
        cost *= inv_parent_area;
        cost *= kCIsec;
        cost += kCTrav;
        if (cost < best.cost) {
          best.bounds_r = right_box;
          best.cost = cost;
          best.position = ev[i].position;
          best.axis = axis;
          best.event = i;
          best.num_left = num_left[i];
          best.num_right = num_right[i];
        }
      }
    }
    return best.event != -1;
  }

  // DynBVH.cc:2310-2370
  Partition partitionSAH(int nodeID, int objBegin, int objEnd) {
    Partition p;
    p.bounds_l.reset();
    p.bounds_r.reset();
    p.split = 0;
    const int n = objEnd - objBegin;
    if (n == 1) {
      p.axis = -1;
      return p;
    }
    Eval best;
    best.cost = kCIsec * n;
    best.axis = -1;
    best.position = FLT_MAX;
    best.event = -1;
    best.bounds_r.reset();
    if (n == 2) return partition2(nodeID, objBegin);
    for (int axis = 0; axis < 3; ++axis) {
      Eval e;
      if (buildEvents(nodeID, objBegin, objEnd, axis, e)) {
        if (e.cost < best.cost) best = e;
      }
    }
    p.bounds_r = best.bounds_r;
    p.axis = best.axis;
    if (p.axis != -1) {
      std::vector<Event> ev;
      makeEvents(objBegin, objEnd, p.axis, ev);
      p.split = objBegin + best.event;
      for (int i = objBegin; i < objEnd; ++i) {
        ids[i] = ev[i - objBegin].obj_id;
        if (i < p.split) p.bounds_l.extendBox(obj_bounds[ids[i]]);
      }
    }
    return p;
  }

  // DynBVH.cc:1801-1857 (exact build, no approximate path, no rotation costs)
  void build(int nodeID, int objBegin, int objEnd) {
    Partition p = partitionSAH(nodeID, objBegin, objEnd);
    Node32& node = nodes[nodeID];
    if (p.axis == -1) {
      node.child = objBegin;
      node.children = (int16_t)(objEnd - objBegin);
      node.axis = 0;
      std::sort(ids.begin() + objBegin, ids.begin() + objEnd);
      num_nodes++;
    } else {
      int my_spot = next_free;
      next_free += 2;
      node.child = my_spot;
      node.children = 0;
      node.axis = (uint8_t)p.axis;
      num_nodes++;
      nodes[my_spot + 0].bounds = p.bounds_l;
      nodes[my_spot + 1].bounds = p.bounds_r;
      build(my_spot + 0, objBegin, p.split);
      build(my_spot + 1, p.split, objEnd);
    }
  }

  // DynBVH.cc:1699-1793 with numProcs == 1
  void rebuild() {
    const int N = mesh.ntris;
    nodes.assign(2 * (size_t)N, Node32());
    for (auto& nd : nodes) {
      nd.bounds.reset();
      nd.child = 0;
      nd.axis = 0;
      nd.bits = 0;
      nd.children = 0;
    }
    ids.resize(N);
    obj_bounds.resize(N);
    obj_centroids.resize(N);
    // Mesh::computeBounds: box of all vertices (Mesh.cc:195-225)
    Box root;
    root.reset();
    for (int v = 0; v < mesh.nverts; ++v) root.extendPoint(mesh.verts + 3 * (size_t)v);
    nodes[0].bounds = root;
    for (int i = 0; i < N; ++i) {
      ids[i] = i;
      obj_bounds[i] = triBox(mesh, i);
      // BBox::center = Interpolate(min, max, 0.5f) (BBox.h:63-65, MiscMath.h:136-139)
      for (int a = 0; a < 3; ++a)
        obj_centroids[i].v[a] = obj_bounds[i].hi[a] * 0.5f + obj_bounds[i].lo[a] * (1 - 0.5f);
    }
    num_nodes = 0;
    next_free = 1;
    build(0, 0, N);
    // computeSubTreeSizes with largeSubtreeSize = nodes.size() / 20 (DynBVH.cc:1776, 1971-1996)
    large_subtree = (unsigned)(nodes.size() / 20);
    subtreeSizes(0);
  }

  unsigned large_subtree = 0;
  unsigned subtreeSizes(int id) {
    Node32& nd = nodes[id];
    if (nd.isLeaf()) {
      nd.bits &= ~2;
      return 1;
    }
    const unsigned size = subtreeSizes(nd.child) + subtreeSizes(nd.child + 1) + 1;
    if (size > large_subtree)
      nodes[id].bits |= 2;
    else
      nodes[id].bits &= ~2;
    return size;
  }
};

// ---------------------------------------------------------------------------
// RTSAH (DynBVH.cc:1594-1678, TreeTraversalProbability.{h,cc})
// ---------------------------------------------------------------------------
struct IcsiLog {  // Core/Math/ExponSSE.h:20-73, precision 15
  int n = 15;
  std::vector<float> table;
  IcsiLog() {
    table.resize(1 << n);
    float oneToTwo = 1.0f + (1.0f / (float)(1 << (n + 1)));
    for (int i = 0; i < (1 << n); ++i) {
      table[i] = logf(oneToTwo) / 0.69314718055995f;
      oneToTwo += 1.0f / (float)(1 << n);
    }
  }
  float log(float val) const {
    int32_t bits;
    std::memcpy(&bits, &val, 4);
    const int e = ((bits >> 23) & 255) - 127;
    const int man = (bits & 0x7FFFFF) >> (23 - n);
    return ((float)(e) + table[man]) * 0.69314718055995f;
  }
};

inline float fastAtan2f(float y, float x) {  // Core/Math/Trig.h:153-178
  const float kPi2 = (float)M_PI_2, kPi = (float)M_PI;
  if (x == 0.0f) {
    if (y > 0.0f) return kPi2;
    if (y == 0.0f) return 0.0f;
    return -kPi2;
  }
  float atan;
  float z = y / x;
  if (fabsf(z) < 1.0f) {
    atan = z / (1.0f + 0.28f * z * z);
    if (x < 0.0f) {
      if (y < 0.0f) return atan - kPi;
      return atan + kPi;
    }
  } else {
    atan = kPi2 - z / (z * z + 0.28f);
    if (y < 0.0f) return atan - kPi;
  }
  return atan;
}

struct TTP {
  IcsiLog L;
  static float altSign(int i) { return i % 2 == 1 ? -1.0f : 1.0f; }
  float perpG(float v, float z, float x, float y, float z0, float x1) const {
    const float a = y - v, b = z0 - z, c = x - x1;
    const float aa = a * a, cc = c * c, bb = b * b;
    const float s = sqrtf(cc + bb);
    return a * s * fastAtan2f(a, s) + .25f * (aa - bb - cc) * L.log(aa + bb + cc);
  }
  static int otherIndex(unsigned a, unsigned b) { return (~(a + b)) & 3; }
  float perpF(const Box& b0, const Box& b1, int d1, int d2) const {
    const int axis = otherIndex(d1, d2);
    const float v[2] = {b1.lo[axis], b1.hi[axis]};
    const float y[2] = {b0.lo[axis], b0.hi[axis]};
    const float x[2] = {b0.lo[d2], b0.hi[d2]};
    const float z[2] = {b1.lo[d1], b1.hi[d1]};
    const float x1 = b1.lo[d2];
    const float z0 = b0.lo[d1];
    float F = 0;
    for (int i = 1; i <= 2; ++i)
      for (int j = 1; j <= 2; ++j)
        for (int k = 1; k <= 2; ++k)
          for (int m = 1; m <= 2; ++m)
            F += perpG(v[i - 1], z[j - 1], x[k - 1], y[m - 1], z0, x1) * altSign(i + j + k + m);
    F *= (float)M_1_PI;
    return fabsf(F);
  }
  float parG(float u, float v, float x, float y, float zi) const {
    const float a = (x - u) * zi, b = (y - v) * zi;
    const float aa = a * a, bb = b * b;
    const float s1 = sqrtf(1 + aa), s2 = sqrtf(1 + bb);
    return b * s1 * fastAtan2f(b, s1) + a * s2 * fastAtan2f(a, s2) - 0.5f * L.log(1 + aa + bb);
  }
  float parF(const Box& b0, const Box& b1, int flat) const {
    const int d1 = (flat + 1) % 3, d2 = (flat + 2) % 3;
    const float y[2] = {b0.lo[d1], b0.hi[d1]};
    const float v[2] = {b1.lo[d1], b1.hi[d1]};
    const float x[2] = {b0.lo[d2], b0.hi[d2]};
    const float u[2] = {b1.lo[d2], b1.hi[d2]};
    const float z = b0.lo[flat] - b1.lo[flat];
    const float zi = 1.0f / z;
    float F = 0;
    for (int i = 1; i <= 2; ++i)
      for (int j = 1; j <= 2; ++j)
        for (int k = 1; k <= 2; ++k)
          for (int m = 1; m <= 2; ++m)
            F += parG(u[i - 1], v[j - 1], x[k - 1], y[m - 1], zi) * altSign(i + j + k + m);
    F *= z * z * (float)M_1_PI;
    return fabsf(F);
  }
  static float overlap(const Box& a, const Box& b) {
    Box o = a;
    for (int i = 0; i < 3; ++i) {
      o.lo[i] = fmax_(o.lo[i], b.lo[i]);
      o.hi[i] = fmin_(o.hi[i], b.hi[i]);
    }
    float d0 = o.hi[0] - o.lo[0], d1 = o.hi[1] - o.lo[1], d2 = o.hi[2] - o.lo[2];
    if (d0 < 0 || d1 < 0 || d2 < 0) return 0;
    return (float)o.area();
  }
  // TreeTraversalProbability.cc:3-95
  void probs(const Box& parent, const Box& bL, const Box& bR, float& lr, float& jl,
             float& jr) const {
    float dl[3] = {bL.hi[0] - bL.lo[0], bL.hi[1] - bL.lo[1], bL.hi[2] - bL.lo[2]};
    float dr[3] = {bR.hi[0] - bR.lo[0], bR.hi[1] - bR.lo[1], bR.hi[2] - bR.lo[2]};
    const bool Linf[3] = {dl[1] == 0 || dl[2] == 0, dl[0] == 0 || dl[2] == 0,
                          dl[0] == 0 || dl[1] == 0};
    const bool Rinf[3] = {dr[1] == 0 || dr[2] == 0, dr[0] == 0 || dr[2] == 0,
                          dr[0] == 0 || dr[1] == 0};
    lr = 0;
    for (int i = 0; i < 6; ++i) {
      const int ii = i % 3;
      if (Linf[ii]) continue;
      if (i < 3) {
        if (bL.hi[ii] >= bR.hi[ii]) continue;
      } else {
        if (bL.lo[ii] <= bR.lo[ii]) continue;
      }
      for (int j = 0; j < 6; ++j) {
        if (i == j) continue;
        const int jj = j % 3;
        if (Rinf[jj]) continue;
        Box pL = bL;
        if (i < 3)
          pL.lo[ii] = pL.hi[ii];
        else
          pL.hi[ii] = pL.lo[ii];
        Box pR = bR;
        if (j < 3)
          pR.lo[jj] = pR.hi[jj];
        else
          pR.hi[jj] = pR.lo[jj];
        if (ii == jj) {
          if (i < 3) {
            if (pL.hi[ii] >= pR.lo[ii]) continue;
            lr += parF(pL, pR, ii);
          } else {
            if (pL.lo[ii] <= pR.hi[ii]) continue;
            lr += parF(pL, pR, ii);
          }
        } else {
          if (j < 3) {
            if (pL.hi[jj] <= pR.hi[jj]) continue;
            pL.lo[jj] = fmax_(pL.lo[jj], pR.hi[jj]);
          } else {
            if (pL.lo[jj] >= pR.lo[jj]) continue;
            pL.hi[jj] = fmin_(pL.hi[jj], pR.lo[jj]);
          }
          if (i < 3) {
            pR.lo[ii] = fmax_(pR.lo[ii], pL.hi[ii]);
            lr += perpF(pL, pR, ii, jj);
          } else {
            pR.hi[ii] = fmin_(pR.hi[ii], pL.lo[ii]);
            lr += perpF(pL, pR, ii, jj);
          }
        }
      }
    }
    const float pinv = (float)(1.0f / parent.area());
    lr += overlap(bL, bR);
    lr *= pinv;
    // computeArea() is double, so the product and the subtraction are in double.
    jl = (float)(bL.area() * (double)pinv - (double)lr);
    jr = (float)(bR.area() * (double)pinv - (double)lr);
  }
};

struct Cost2 {
  float c[2];
};

// Per record (c0, c1, lCost, rCost) of the last subtreeCost pass, when a caller asks.
float* g_cost_out = nullptr;

// DynBVH.cc:1594-1678 with opaque materials (canAttenuateShadows() == false).
Cost2 subtreeCost(std::vector<Node32>& nodes, const TTP& ttp, int id) {
  const float ISEC = 3, TRAV = 1;
  Node32& node = nodes[id];
  if (node.isLeaf()) {
    Cost2 r;
    r.c[0] = node.children * ISEC + TRAV;
    r.c[1] = 0;
    if (g_cost_out) {
      float* o = g_cost_out + 4 * (size_t)id;
      o[0] = r.c[0];
      o[1] = r.c[1];
      o[2] = o[3] = 0.f;
    }
    return r;
  }
  float Plr, Pjl, Pjr;
  ttp.probs(nodes[id].bounds, nodes[node.child].bounds, nodes[node.child + 1].bounds, Plr, Pjl,
            Pjr);
  const float Pl = Pjl + Plr;
  const float Pr = Pjr + Plr;
  Cost2 Cl = subtreeCost(nodes, ttp, node.child);
  Cost2 Cr = subtreeCost(nodes, ttp, node.child + 1);
  const float Pe = 1 - Pjr - Pjl - Plr;
  const float lCost = Pl * Cl.c[0] + (Pjr + Plr * Cl.c[1]) * Cr.c[0] + (1 + Pl * Cl.c[1] + (1 - Pl)) * TRAV;
  const float rCost = Pr * Cr.c[0] + (Pjl + Plr * Cr.c[1]) * Cl.c[0] + (1 + Pr * Cr.c[1] + (1 - Pr)) * TRAV;
  Cost2 cost;
  cost.c[0] = fmin_(lCost, rCost);
  float c1 = Pjl * Cl.c[1] + Pjr * Cr.c[1] + Plr * Cl.c[1] * Cr.c[1] + Pe;
  c1 = c1 <= 0.0f ? 0.0f : (c1 >= 1.0f ? 1.0f : c1);  // Clamp (Core/Math/MiscMath.h:82-85)
  cost.c[1] = c1;
  if (lCost < rCost)
    node.bits |= 1;
  else
    node.bits &= ~1;
  if (g_cost_out) {
    float* o = g_cost_out + 4 * (size_t)id;
    o[0] = cost.c[0];
    o[1] = cost.c[1];
    o[2] = lCost;
    o[3] = rCost;
  }
  return cost;
}

// ---------------------------------------------------------------------------
// Triangle tests
// ---------------------------------------------------------------------------
// RayPacketData, reduced to the fields the path touches (Interface/RayPacket.h:58-130).
// state: 0 = not hit (hitMatl == 0), 1 = hit (hitMatl = mesh material), -1 = masked
// (hitMatl == ~0, RayPacket.h:488-502).
struct alignas(16) Packet {
  float org[3][64];
  float dir[3][64];
  float inv[3][64];
  float minT[64];
  int state[64];
  int prim[64];
  float u[64], v[64];
  bool anyHit;
};

inline bool masked(const Packet& p, int i) { return p.state[i] == -1; }
inline bool wasHit(const Packet& p, int i) { return p.state[i] != 0; }

// RayPacket::hit (RayPacket.h:531-542)
inline bool packetHit(Packet& p, int i, float t, int prim) {
  if (t > kTEps && t < p.minT[i]) {
    p.minT[i] = t;
    p.state[i] = 1;
    p.prim[i] = prim;
    return true;
  }
  return false;
}

// Per-triangle Wald data (WaldTriangle.cc:64-127)
struct Wald {
  float n_u, n_v, n_d;
  int k;
  float b_nu, b_nv, b_d, n_k, c_nu, c_nv, c_d;
};

Wald waldSetup(const float* P1, const float* P2, const float* P3) {
  Vec p1, p2, p3;
  for (int i = 0; i < 3; ++i) {
    p1.v[i] = P1[i];
    p2.v[i] = P2[i];
    p3.v[i] = P3[i];
  }
  Vec nrm = vcross(vsub(p2, p1), vsub(p3, p1));
  // Vector::normalize (Core/Geometry/Vector.h:322-331)
  float l = sqrtf(nrm.v[0] * nrm.v[0] + nrm.v[1] * nrm.v[1] + nrm.v[2] * nrm.v[2]);
  float scale = 1 / l;
  nrm.v[0] *= scale;
  nrm.v[1] *= scale;
  nrm.v[2] *= scale;
  unsigned n = 0, u, v;
  if (fabsf(nrm.v[1]) > fabsf(nrm.v[n])) n = 1;
  if (fabsf(nrm.v[2]) > fabsf(nrm.v[n])) n = 2;
  switch (n) {
    case 0: u = 1; v = 2; break;
    case 1: u = 2; v = 0; break;
    default: u = 0; v = 1; break;
  }
  Wald w;
  w.k = (int)n;
  float inv_nk = 1 / nrm.v[n];
  w.n_u = nrm.v[u] * inv_nk;
  w.n_v = nrm.v[v] * inv_nk;
  w.n_d = p1.v[u] * w.n_u + p1.v[v] * w.n_v + p1.v[n];
  float s;
  w.c_nu = +p2.v[v] - p1.v[v];
  w.c_nv = -p2.v[u] + p1.v[u];
  w.c_d = -(p1.v[u] * w.c_nu + p1.v[v] * w.c_nv);
  s = 1.f / (p3.v[u] * w.c_nu + p3.v[v] * w.c_nv + w.c_d);
  w.c_nu *= s;
  w.c_nv *= s;
  w.c_d *= s;
  w.b_nu = +p3.v[v] - p1.v[v];
  w.b_nv = -p3.v[u] + p1.v[u];
  w.b_d = -(p1.v[u] * w.b_nu + p1.v[v] * w.b_nv);
  s = 1.f / (p2.v[u] * w.b_nu + p2.v[v] * w.b_nv + w.b_d);
  w.b_nu *= s;
  w.b_nv *= s;
  w.b_d *= s;
  w.n_k = nrm.v[n];
  return w;
}

struct Scene {
  Mesh mesh;
  const Node32* nodes;
  const int32_t* ids;
  int prim_kind;  // 0 = Kensler-Shirley, 1 = Wald
  std::vector<Wald> wald;
};

// KenslerShirleyTriangle::intersect, scalar branch (KenslerShirleyTriangle.cc:336-377)
void intersectKS(const Scene& sc, int tri, Packet& p, int b, int e) {
  Vec p0, p1, p2;
  for (int i = 0; i < 3; ++i) {
    p0.v[i] = sc.mesh.vtx(tri, 0)[i];
    p1.v[i] = sc.mesh.vtx(tri, 1)[i];
    p2.v[i] = sc.mesh.vtx(tri, 2)[i];
  }
  const Vec edge0 = vsub(p1, p0);
  const Vec edge1 = vsub(p0, p2);
  const Vec normal = vcross(edge0, edge1);
  for (int r = b; r < e; ++r) {
    Vec o, d;
    for (int i = 0; i < 3; ++i) {
      o.v[i] = p.org[i][r];
      d.v[i] = p.dir[i][r];
    }
    const float oldt = p.minT[r];
    const float rcp = 1.0f / vdot(normal, d);
    const Vec edge2 = vsub(p0, o);
    const float toverd = vdot(normal, edge2) * rcp;
    if (toverd > oldt - kTEps || toverd < kTEps) continue;
    const Vec interm = vcross(edge2, d);
    const float uoverd = vdot(interm, edge1) * rcp;
    if (uoverd < 0.0f) continue;
    const float voverd = vdot(interm, edge0) * rcp;
    if (uoverd + voverd > 1.0f || voverd < 0.0f) continue;
    if (packetHit(p, r, toverd, tri)) {
      p.u[r] = uoverd;
      p.v[r] = voverd;
    }
  }
}

// WaldTriangle::intersect, scalar branch (WaldTriangle.cc:489-555), per-ray origin.
// (The ConstantOrigin shortcut computes the same f0 when all origins are equal.)
void intersectWald(const Scene& sc, int tri, Packet& p, int b, int e) {
  const Wald& w = sc.wald[tri];
  const int axis = w.k;
  const int ku = (w.k == 2) ? 0 : w.k + 1;
  const int kv = (w.k == 0) ? 2 : w.k - 1;
  for (int i = b; i < e; ++i) {
    const float nd0 = w.n_u * p.dir[ku][i] + w.n_v * p.dir[kv][i] + p.dir[axis][i];
    const float nd = 1.f / nd0;
    const float org_k = p.org[axis][i], org_ku = p.org[ku][i], org_kv = p.org[kv][i];
    const float f0 = w.n_d - (org_k + w.n_u * org_ku + w.n_v * org_kv);
    const float f = f0 * nd;
    if (f < kTEps || f > p.minT[i]) continue;
    const float hu = org_ku + f * p.dir[ku][i];
    const float hv = org_kv + f * p.dir[kv][i];
    const float lambda = w.b_d + hu * w.b_nu + hv * w.b_nv;
    if (lambda < 0.f) continue;
    const float mue = w.c_d + hu * w.c_nu + hv * w.c_nv;
    if (mue < 0.f || mue + lambda > 1.f) continue;
    if (packetHit(p, i, f, tri)) {
      p.u[i] = lambda;
      p.v[i] = mue;
    }
  }
}

inline void intersectPrim(const Scene& sc, int tri, Packet& p, int b, int e) {
  if (sc.prim_kind == 1)
    intersectWald(sc, tri, p, b, e);
  else
    intersectKS(sc, tri, p, b, e);
}

// ---------------------------------------------------------------------------
// Packet traversal (DynBVH.cc:101-404, scalar firstIntersects 655-725, lastIntersects 882-905)
// ---------------------------------------------------------------------------
struct IA {
  float min_rcp[3], max_rcp[3], min_org[3], max_org[3];
};

struct Counters {
  int64_t box_tests = 0;  // per-ray slab evaluations (single-ray traversal)
  int64_t tri_tests = 0;  // per-ray triangle tests
};

inline bool slab(const Box& box, const Packet& p, int i) {
  float tmin = kTEps;
  float tmax = p.minT[i];
  for (int c = 0; c < 3; ++c) {
    float t0 = (box.lo[c] - p.org[c][i]) * p.inv[c][i];
    float t1 = (box.hi[c] - p.org[c][i]) * p.inv[c][i];
    float near_ = (t0 < t1) ? t0 : t1;
    float far_ = (t0 < t1) ? t1 : t0;
    tmin = (tmin < near_) ? near_ : tmin;
    tmax = (far_ < tmax) ? far_ : tmax;
  }
  return tmin <= tmax;
}

int firstIntersects(const Box& box, const Packet& p, int rb, int re, const IA& ia,
                    Counters* cnt) {
  int rayBegin = rb;
  for (; rayBegin < re; rayBegin++) {
    if (masked(p, rayBegin)) continue;
    if (p.anyHit && wasHit(p, rayBegin)) continue;
    break;
  }
  for (int i = rayBegin; i < re; ++i) {
    if (masked(p, i)) continue;
    if (p.anyHit && wasHit(p, i)) continue;
    if (cnt) cnt->box_tests++;
    if (slab(box, p, i)) return i;
    if (i == rayBegin) {
      float tmin_f = (float)1e-5;
      float tmax_f = FLT_MAX;
      for (int axis = 0; axis < 3; ++axis) {
        float a = box.lo[axis] - ia.max_org[axis];
        float b = box.hi[axis] - ia.min_org[axis];
        float ar0 = a * ia.min_rcp[axis];
        float ar1 = a * ia.max_rcp[axis];
        float br0 = b * ia.min_rcp[axis];
        float br1 = b * ia.max_rcp[axis];
        a = (br0 < br1) ? br0 : br1;
        a = (a < ar1) ? a : ar1;
        a = (a < ar0) ? a : ar0;
        b = (br0 < br1) ? br1 : br0;
        b = (b < ar1) ? ar1 : b;
        b = (b < ar0) ? ar0 : b;
        tmin_f = (tmin_f < a) ? a : tmin_f;
        tmax_f = (tmax_f > b) ? b : tmax_f;
      }
      if (tmin_f > tmax_f) return re;
    }
  }
  return re;
}

int lastIntersects(const Box& box, const Packet& p, int rb, int re) {
  int rayEnd = re;
  for (; rb < rayEnd; rayEnd--) {
    if (masked(p, rayEnd - 1)) continue;
    if (p.anyHit && wasHit(p, rayEnd - 1)) continue;
    break;
  }
  for (int i = rayEnd - 1; i > rb; --i) {
    if (p.anyHit && wasHit(p, i)) continue;
    if (slab(box, p, i)) return i;
  }
  return rb;
}

// ---------------------------------------------------------------------------
// SSE build (MANTA_SSE): 4-wide groups of aligned rays. _mm_min_ps(a, b) is a < b ? a : b
// and _mm_max_ps(a, b) is a > b ? a : b (the second operand when either is NaN), which the
// intrinsics below give exactly.
// ---------------------------------------------------------------------------
// oneOver (Core/Math/SSEDefs.h:193-197): rcpps and one Newton step, 2r - r*r*v.
inline __m128 oneOver4(__m128 v) {
  const __m128 r = _mm_rcp_ps(v);
  return _mm_sub_ps(_mm_add_ps(r, r), _mm_mul_ps(_mm_mul_ps(r, r), v));
}

// min4f / max4f (SSEDefs.h:269-299)
inline float min4f(__m128 t) {
  const __m128 c = _mm_min_ps(_mm_unpacklo_ps(t, t), _mm_unpackhi_ps(t, t));
  return _mm_cvtss_f32(_mm_min_ss(c, _mm_movehl_ps(c, c)));
}
inline float max4f(__m128 t) {
  const __m128 c = _mm_max_ps(_mm_unpacklo_ps(t, t), _mm_unpackhi_ps(t, t));
  return _mm_cvtss_f32(_mm_max_ss(c, _mm_movehl_ps(c, c)));
}

inline float lane(__m128 v, int k) {
  alignas(16) float a[4];
  _mm_store_ps(a, v);
  return a[k];
}

// KenslerShirleyTriangle::intersect, SSE branch (KenslerShirleyTriangle.cc:134-332): the
// aligned groups of [b, e) four at a time, the rest with the scalar test.
void intersectKS_SSE(const Scene& sc, int tri, Packet& p, int b, int e) {
  const int sb = (b + 3) & ~3, se = e & ~3;
  if (sb >= se) {
    intersectKS(sc, tri, p, b, e);
    return;
  }
  const __m128 eps4 = _mm_set1_ps(kTEps);
  const __m128 zero = _mm_setzero_ps();
  const float* P0 = sc.mesh.vtx(tri, 0);
  const float* P1 = sc.mesh.vtx(tri, 1);
  const float* P2 = sc.mesh.vtx(tri, 2);
  const __m128 p0x = _mm_set1_ps(P0[0]), p0y = _mm_set1_ps(P0[1]), p0z = _mm_set1_ps(P0[2]);
  const __m128 e0x = _mm_sub_ps(_mm_set1_ps(P1[0]), p0x);
  const __m128 e0y = _mm_sub_ps(_mm_set1_ps(P1[1]), p0y);
  const __m128 e0z = _mm_sub_ps(_mm_set1_ps(P1[2]), p0z);
  const __m128 e1x = _mm_sub_ps(p0x, _mm_set1_ps(P2[0]));
  const __m128 e1y = _mm_sub_ps(p0y, _mm_set1_ps(P2[1]));
  const __m128 e1z = _mm_sub_ps(p0z, _mm_set1_ps(P2[2]));
  // cross4 (:127-132): x = y1 z2 - z1 y2, y = z1 x2 - x1 z2, z = x1 y2 - y1 x2
  const __m128 nx = _mm_sub_ps(_mm_mul_ps(e0y, e1z), _mm_mul_ps(e0z, e1y));
  const __m128 ny = _mm_sub_ps(_mm_mul_ps(e0z, e1x), _mm_mul_ps(e0x, e1z));
  const __m128 nz = _mm_sub_ps(_mm_mul_ps(e0x, e1y), _mm_mul_ps(e0y, e1x));
  for (int r = sb; r < se; r += 4) {
    const __m128 ox = _mm_load_ps(&p.org[0][r]), oy = _mm_load_ps(&p.org[1][r]),
                 oz = _mm_load_ps(&p.org[2][r]);
    const __m128 oldt = _mm_load_ps(&p.minT[r]);
    const __m128 dx = _mm_load_ps(&p.dir[0][r]), dy = _mm_load_ps(&p.dir[1][r]),
                 dz = _mm_load_ps(&p.dir[2][r]);
    // dot4(o, v) (SSEDefs.h:148-154) = (vx ox + vy oy) + vz oz
    const __m128 det = _mm_add_ps(_mm_add_ps(_mm_mul_ps(dx, nx), _mm_mul_ps(dy, ny)),
                                  _mm_mul_ps(dz, nz));
    const __m128 rcp = oneOver4(det);
    const __m128 g0 = _mm_sub_ps(p0x, ox), g1 = _mm_sub_ps(p0y, oy), g2 = _mm_sub_ps(p0z, oz);
    const __m128 t = _mm_add_ps(_mm_add_ps(_mm_mul_ps(g0, nx), _mm_mul_ps(g1, ny)),
                                _mm_mul_ps(g2, nz));
    const __m128 toverd = _mm_mul_ps(rcp, t);
    __m128 mask = _mm_and_ps(_mm_cmpgt_ps(toverd, eps4),
                             _mm_cmplt_ps(toverd, _mm_sub_ps(oldt, eps4)));
    if (_mm_movemask_ps(mask) == 0) continue;
    const __m128 ix = _mm_sub_ps(_mm_mul_ps(g1, dz), _mm_mul_ps(g2, dy));
    const __m128 iy = _mm_sub_ps(_mm_mul_ps(g2, dx), _mm_mul_ps(g0, dz));
    const __m128 iz = _mm_sub_ps(_mm_mul_ps(g0, dy), _mm_mul_ps(g1, dx));
    const __m128 u = _mm_add_ps(_mm_add_ps(_mm_mul_ps(e1x, ix), _mm_mul_ps(e1y, iy)),
                                _mm_mul_ps(e1z, iz));
    const __m128 uoverd = _mm_mul_ps(rcp, u);
    mask = _mm_and_ps(mask, _mm_cmpge_ps(uoverd, zero));
    if (_mm_movemask_ps(mask) == 0) continue;
    const __m128 v = _mm_add_ps(_mm_add_ps(_mm_mul_ps(e0x, ix), _mm_mul_ps(e0y, iy)),
                                _mm_mul_ps(e0z, iz));
    // u + v <= 1 tested as (u' + v') det <= det^2 on the undivided values (:250)
    const __m128 uvmask = _mm_cmple_ps(_mm_mul_ps(_mm_add_ps(u, v), det), _mm_mul_ps(det, det));
    const __m128 voverd = _mm_mul_ps(rcp, v);
    mask = _mm_and_ps(mask, uvmask);
    mask = _mm_and_ps(mask, _mm_cmpge_ps(voverd, zero));
    const int m = _mm_movemask_ps(mask);
    if (m == 0) continue;
    // hitWithoutTminCheck -> hitFinal (RayPacket.h:545-605): t < minT as well; the
    // scratchpad u, v are stored under the triangle's mask.
    const int hit = _mm_movemask_ps(_mm_and_ps(mask, _mm_cmplt_ps(toverd, _mm_load_ps(&p.minT[r]))));
    for (int k = 0; k < 4; ++k) {
      if (hit & (1 << k)) {
        p.minT[r + k] = lane(toverd, k);
        p.state[r + k] = 1;
        p.prim[r + k] = tri;
      }
      if (m & (1 << k)) {
        p.u[r + k] = lane(uoverd, k);
        p.v[r + k] = lane(voverd, k);
      }
    }
  }
  if (b < sb) intersectKS(sc, tri, p, b, sb);
  if (se < e) intersectKS(sc, tri, p, se, e);
}

// The one-ray SSE slab test of firstIntersects (DynBVH.cc:431-456, 458-483): _ss forms.
inline bool slabSS(const Box& box, const Packet& p, int i) {
  auto mn = [](float a, float b) { return a < b ? a : b; };  // _mm_min_ss
  auto mx = [](float a, float b) { return a > b ? a : b; };  // _mm_max_ss
  float tmin = kTEps, tmax = p.minT[i];
  for (int c = 0; c < 3; ++c) {
    const float t0 = (box.lo[c] - p.org[c][i]) * p.inv[c][i];
    const float t1 = (box.hi[c] - p.org[c][i]) * p.inv[c][i];
    tmin = mx(mn(t0, t1), tmin);
    tmax = mn(mx(t0, t1), tmax);
  }
  return tmin <= tmax;
}

// One aligned group of four rays (DynBVH.cc:583-604 / 817-845): the rays whose slab test
// passes, as a movemask.
inline int slab4(const Box& box, const Packet& p, int i) {
  const __m128 x0 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(box.lo[0]), _mm_load_ps(&p.org[0][i])),
                               _mm_load_ps(&p.inv[0][i]));
  const __m128 x1 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(box.hi[0]), _mm_load_ps(&p.org[0][i])),
                               _mm_load_ps(&p.inv[0][i]));
  const __m128 y0 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(box.lo[1]), _mm_load_ps(&p.org[1][i])),
                               _mm_load_ps(&p.inv[1][i]));
  const __m128 y1 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(box.hi[1]), _mm_load_ps(&p.org[1][i])),
                               _mm_load_ps(&p.inv[1][i]));
  const __m128 z0 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(box.lo[2]), _mm_load_ps(&p.org[2][i])),
                               _mm_load_ps(&p.inv[2][i]));
  const __m128 z1 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(box.hi[2]), _mm_load_ps(&p.org[2][i])),
                               _mm_load_ps(&p.inv[2][i]));
  const __m128 mxmn = _mm_max_ps(_mm_min_ps(x0, x1),
                                 _mm_max_ps(_mm_min_ps(y0, y1),
                                            _mm_max_ps(_mm_min_ps(z0, z1), _mm_set1_ps(kTEps))));
  const __m128 mnmx = _mm_min_ps(_mm_max_ps(x0, x1),
                                 _mm_min_ps(_mm_max_ps(y0, y1),
                                            _mm_min_ps(_mm_max_ps(z0, z1),
                                                       _mm_load_ps(&p.minT[i]))));
  return _mm_movemask_ps(_mm_cmple_ps(mxmn, mnmx));
}

inline bool groupMasked(const Packet& p, int i) {
  return masked(p, i) && masked(p, i + 1) && masked(p, i + 2) && masked(p, i + 3);
}
// RayPacket::wereNotHitSSE: hitMatl == 0 (masked rays count as hit)
inline int notHit4(const Packet& p, int i) {
  int m = 0;
  for (int k = 0; k < 4; ++k) m |= (p.state[i + k] == 0) << k;
  return m;
}

// DynBVH::firstIntersects, SSE branch (DynBVH.cc:407-629)
int firstIntersectsSSE(const Box& box, const Packet& p, int rb, int re, const IA& ia) {
  int rayBegin = rb;
  const int rayEnd = re;
  for (; rayBegin < rayEnd; rayBegin++) {
    if (masked(p, rayBegin)) continue;
    if (p.anyHit && wasHit(p, rayBegin)) continue;
    break;
  }
  const int b = (rayBegin + 3) & ~3, e = rayEnd & ~3;
  if (b >= e) {
    for (int i = rayBegin; i < rayEnd; i++) {
      if (masked(p, i)) continue;
      if (p.anyHit && wasHit(p, i)) continue;
      if (slabSS(box, p, i)) return i;
    }
    return rayEnd;
  }
  int i = rayBegin;
  if (slabSS(box, p, i)) return i;
  {  // frustum miss (the SSE branch starts from T_EPSILON)
    float tmin_f = kTEps, tmax_f = FLT_MAX;
    for (int axis = 0; axis < 3; ++axis) {
      float a = box.lo[axis] - ia.max_org[axis];
      float bb = box.hi[axis] - ia.min_org[axis];
      const float ar0 = a * ia.min_rcp[axis], ar1 = a * ia.max_rcp[axis];
      const float br0 = bb * ia.min_rcp[axis], br1 = bb * ia.max_rcp[axis];
      a = (br0 < br1) ? br0 : br1;
      a = (a < ar1) ? a : ar1;
      a = (a < ar0) ? a : ar0;
      bb = (br0 < br1) ? br1 : br0;
      bb = (bb < ar1) ? ar1 : bb;
      bb = (bb < ar0) ? ar0 : bb;
      tmin_f = (tmin_f < a) ? a : tmin_f;
      tmax_f = (tmax_f > bb) ? bb : tmax_f;
    }
    if (tmin_f > tmax_f) return rayEnd;
  }
  if (i < b) i++;
  for (; i < b; i++) {
    if (masked(p, i)) continue;
    if (p.anyHit && wasHit(p, i)) continue;
    if (slab(box, p, i)) return i;
  }
  for (; i < e; i += 4) {
    if (groupMasked(p, i)) continue;
    int valid = 15;
    if (p.anyHit) {
      valid = notHit4(p, i);
      if (valid == 0) continue;
    }
    const int mask = valid & slab4(box, p, i);
    if (mask) {
      for (int k = 0; k < 4; ++k)
        if ((mask & (1 << k)) && !masked(p, i + k)) return i + k;
    }
  }
  for (; i < rayEnd; i++) {
    if (masked(p, i)) continue;
    if (p.anyHit && wasHit(p, i)) continue;
    if (slab(box, p, i)) return i;
  }
  return rayEnd;
}

// DynBVH::lastIntersects, SSE branch (DynBVH.cc:728-879)
int lastIntersectsSSE(const Box& box, const Packet& p, int rb, int re) {
  const int rayBegin = rb;
  int rayEnd = re;
  for (; rayBegin < rayEnd; rayEnd--) {
    if (masked(p, rayEnd - 1)) continue;
    if (p.anyHit && wasHit(p, rayEnd - 1)) continue;
    break;
  }
  const int last_simd = rayEnd & ~3, first_simd = (rayBegin + 3) & ~3;
  if (first_simd >= last_simd) {
    for (int i = rayEnd - 1; i > rayBegin; i--) {
      if (masked(p, i)) continue;
      if (p.anyHit && wasHit(p, i)) continue;
      if (slab(box, p, i)) return i;
    }
    return rayBegin;
  }
  int i = rayEnd;
  while (i > last_simd) {
    i--;
    if (masked(p, i)) continue;
    if (p.anyHit && wasHit(p, i)) continue;
    if (slab(box, p, i)) return i;
  }
  while (i > first_simd) {
    i -= 4;
    if (groupMasked(p, i)) continue;
    int valid = 15;
    if (p.anyHit) {
      valid = notHit4(p, i);
      if (valid == 0) continue;
    }
    const int mask = valid & slab4(box, p, i);
    if (mask) {
      for (int k = 3; k >= 0; --k)
        if ((mask & (1 << k)) && !masked(p, i + k)) return i + k;
    }
  }
  while (i > rayBegin) {
    i--;
    if (p.anyHit && wasHit(p, i)) continue;
    if (slab(box, p, i)) return i;
  }
  return rayBegin;
}

template <bool SSE>
void intersectNode(const Scene& sc, int nodeID, Packet& p, int rb, int re, const IA& ia,
                   Counters* cnt) {
  const Node32& node = sc.nodes[nodeID];
  int firstActive = SSE ? firstIntersectsSSE(node.bounds, p, rb, re, ia)
                        : firstIntersects(node.bounds, p, rb, re, ia, cnt);
  if (firstActive == re) return;
  if (node.isLeaf()) {
    int lastActive = SSE ? lastIntersectsSSE(node.bounds, p, rb, re)
                         : lastIntersects(node.bounds, p, rb, re);
    int sb = firstActive, se = lastActive + 1;
    for (int k = 0; k < node.children; ++k) {
      const int object_id = sc.ids[node.child + k];
      if (cnt) cnt->tri_tests += se - sb;
      if (p.anyHit) {
        // DynBVH.cc:315-375; every material is opaque, so attenuateShadows turns each
        // newly hit ray black and the ray's minT is overridden with T_EPSILON.
        if (SSE)
          intersectKS_SSE(sc, object_id, p, sb, se);
        else
          intersectPrim(sc, object_id, p, sb, se);
        bool terminated = false;
        for (int i = sb; i < se; ++i) {
          if (!wasHit(p, i) || masked(p, i)) continue;
          terminated = true;
          p.minT[i] = kTEps;
        }
        if (terminated) {
          int start, end;
          for (start = sb; start < se; ++start)
            if (!wasHit(p, start) && !masked(p, start)) break;
          for (end = se - 1; end > start; --end)
            if (!wasHit(p, end) && !masked(p, end)) break;
          sb = start;
          se = end + 1;
          if (end < start) break;
        }
      } else if (SSE) {
        intersectKS_SSE(sc, object_id, p, sb, se);
      } else {
        intersectPrim(sc, object_id, p, sb, se);
      }
    }
  } else {
    int front;
    if (p.anyHit)
      front = node.leftCheaper() ? 0 : 1;
    else
      front = p.dir[node.axis][firstActive] > 0 ? 0 : 1;
    intersectNode<SSE>(sc, node.child + front, p, firstActive, re, ia, cnt);
    intersectNode<SSE>(sc, node.child + 1 - front, p, firstActive, re, ia, cnt);
  }
}

inline void iaScalar(IA& ia, const Packet& p, int r) {
  for (int a = 0; a < 3; ++a) {
    const float rc = p.inv[a][r], o = p.org[a][r];
    ia.min_rcp[a] = (ia.min_rcp[a] < rc) ? ia.min_rcp[a] : rc;
    ia.max_rcp[a] = (ia.max_rcp[a] < rc) ? rc : ia.max_rcp[a];
    ia.min_org[a] = (ia.min_org[a] < o) ? ia.min_org[a] : o;
    ia.max_org[a] = (ia.max_org[a] < o) ? o : ia.max_org[a];
  }
}

// DynBVH::intersect (DynBVH.cc:101-272), scalar branch
void intersectPacket(const Scene& sc, Packet& p, int n, Counters* cnt) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) p.inv[j][i] = (float)1 / p.dir[j][i];
  IA ia;
  for (int a = 0; a < 3; ++a) {
    ia.min_rcp[a] = FLT_MAX;
    ia.max_rcp[a] = -FLT_MAX;
    ia.min_org[a] = FLT_MAX;
    ia.max_org[a] = -FLT_MAX;
  }
  for (int r = 0; r < n; ++r) {
    if (masked(p, r)) continue;
    iaScalar(ia, p, r);
  }
  intersectNode<false>(sc, 0, p, 0, n, ia, cnt);
}

// DynBVH::intersect, SSE branch: inverse directions by oneOver on the aligned groups
// (RayPacket.cc:80-104), interval data by 4-wide min/max over the aligned groups, masked
// rays included (DynBVH.cc:130-236), then the SSE recursion.
void intersectPacketSSE(const Scene& sc, Packet& p, int n) {
  const int e = n & ~3;
  if (e <= 0) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < 3; ++j) p.inv[j][i] = (float)1 / p.dir[j][i];
  } else {
    for (int i = 0; i < e; i += 4)
      for (int j = 0; j < 3; ++j)
        _mm_store_ps(&p.inv[j][i], oneOver4(_mm_load_ps(&p.dir[j][i])));
    for (int i = e; i < n; ++i)
      for (int j = 0; j < 3; ++j) p.inv[j][i] = (float)1 / p.dir[j][i];
  }
  IA ia;
  for (int a = 0; a < 3; ++a) {
    ia.min_rcp[a] = FLT_MAX;
    ia.max_rcp[a] = -FLT_MAX;
    ia.min_org[a] = FLT_MAX;
    ia.max_org[a] = -FLT_MAX;
  }
  if (e <= 0) {
    for (int r = 0; r < n; ++r)
      if (!masked(p, r)) iaScalar(ia, p, r);
  } else {
    for (int a = 0; a < 3; ++a) {
      __m128 mnr = _mm_set1_ps(ia.min_rcp[a]), mxr = _mm_set1_ps(ia.max_rcp[a]);
      __m128 mno = _mm_set1_ps(ia.min_org[a]), mxo = _mm_set1_ps(ia.max_org[a]);
      for (int r = 0; r < e; r += 4) {
        const __m128 rc = _mm_load_ps(&p.inv[a][r]), o = _mm_load_ps(&p.org[a][r]);
        mnr = _mm_min_ps(mnr, rc);
        mxr = _mm_max_ps(mxr, rc);
        mno = _mm_min_ps(mno, o);
        mxo = _mm_max_ps(mxo, o);
      }
      ia.min_rcp[a] = min4f(mnr);
      ia.max_rcp[a] = max4f(mxr);
      ia.min_org[a] = min4f(mno);
      ia.max_org[a] = max4f(mxo);
    }
    for (int r = e; r < n; ++r)
      if (!masked(p, r)) iaScalar(ia, p, r);
  }
  intersectNode<true>(sc, 0, p, 0, n, ia, nullptr);
}

struct RayIO {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  const uint8_t* masked;  // may be null
  int32_t* prim;
  float *t, *u, *v;
  uint8_t* occluded;  // may be null
};

void tracePacketRange(const Scene& sc, const RayIO& io, int64_t first, int n, bool anyHit,
                      Counters* cnt, bool sse = false) {
  Packet p;
  p.anyHit = anyHit;
  for (int i = 0; i < n; ++i) {
    const int64_t r = first + i;
    p.org[0][i] = io.ox[r];
    p.org[1][i] = io.oy[r];
    p.org[2][i] = io.oz[r];
    p.dir[0][i] = io.dx[r];
    p.dir[1][i] = io.dy[r];
    p.dir[2][i] = io.dz[r];
    p.prim[i] = -1;
    p.u[i] = p.v[i] = 0.f;
    if (io.masked && io.masked[r]) {
      p.state[i] = -1;          // RayPacket::maskRay
      p.minT[i] = -kMaxT;
    } else {
      p.state[i] = 0;
      p.minT[i] = io.tmax[r];
    }
  }
  if (sse)
    intersectPacketSSE(sc, p, n);
  else
    intersectPacket(sc, p, n, cnt);
  for (int i = 0; i < n; ++i) {
    const int64_t r = first + i;
    if (io.prim) io.prim[r] = (p.state[i] == 1) ? p.prim[i] : (p.state[i] == -1 ? -2 : -1);
    if (io.t) io.t[r] = p.minT[i];
    if (io.u) io.u[r] = p.u[i];
    if (io.v) io.v[r] = p.v[i];
    if (io.occluded) io.occluded[r] = wasHit(p, i) ? 1 : 0;
  }
}

template <class F>
void parallelFor(int64_t n, int nthreads, F f) {
  if (nthreads <= 1 || n <= 1) {
    for (int64_t i = 0; i < n; ++i) f(i, 0);
    return;
  }
  std::atomic<int64_t> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t]() {
      for (;;) {
        int64_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i, t);
      }
    });
  for (auto& x : th) x.join();
}

}  // namespace

// ---------------------------------------------------------------------------
// Refit with tree rotations and leaf splits: DynBVH::update with TREE_ROT set
// (Model/Groups/DynBVH.h:377), numProcs == 1: parallelUpdateBounds (DynBVH.cc:2012-2076)
// -> updateBounds<false> (DynBVH.cc:2078-2123) -> rotateNode (DynBVH.cc:3051-3372), which
// splits leaf children with the approximate SAH build (DynBVH.cc:1801-1857 with
// useApproximateBuild, partitionApproxSAH DynBVH.cc:2125-2275, partitionObjects
// DynBVH.cc:1369-1396, partition2Objs DynBVH.cc:2277-2308). Real is float
// (include/CMakeLists.txt:9); BBox::computeArea's double is narrowed where the
// reference assigns it to Real.
// ---------------------------------------------------------------------------
struct Updater {
  const Mesh& mesh;
  Node32* nodes;
  int32_t* ids;
  float* costs;
  uint32_t* ssize;
  int& next_free;
  uint32_t large_subtree;
  std::vector<Box> obj_bounds;
  std::vector<Vec> obj_centroids;

  Updater(const Mesh& m, Node32* n, int32_t* i, float* c, uint32_t* z, int& nf, uint32_t large)
      : mesh(m), nodes(n), ids(i), costs(c), ssize(z), next_free(nf), large_subtree(large) {}

  struct Partition {
    Box bounds_l, bounds_r;
    int axis;
    int split;
  };

  static int longestAxis(const Box& b) {  // BBox.h:123-129
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx > dy) return dx > dz ? 0 : 2;
    return dy > dz ? 1 : 2;
  }
  static int indexOfMax(float a, float b, float c) {  // Vector.h:382-393
    if (a > b) return a > c ? 0 : 2;
    return b > c ? 1 : 2;
  }

  // DynBVH.cc:2277-2308
  Partition partition2(int nodeID, int objBegin) {
    Partition p;
    p.bounds_l.reset();
    p.bounds_r.reset();
    const float nodeArea = (float)nodes[nodeID].bounds.area();
    const float leftArea = (float)obj_bounds[ids[objBegin]].area();
    const float rightArea = (float)obj_bounds[ids[objBegin + 1]].area();
    const float leafCost = kCIsec * 2;
    const float newCost = kCTrav + (leftArea * kCIsec + rightArea * kCIsec) / nodeArea;
    if (newCost < leafCost) {
      const Box& nb = nodes[nodeID].bounds;
      const int axis = indexOfMax(nb.hi[0] - nb.lo[0], nb.hi[1] - nb.lo[1], nb.hi[2] - nb.lo[2]);
      p.axis = axis;
      p.split = objBegin + 1;
      if (obj_bounds[ids[objBegin]].lo[axis] > obj_bounds[ids[objBegin + 1]].lo[axis])
        std::swap(ids[objBegin], ids[objBegin + 1]);
      p.bounds_l = obj_bounds[ids[objBegin]];
      p.bounds_r = obj_bounds[ids[objBegin + 1]];
      return p;
    }
    p.axis = -1;
    return p;
  }

  // DynBVH.cc:1369-1396
  int partitionObjects(int objBegin, int objEnd, int axis, float position) {
    int first = objBegin - 1, last = objEnd;
    for (;;) {
      for (++first; first < last; first++)
        if (obj_centroids[ids[first]].v[axis] >= position) break;
      for (--last; first < last; --last)
        if (obj_centroids[ids[last]].v[axis] < position) break;
      if (first < last)
        std::swap(ids[first], ids[last]);
      else
        return first;
    }
  }

  // DynBVH.cc:2125-2275 (LONGEST_AXIS is defined, so only the centroid box's longest axis)
  Partition partitionApprox(int nodeID, int objBegin, int objEnd) {
    Partition p;
    p.bounds_l.reset();
    p.bounds_r.reset();
    const int n = objEnd - objBegin;
    if (n == 1) {
      p.axis = -1;
      p.split = -1;
      return p;
    }
    if (n == 2) return partition2(nodeID, objBegin);
    Box cb;
    cb.reset();
    for (int i = objBegin; i < objEnd; ++i) cb.extendPoint(obj_centroids[ids[i]].v);
    const float inv_parent_area = (float)(1.0f / nodes[nodeID].bounds.area());
    float best_cost = n * kCIsec;
    int best_axis = -1;
    float best_pos = FLT_MAX;
    const int S = 16;  // BVH_num_samples (DynBVH.cc:34)
    float width[3], scale[3];
    for (int a = 0; a < 3; ++a) {
      width[a] = cb.hi[a] - cb.lo[a];
      scale[a] = ((float)S / width[a]) * .999f;
    }
    const int axis = longestAxis(cb);
    Box bins[S];
    int counts[S];
    for (int b = 0; b < S; ++b) {
      bins[b].reset();
      counts[b] = 0;
    }
    for (int i = objBegin; i < objEnd; ++i) {
      const Box& ob = obj_bounds[ids[i]];
      const float c = obj_centroids[ids[i]].v[axis];
      const float f = (c - cb.lo[axis]) * scale[axis];
      // int(f) as the reference's x86 build converts it (cvttss2si): NaN and values outside
      // the int range give INT_MIN, i.e. bin 0 after the clamp.
      int which = (f >= -2147483648.f && f < 2147483648.f) ? (int)f : INT_MIN;
      which = which < 0 ? 0 : (which > S - 1 ? S - 1 : which);
      counts[which]++;
      bins[which].extendBox(ob);
    }
    float left_areas[S], right_areas[S];
    int left_counts[S], right_counts[S];
    Box left_box, right_box;
    left_box.reset();
    right_box.reset();
    int num_left = 0, num_right = n;
    for (int i = 0; i < S; ++i) {
      num_left += counts[i];
      num_right -= counts[i];
      left_box.extendBox(bins[i]);
      left_counts[i] = num_left;
      right_counts[i] = num_right;
      left_areas[i] = (float)left_box.area();
    }
    for (int i = S - 1; i >= 0; --i) {
      right_areas[i] = (i == S - 1) ? FLT_MAX : (float)right_box.area();
      right_box.extendBox(bins[i]);
    }
    for (int i = 0; i < S - 1; ++i) {
      const float cost = (left_areas[i] * left_counts[i] + right_areas[i] * right_counts[i]) *
                             inv_parent_area * kCIsec + kCTrav;
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = axis;
        best_pos = cb.lo[axis] + width[axis] * (static_cast<float>(i + 1) / S);
      }
    }
    p.axis = best_axis;
    if (p.axis != -1) {
      int middle = partitionObjects(objBegin, objEnd, p.axis, best_pos);
      if (middle == objBegin || middle == objEnd) {
        if (n < 3) {
          p.axis = -1;
          p.split = 0;
          return p;
        }
        middle = objBegin + n / 2;
      }
      for (int i = objBegin; i < middle; ++i) p.bounds_l.extendBox(obj_bounds[ids[i]]);
      for (int i = middle; i < objEnd; ++i) p.bounds_r.extendBox(obj_bounds[ids[i]]);
      p.split = middle;
      return p;
    }
    p.split = 0;
    return p;
  }

  // DynBVH::build with useApproximateBuild (DynBVH.cc:1801-1857, TREE_ROT parts)
  void buildApprox(int nodeID, int objBegin, int objEnd, bool exists) {
    Partition p = partitionApprox(nodeID, objBegin, objEnd);
    Node32& node = nodes[nodeID];
    if (p.axis == -1) {
      if (exists) return;
      node.child = objBegin;
      node.children = (int16_t)(objEnd - objBegin);
      node.axis = 0;
      costs[nodeID] = node.children * kCIsec;
      ssize[nodeID] = 1;
      std::sort(ids + objBegin, ids + objEnd);
    } else {
      const int my_spot = next_free;
      next_free += 2;
      node.child = my_spot;
      node.children = 0;
      node.axis = (uint8_t)p.axis;
      nodes[my_spot + 0].bounds = p.bounds_l;
      nodes[my_spot + 1].bounds = p.bounds_r;
      buildApprox(my_spot + 0, objBegin, p.split, false);
      buildApprox(my_spot + 1, p.split, objEnd, false);
      ssize[nodeID] = 1 + ssize[my_spot + 0] + ssize[my_spot + 1];
    }
  }

  void swapRec(int a, int b) {
    std::swap(nodes[a], nodes[b]);
    std::swap(costs[a], costs[b]);
    std::swap(ssize[a], ssize[b]);
  }

  // DynBVH.cc:3051-3372
  void rotateNode(int nodeID) {
    const int leftID = nodes[nodeID].child;
    const int rightID = nodes[nodeID].child + 1;
    if (nodes[leftID].isLeaf() && nodes[leftID].children > 1)
      buildApprox(leftID, nodes[leftID].child, nodes[leftID].child + nodes[leftID].children, true);
    if (nodes[rightID].isLeaf() && nodes[rightID].children > 1)
      buildApprox(rightID, nodes[rightID].child, nodes[rightID].child + nodes[rightID].children,
                  true);
    const int LL = nodes[leftID].child, LR = nodes[leftID].child + 1;
    const int RL = nodes[rightID].child, RR = nodes[rightID].child + 1;
    const float nodeArea = (float)nodes[nodeID].bounds.area();
    const float leftArea = (float)nodes[leftID].bounds.area();
    const float rightArea = (float)nodes[rightID].bounds.area();
    float llA = 0.f, lrA = 0.f, rlA = 0.f, rrA = 0.f;
    Box boxA, boxB;
    boxA.reset();
    boxB.reset();
    float areaA = 0.f, areaB = 0.f;
    int best = 0;
    float bestCost = nodeArea * kCTrav + leftArea * costs[leftID] + rightArea * costs[rightID];
    auto unite = [&](int a, int b) {
      Box x = nodes[a].bounds;
      x.extendBox(nodes[b].bounds);
      return x;
    };
    if (!nodes[leftID].isLeaf()) {
      llA = (float)nodes[LL].bounds.area();
      lrA = (float)nodes[LR].bounds.area();
      const float partial = nodeArea * kCTrav + rightArea * costs[rightID] + lrA * costs[LR] +
                            llA * costs[LL];
      const Box b1 = unite(rightID, LR);
      const float a1 = (float)b1.area();
      const float c1 = a1 * kCTrav + partial;
      if (c1 < bestCost) { best = 1; bestCost = c1; boxA = b1; areaA = a1; }
      const Box b2 = unite(rightID, LL);
      const float a2 = (float)b2.area();
      const float c2 = a2 * kCTrav + partial;
      if (c2 < bestCost) { best = 2; bestCost = c2; boxA = b2; areaA = a2; }
    }
    if (!nodes[rightID].isLeaf()) {
      rlA = (float)nodes[RL].bounds.area();
      rrA = (float)nodes[RR].bounds.area();
      const float partial = nodeArea * kCTrav + leftArea * costs[leftID] + rrA * costs[RR] +
                            rlA * costs[RL];
      const Box b3 = unite(leftID, RR);
      const float a3 = (float)b3.area();
      const float c3 = a3 * kCTrav + partial;
      if (c3 < bestCost) { best = 3; bestCost = c3; boxA = b3; areaA = a3; }
      const Box b4 = unite(leftID, RL);
      const float a4 = (float)b4.area();
      const float c4 = a4 * kCTrav + partial;
      if (c4 < bestCost) { best = 4; bestCost = c4; boxA = b4; areaA = a4; }
      if (!nodes[leftID].isLeaf()) {
        const float partial2 = nodeArea * kCTrav + llA * costs[LL] + lrA * costs[LR] +
                               rlA * costs[RL] + rrA * costs[RR];
        const Box b5l = unite(LL, RL), b5r = unite(LR, RR);
        const float a5l = (float)b5l.area(), a5r = (float)b5r.area();
        const float c5 = (a5l + a5r) * kCTrav + partial2;
        if (c5 < bestCost) { best = 5; bestCost = c5; boxA = b5l; areaA = a5l; boxB = b5r; areaB = a5r; }
        const Box b6l = unite(LL, RR), b6r = unite(RL, LR);
        const float a6l = (float)b6l.area(), a6r = (float)b6r.area();
        const float c6 = (a6l + a6r) * kCTrav + partial2;
        if (c6 < bestCost) { best = 6; bestCost = c6; boxA = b6l; areaA = a6l; boxB = b6r; areaB = a6r; }
      }
    }
    costs[nodeID] = bestCost / nodeArea;
    if (best == 0) return;
    const float invA = 1 / areaA;
    switch (best) {
      case 1:
        nodes[leftID].bounds = boxA;
        costs[leftID] = kCTrav + ((lrA * costs[LR] + rightArea * costs[rightID]) * invA);
        swapRec(LL, rightID);
        ssize[leftID] = ssize[LL] + ssize[LR] + 1;
        break;
      case 2:
        nodes[leftID].bounds = boxA;
        costs[leftID] = kCTrav + ((llA * costs[LL] + rightArea * costs[rightID]) * invA);
        swapRec(LR, rightID);
        ssize[leftID] = ssize[LL] + ssize[LR] + 1;
        break;
      case 3:
        nodes[rightID].bounds = boxA;
        costs[rightID] = kCTrav + ((rrA * costs[RR] + leftArea * costs[leftID]) * invA);
        swapRec(RL, leftID);
        ssize[rightID] = ssize[RL] + ssize[RR] + 1;
        break;
      case 4:
        nodes[rightID].bounds = boxA;
        costs[rightID] = kCTrav + ((rlA * costs[RL] + leftArea * costs[leftID]) * invA);
        swapRec(RR, leftID);
        ssize[rightID] = ssize[RL] + ssize[RR] + 1;
        break;
      case 5:
        nodes[leftID].bounds = boxA;
        nodes[rightID].bounds = boxB;
        costs[leftID] = kCTrav + ((llA * costs[LL] + rlA * costs[RL]) * invA);
        costs[rightID] = kCTrav + ((lrA * costs[LR] + rrA * costs[RR]) / areaB);
        swapRec(LR, RL);
        ssize[leftID] = ssize[LL] + ssize[LR] + 1;
        ssize[rightID] = ssize[RL] + ssize[RR] + 1;
        break;
      default:  // 6
        nodes[leftID].bounds = boxA;
        nodes[rightID].bounds = boxB;
        costs[leftID] = kCTrav + ((llA * costs[LL] + rrA * costs[RR]) * invA);
        costs[rightID] = kCTrav + ((lrA * costs[LR] + rlA * costs[RL]) / areaB);
        swapRec(LR, RR);
        ssize[leftID] = ssize[LL] + ssize[LR] + 1;
        ssize[rightID] = ssize[RL] + ssize[RR] + 1;
        break;
    }
    // Step 3: guess split axes (overlap of the two children, DynBVH.cc:3305-3352)
    auto guess = [&](int parent, int a, int b) {
      const Box& x = nodes[a].bounds;
      const Box& y = nodes[b].bounds;
      float o[3];
      for (int k = 0; k < 3; ++k) o[k] = fmax_(x.lo[k], y.lo[k]) - fmin_(x.hi[k], y.hi[k]);
      const int ax = indexOfMax(o[0], o[1], o[2]);
      nodes[parent].axis = (uint8_t)ax;
      if (nodes[a].bounds.hi[ax] > nodes[b].bounds.hi[ax]) swapRec(a, b);
    };
    if (best <= 2 || best >= 5) guess(leftID, LL, LR);
    if (best >= 3) guess(rightID, RL, RR);
    guess(nodeID, leftID, rightID);
    if (nodes[nodeID].bits & 2) {
      auto mark = [&](int r) {
        if (ssize[r] > large_subtree)
          nodes[r].bits |= 2;
        else
          nodes[r].bits &= ~2;
      };
      mark(leftID);
      mark(rightID);
      if (best <= 2 || best >= 5) {
        mark(LL);
        mark(LR);
      }
      if (best >= 3) {
        mark(RL);
        mark(RR);
      }
    }
  }

  // DynBVH.cc:2078-2123, updateBounds<false> with TREE_ROT
  void updateBounds(int id) {
    Node32& node = nodes[id];
    if (node.isLeaf()) {
      node.bounds.reset();
      for (int i = 0; i < node.children; ++i) node.bounds.extendBox(obj_bounds[ids[node.child + i]]);
      node.axis = 0;
    } else {
      const int l = node.child, r = node.child + 1;
      updateBounds(l);
      updateBounds(r);
      node.bounds.reset();
      node.bounds.extendBox(nodes[l].bounds);
      node.bounds.extendBox(nodes[r].bounds);
      rotateNode(id);
    }
  }

  // updateBounds<true> (DynBVH.cc:2098-2108 with bottomSubtreesDone): recurses only into
  // children marked isLargeSubtree, then refits and rotates the node.
  void updateBoundsLarge(int id) {
    Node32& node = nodes[id];
    if (node.isLeaf()) {
      node.bounds.reset();
      for (int i = 0; i < node.children; ++i) node.bounds.extendBox(obj_bounds[ids[node.child + i]]);
      node.axis = 0;
      return;
    }
    const int l = node.child, r = node.child + 1;
    if (nodes[l].bits & 2) updateBoundsLarge(l);
    if (nodes[r].bits & 2) updateBoundsLarge(r);
    node.bounds.reset();
    node.bounds.extendBox(nodes[l].bounds);
    node.bounds.extendBox(nodes[r].bounds);
    rotateNode(id);
  }

  // seedSubtrees (DynBVH.cc:1998-2010): the maximal subtrees not marked isLargeSubtree,
  // in the order proc 0 appends them to the shared list.
  void seedSubtrees(int id, std::vector<int>& list) {
    if (nodes[id].bits & 2) {
      seedSubtrees(nodes[id].child, list);
      seedSubtrees(nodes[id].child + 1, list);
    } else {
      list.push_back(id);
    }
  }

  void objectBounds() {
    const int N = mesh.ntris;
    obj_bounds.resize(N);
    obj_centroids.resize(N);
    for (int i = 0; i < N; ++i) {
      obj_bounds[i] = triBox(mesh, i);
      for (int a = 0; a < 3; ++a)
        obj_centroids[i].v[a] = obj_bounds[i].hi[a] * 0.5f + obj_bounds[i].lo[a] * (1 - 0.5f);
    }
  }

  // parallelUpdateBounds with numProcs > 1 (DynBVH.cc:2034-2065), its workers' subtree
  // tasks run one after another as one worker popping the shared list from the back
  // would: every seeded subtree by updateBounds<false>, then proc 0's updateBounds<true>
  // from the root over the large nodes.
  void updateWorkers() {
    objectBounds();
    std::vector<int> list;
    seedSubtrees(0, list);
    while (!list.empty()) {
      const int id = list.back();
      list.pop_back();
      updateBounds(id);
    }
    updateBoundsLarge(0);
  }

  // DynBVH.cc:2018-2030 then the numProcs == 1 branch
  void update() {
    const int N = mesh.ntris;
    obj_bounds.resize(N);
    obj_centroids.resize(N);
    for (int i = 0; i < N; ++i) {
      obj_bounds[i] = triBox(mesh, i);
      for (int a = 0; a < 3; ++a)
        obj_centroids[i].v[a] = obj_bounds[i].hi[a] * 0.5f + obj_bounds[i].lo[a] * (1 - 0.5f);
    }
    updateBounds(0);
  }
};

// DynBVH::computeCost<true> (DynBVH.cc:3388-3415) and subtree sizes (DynBVH.cc:1971-1996).
float rotCost(const Node32* nodes, float* costs, int id) {
  const Node32& node = nodes[id];
  float cost;
  if (node.isLeaf()) {
    cost = node.children * kCIsec;
  } else {
    const float lc = rotCost(nodes, costs, node.child);
    const float rc = rotCost(nodes, costs, node.child + 1);
    const float area = (float)node.bounds.area();
    const float la = (float)nodes[node.child].bounds.area();
    const float ra = (float)nodes[node.child + 1].bounds.area();
    cost = kCTrav + (la * lc + ra * rc) / area;
  }
  costs[id] = cost;
  return cost;
}
uint32_t rotSizes(const Node32* nodes, uint32_t* ssize, int id) {
  const Node32& node = nodes[id];
  uint32_t s = 1;
  if (!node.isLeaf()) s += rotSizes(nodes, ssize, node.child) + rotSizes(nodes, ssize, node.child + 1);
  ssize[id] = s;
  return s;
}

extern "C" {

int oracle_node_size(void) { return (int)sizeof(Node32); }

// Builds a DynBVH over the mesh exactly as DynBVH::rebuild(0,1) does, then runs the RTSAH
// pass when rtsah != 0. nodes_out must hold 2*ntris 32-byte nodes, ids_out ntris ints.
// Returns the number of nodes in use (num_nodes).
int oracle_build(const float* verts, int nverts, const uint32_t* idx, int ntris, int rtsah,
                 void* nodes_out, int32_t* ids_out) {
  if (ntris <= 0) return -1;
  Mesh m{verts, nverts, idx, ntris};
  std::vector<Node32> nodes;
  std::vector<int32_t> ids;
  Builder b(m, nodes, ids);
  b.rebuild();
  if (rtsah) {
    TTP ttp;
    subtreeCost(nodes, ttp, 0);
  }
  std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(Node32));
  std::memcpy(ids_out, ids.data(), ids.size() * sizeof(int32_t));
  return b.num_nodes;
}

// DynBVH::rebuild(0, 1) built with USE_APPROXIMATE_BUILD set (the MANTA_USE_DYNBVH_APPROXIMATE
// option, include/CMakeLists.txt:134-147, DynBVH.cc:1815-1822): the same rebuild
// (DynBVH.cc:1699-1793) whose every node is partitioned by partitionApproxSAH; leaves sorted,
// computeSubTreeSizes marks, then the RTSAH pass when rtsah != 0. Returns num_nodes.
int oracle_build_approx(const float* verts, int nverts, const uint32_t* idx, int ntris, int rtsah,
                        void* nodes_out, int32_t* ids_out) {
  if (ntris <= 0) return -1;
  Mesh m{verts, nverts, idx, ntris};
  const size_t nrec = 2 * (size_t)ntris;
  std::vector<Node32> nodes(nrec);
  for (auto& nd : nodes) {
    nd.bounds.reset();
    nd.child = 0;
    nd.axis = 0;
    nd.bits = 0;
    nd.children = 0;
  }
  std::vector<int32_t> ids(ntris);
  std::vector<float> costs(nrec, 0.f);
  std::vector<uint32_t> ssize(nrec, 0);
  Box root;  // Mesh::computeBounds: box of all vertices (Mesh.cc:195-225)
  root.reset();
  for (int v = 0; v < nverts; ++v) root.extendPoint(verts + 3 * (size_t)v);
  nodes[0].bounds = root;
  int next_free = 1;
  Updater u(m, nodes.data(), ids.data(), costs.data(), ssize.data(), next_free,
            (uint32_t)(nrec / 20));
  u.obj_bounds.resize(ntris);
  u.obj_centroids.resize(ntris);
  for (int i = 0; i < ntris; ++i) {
    ids[i] = i;
    u.obj_bounds[i] = triBox(m, i);
    for (int a = 0; a < 3; ++a)
      u.obj_centroids[i].v[a] = u.obj_bounds[i].hi[a] * 0.5f + u.obj_bounds[i].lo[a] * (1 - 0.5f);
  }
  u.buildApprox(0, 0, ntris, false);
  // computeSubTreeSizes (DynBVH.cc:1971-1996), largeSubtreeSize = nodes.size() / 20
  const uint32_t large = (uint32_t)(nrec / 20);
  struct Marks {
    std::vector<Node32>& n;
    uint32_t large;
    uint32_t go(int id) {
      if (n[id].isLeaf()) {
        n[id].bits &= ~2;
        return 1;
      }
      const uint32_t s = go(n[id].child) + go(n[id].child + 1) + 1;
      if (s > large)
        n[id].bits |= 2;
      else
        n[id].bits &= ~2;
      return s;
    }
  } marks{nodes, large};
  marks.go(0);
  if (rtsah) {
    TTP ttp;
    subtreeCost(nodes, ttp, 0);
  }
  std::memcpy(nodes_out, nodes.data(), nrec * sizeof(Node32));
  std::memcpy(ids_out, ids.data(), ids.size() * sizeof(int32_t));
  return next_free;  // every allocated record is in the tree
}

// What shading reads from a hit, per ray: the hit position
// (RayPacket::actualComputeHitPositions, Interface/RayPacket.cc:114-146), the shading
// normal (KenslerShirleyTriangle::computeNormal, KenslerShirleyTriangle.cc:381-405;
// WaldTriangle::computeNormal, WaldTriangle.cc:558-590), the geometric normal
// (KenslerShirleyTriangle.cc:407-420; WaldTriangle.cc:592-609) and the texture coordinates
// (computeTexCoords2, KenslerShirleyTriangle.cc:85-123; WaldTriangle.cc:131-169). u and v
// are the scratchpad weights of vertices 1 and 2. normals/nidx (Mesh::vertexNormals and
// normal_indices) and tex/tidx (Mesh::texCoords and texture_indices) may be null. out
// holds 12 floats per ray: position, normal, geometric normal, texture coordinates; rays
// that missed (prim < 0) get zeros.
void oracle_shade(const float* verts, int nverts, const uint32_t* idx, int ntris, int prim_kind,
                  const float* normals, const uint32_t* nidx, const float* tex,
                  const uint32_t* tidx, const float* const* ray6, const int32_t* prim,
                  const float* t, const float* u, const float* v, int64_t n, float* out) {
  Mesh m{verts, nverts, idx, ntris};
  for (int64_t r = 0; r < n; ++r) {
    float* o = out + 12 * r;
    for (int k = 0; k < 12; ++k) o[k] = 0.f;
    const int id = prim[r];
    if (id < 0) continue;
    for (int j = 0; j < 3; ++j) o[j] = ray6[j][r] + ray6[3 + j][r] * t[r];
    const float a = u[r], b = v[r];
    const float c = (float)(1.0 - a - b);  // Real c = (1.0 - a - b): double, then narrowed
    const float* p0 = m.vtx(id, 0);
    const float* p1 = m.vtx(id, 1);
    const float* p2 = m.vtx(id, 2);
    float face[3];
    if (prim_kind == 1) {
      const Wald w = waldSetup(p0, p1, p2);
      const int ku = (w.k == 2) ? 0 : w.k + 1, kv = (w.k == 0) ? 2 : w.k - 1;
      face[w.k] = w.n_k;
      face[ku] = w.n_k * w.n_u;
      face[kv] = w.n_k * w.n_v;
    } else {  // Cross(p1 - p0, p2 - p0)
      const float e0[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      const float e1[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      face[0] = e0[1] * e1[2] - e0[2] * e1[1];
      face[1] = e0[2] * e1[0] - e0[0] * e1[2];
      face[2] = e0[0] * e1[1] - e0[1] * e1[0];
    }
    for (int j = 0; j < 3; ++j) o[6 + j] = face[j];
    if (normals && nidx) {  // (n1 * a) + (n2 * b) + (n0 * c)
      const float* n0 = normals + 3 * (size_t)nidx[3 * (size_t)id + 0];
      const float* n1 = normals + 3 * (size_t)nidx[3 * (size_t)id + 1];
      const float* n2 = normals + 3 * (size_t)nidx[3 * (size_t)id + 2];
      for (int j = 0; j < 3; ++j) o[3 + j] = (n1[j] * a + n2[j] * b) + n0[j] * c;
    } else {
      for (int j = 0; j < 3; ++j) o[3 + j] = face[j];
    }
    const uint32_t t0 = (tex && tidx) ? tidx[3 * (size_t)id] : 0xFFFFFFFFu;
    if (t0 == 0xFFFFFFFFu) {  // Mesh::kNoTextureIndex: the barycentric weights
      o[9] = a;
      o[10] = b;
      o[11] = c;
    } else {
      const float* x0 = tex + 3 * (size_t)t0;
      const float* x1 = tex + 3 * (size_t)tidx[3 * (size_t)id + 1];
      const float* x2 = tex + 3 * (size_t)tidx[3 * (size_t)id + 2];
      for (int j = 0; j < 3; ++j) o[9 + j] = (x1[j] * a + x2[j] * b) + x0[j] * c;
    }
  }
}

// partitionObjects (DynBVH.cc:1369-1396) alone: slots 0..n-1 hold objects 0..n-1 whose
// centroid on axis 0 is keys[i]; perm receives the object order after the partition.
// Returns the split (first slot of the right side).
int oracle_partition_objects(const float* keys, int n, float position, int32_t* perm) {
  Mesh m{nullptr, 0, nullptr, 0};
  std::vector<Node32> nodes(1);
  std::vector<float> costs(1);
  std::vector<uint32_t> ssize(1);
  int nf = 1;
  Updater u(m, nodes.data(), perm, costs.data(), ssize.data(), nf, 0);
  u.obj_centroids.resize(n);
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    u.obj_centroids[i].v[0] = keys[i];
    u.obj_centroids[i].v[1] = u.obj_centroids[i].v[2] = 0.f;
  }
  return u.partitionObjects(0, n, 0, position);
}

// TreeTraversalProb::computeProbs for one parent/left/right triple; boxes are
// lo.xyz, hi.xyz. out3 = {lr, jl, jr}.
void oracle_ttp_probs(const float* parent6, const float* left6, const float* right6,
                      float* out3) {
  auto mk = [](const float* b) {
    Box x;
    for (int i = 0; i < 3; ++i) {
      x.lo[i] = b[i];
      x.hi[i] = b[3 + i];
    }
    return x;
  };
  TTP ttp;
  ttp.probs(mk(parent6), mk(left6), mk(right6), out3[0], out3[1], out3[2]);
}

// Recomputes isLeftCheaper on an existing tree (DynBVH::computeTraversalCost).
void oracle_rtsah(void* nodes, int n_nodes) {
  std::vector<Node32> nv((Node32*)nodes, (Node32*)nodes + n_nodes);
  TTP ttp;
  subtreeCost(nv, ttp, 0);
  std::memcpy(nodes, nv.data(), nv.size() * sizeof(Node32));
}

// The same pass, also writing per record (c0, c1, lCost, rCost) to costs4 (4 floats per
// record; leaves: their (c0, c1) and zeros; records the walk does not reach: untouched).
// c0/c1 are computeSubTreeTraversalCost's cost[0]/cost[1] (DynBVH.cc:1594-1678).
void oracle_rtsah_costs(void* nodes, int n_nodes, float* costs4) {
  std::vector<Node32> nv((Node32*)nodes, (Node32*)nodes + n_nodes);
  TTP ttp;
  g_cost_out = costs4;
  subtreeCost(nv, ttp, 0);
  g_cost_out = nullptr;
  std::memcpy(nodes, nv.data(), nv.size() * sizeof(Node32));
}

// Traces n rays through the tree in packets of packet_size (1..64) consecutive rays.
// prim: triangle id, -1 no hit, -2 masked. t: final minT. occluded: RayPacket::wasHit.
// counters (optional, int64[2]): total box tests and triangle tests.
// mode 0: the scalar build (default); mode 1: the SSE build (Kensler-Shirley only, no
// counters).
int oracle_trace_mode(const float* verts, int nverts, const uint32_t* idx, int ntris,
                      const void* nodes, const int32_t* ids, int prim_kind, const float* ox,
                      const float* oy, const float* oz, const float* dx, const float* dy,
                      const float* dz, const float* tmax, const uint8_t* maskd, int64_t n,
                      int packet_size, int any_hit, int nthreads, int32_t* prim, float* t,
                      float* u, float* v, uint8_t* occluded, int64_t* counters, int mode) {
  if (packet_size < 1 || packet_size > 64) return -1;
  if (mode != 0 && mode != 1) return -1;
  if (mode == 1 && (prim_kind != 0 || counters)) return -1;
  Scene sc;
  sc.mesh = Mesh{verts, nverts, idx, ntris};
  sc.nodes = (const Node32*)nodes;
  sc.ids = ids;
  sc.prim_kind = prim_kind;
  if (prim_kind == 1) {
    sc.wald.resize(ntris);
    for (int i = 0; i < ntris; ++i)
      sc.wald[i] = waldSetup(sc.mesh.vtx(i, 0), sc.mesh.vtx(i, 1), sc.mesh.vtx(i, 2));
  }
  RayIO io{ox, oy, oz, dx, dy, dz, tmax, maskd, prim, t, u, v, occluded};
  const int64_t npk = (n + packet_size - 1) / packet_size;
  int nt = nthreads < 1 ? 1 : nthreads;
  std::vector<Counters> per(nt);
  parallelFor(npk, nt, [&](int64_t k, int tid) {
    int64_t first = k * packet_size;
    int cnt = (int)std::min<int64_t>(packet_size, n - first);
    tracePacketRange(sc, io, first, cnt, any_hit != 0, counters ? &per[tid] : nullptr,
                     mode == 1);
  });
  if (counters) {
    counters[0] = counters[1] = 0;
    for (auto& c : per) {
      counters[0] += c.box_tests;
      counters[1] += c.tri_tests;
    }
  }
  return 0;
}

int oracle_trace(const float* verts, int nverts, const uint32_t* idx, int ntris,
                 const void* nodes, const int32_t* ids, int prim_kind, const float* ox,
                 const float* oy, const float* oz, const float* dx, const float* dy,
                 const float* dz, const float* tmax, const uint8_t* maskd, int64_t n,
                 int packet_size, int any_hit, int nthreads, int32_t* prim, float* t,
                 float* u, float* v, uint8_t* occluded, int64_t* counters) {
  return oracle_trace_mode(verts, nverts, idx, ntris, nodes, ids, prim_kind, ox, oy, oz, dx,
                           dy, dz, tmax, maskd, n, packet_size, any_hit, nthreads, prim, t, u,
                           v, occluded, counters, 0);
}

// Brute force over all triangles in ascending id order with the same triangle test and
// RayPacket::hit rule; no tree. Used to pin the traversal.
int oracle_bruteforce(const float* verts, int nverts, const uint32_t* idx, int ntris,
                      int prim_kind, const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz, const float* tmax,
                      const uint8_t* maskd, int64_t n, int any_hit, int nthreads, int32_t* prim,
                      float* t, float* u, float* v, uint8_t* occluded) {
  Scene sc;
  sc.mesh = Mesh{verts, nverts, idx, ntris};
  sc.nodes = nullptr;
  sc.ids = nullptr;
  sc.prim_kind = prim_kind;
  if (prim_kind == 1) {
    sc.wald.resize(ntris);
    for (int i = 0; i < ntris; ++i)
      sc.wald[i] = waldSetup(sc.mesh.vtx(i, 0), sc.mesh.vtx(i, 1), sc.mesh.vtx(i, 2));
  }
  parallelFor(n, nthreads < 1 ? 1 : nthreads, [&](int64_t r, int) {
    Packet p;
    p.anyHit = any_hit != 0;
    p.org[0][0] = ox[r];
    p.org[1][0] = oy[r];
    p.org[2][0] = oz[r];
    p.dir[0][0] = dx[r];
    p.dir[1][0] = dy[r];
    p.dir[2][0] = dz[r];
    p.prim[0] = -1;
    p.u[0] = p.v[0] = 0;
    if (maskd && maskd[r]) {
      p.state[0] = -1;
      p.minT[0] = -kMaxT;
    } else {
      p.state[0] = 0;
      p.minT[0] = tmax[r];
    }
    if (p.state[0] == 0) {
      for (int tri = 0; tri < ntris; ++tri) {
        intersectPrim(sc, tri, p, 0, 1);
        if (p.anyHit && p.state[0] == 1) break;
      }
    }
    if (prim) prim[r] = (p.state[0] == 1) ? p.prim[0] : (p.state[0] == -1 ? -2 : -1);
    if (t) t[r] = p.minT[0];
    if (u) u[r] = p.u[0];
    if (v) v[r] = p.v[0];
    if (occluded) occluded[r] = wasHit(p, 0) ? 1 : 0;
  });
  return 0;
}

// Refit without tree rotations: DynBVH::parallelUpdateBounds(numProcs == 1) ->
// updateBounds<false>(0) (DynBVH.cc:2012-2123) on the current vertices, in place.
void oracle_refit(const float* verts, int nverts, const uint32_t* idx, int ntris, void* nodes_v,
                  const int32_t* ids) {
  Mesh m{verts, nverts, idx, ntris};
  Node32* nodes = (Node32*)nodes_v;
  std::vector<Box> ob(ntris);
  for (int i = 0; i < ntris; ++i) ob[i] = triBox(m, i);
  struct R {
    Node32* nodes;
    const int32_t* ids;
    const std::vector<Box>& ob;
    void go(int id) {
      Node32& node = nodes[id];
      if (node.isLeaf()) {
        node.bounds.reset();
        for (int i = 0; i < node.children; ++i) node.bounds.extendBox(ob[ids[node.child + i]]);
        node.axis = 0;
      } else {
        go(node.child);
        go(node.child + 1);
        node.bounds.reset();
        node.bounds.extendBox(nodes[node.child].bounds);
        node.bounds.extendBox(nodes[node.child + 1].bounds);
      }
    }
  } rr{nodes, ids, ob};
  rr.go(0);
}

// Rotation state after DynBVH::rebuild with TREE_ROT: costs (computeCost<true>(0)) and
// subtree sizes of the reachable records; the rest stay zero.
void oracle_rot_init(const void* nodes_v, int n_records, float* costs, uint32_t* ssize) {
  const Node32* nodes = (const Node32*)nodes_v;
  for (int i = 0; i < n_records; ++i) {
    costs[i] = 0.f;
    ssize[i] = 0;
  }
  rotCost(nodes, costs, 0);
  rotSizes(nodes, ssize, 0);
}

// One DynBVH::update (TREE_ROT, numProcs == 1) on the current vertices, in place: nodes
// (2T records), object ids, costs, subtree sizes and the next free record pair.
void oracle_update(const float* verts, int nverts, const uint32_t* idx, int ntris, void* nodes_v,
                   int32_t* ids, float* costs, uint32_t* ssize, int* next_free) {
  Mesh m{verts, nverts, idx, ntris};
  Updater u(m, (Node32*)nodes_v, ids, costs, ssize, *next_free, (uint32_t)(2 * (size_t)ntris / 20));
  u.update();
}

// DynBVH::update with numProcs workers (TREE_ROT): largeSubtreeSize = nodes.size() /
// (20 * numProcs) (DynBVH.cc:945, 1776); one worker runs the numProcs == 1 branch, more run
// the seeded-subtree schedule.
void oracle_update_workers(const float* verts, int nverts, const uint32_t* idx, int ntris,
                           void* nodes_v, int32_t* ids, float* costs, uint32_t* ssize,
                           int* next_free, int num_procs) {
  Mesh m{verts, nverts, idx, ntris};
  const uint32_t large = (uint32_t)(2 * (size_t)ntris / (20 * (size_t)num_procs));
  Updater u(m, (Node32*)nodes_v, ids, costs, ssize, *next_free, large);
  if (num_procs <= 1)
    u.update();
  else
    u.updateWorkers();
}

// computeSubTreeSizes (DynBVH.cc:1971-1996) with largeSubtreeSize = n_records /
// (20 * numProcs): the isLargeSubtree bits rebuild(proc, numProcs) leaves (DynBVH.cc:1776)
// and preprocess sets on a loaded tree (DynBVH.cc:941-950).
void oracle_mark_large(void* nodes_v, int n_records, int num_procs) {
  Node32* nodes = (Node32*)nodes_v;
  const uint32_t large = (uint32_t)((size_t)n_records / (20 * (size_t)num_procs));
  struct Marks {
    Node32* n;
    uint32_t large;
    uint32_t go(int id) {
      if (n[id].isLeaf()) {
        n[id].bits &= ~2;
        return 1;
      }
      const uint32_t sz = go(n[id].child) + go(n[id].child + 1) + 1;
      if (sz > large)
        n[id].bits |= 2;
      else
        n[id].bits &= ~2;
      return sz;
    }
  } marks{nodes, large};
  marks.go(0);
}

// Wald per-triangle constants (11 floats + k as float) for tests of the setup pass.
void oracle_wald_setup(const float* verts, const uint32_t* idx, int ntris, float* out12) {
  Mesh m{verts, 0, idx, ntris};
  for (int i = 0; i < ntris; ++i) {
    Wald w = waldSetup(m.vtx(i, 0), m.vtx(i, 1), m.vtx(i, 2));
    float* o = out12 + 12 * (size_t)i;
    o[0] = w.n_u; o[1] = w.n_v; o[2] = w.n_d; o[3] = (float)w.k;
    o[4] = w.b_nu; o[5] = w.b_nv; o[6] = w.b_d; o[7] = w.n_k;
    o[8] = w.c_nu; o[9] = w.c_nv; o[10] = w.c_d; o[11] = 0.f;
  }
}

}  // extern "C"
