// RTSAH child order (DynBVH::computeTraversalCost, Model/Groups/DynBVH.cc:1583-1678, with
// TreeTraversalProb::computeProbs, Model/Groups/TreeTraversalProbability.cc:3-95), shared
// by the host builder and the device build so both set the same isLeftCheaper bits.
//
// Plain binary32 operations only (sqrtf and division correctly rounded on both sides,
// every unit compiled without contraction); the logarithm is the reference's table
// logarithm, whose table the host fills once with logf and hands to the device.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "sah_builder.h"

#if defined(__HIPCC__)
#define MH_HD __host__ __device__
#else
#define MH_HD
#endif

namespace mh {
namespace rtsah {

constexpr int kLogBits = 15;  // Core/Math/ExponSSE.h:20-73 at precision 15
constexpr int kLogTable = 1 << kLogBits;

// The table of log2(1 + (i + 1/2) / 2^kLogBits) (host only).
inline void fill_log_table(float* tab) {
  float x = 1.0f + (1.0f / (float)(1 << (kLogBits + 1)));
  for (int i = 0; i < kLogTable; ++i) {
    tab[i] = logf(x) / 0.69314718055995f;
    x += 1.0f / (float)(1 << kLogBits);
  }
}

MH_HD inline float fast_log(const float* tab, float val) {
  int32_t w;
  std::memcpy(&w, &val, 4);
  const int ex = ((w >> 23) & 255) - 127;
  const int mi = (w & 0x7FFFFF) >> (23 - kLogBits);
  return ((float)ex + tab[mi]) * 0.69314718055995f;
}

MH_HD inline float approx_atan2(float y, float x) {  // Core/Math/Trig.h:153-178
  const float half_pi = (float)M_PI_2, pi = (float)M_PI;
  if (x == 0.0f) return y > 0.0f ? half_pi : (y == 0.0f ? 0.0f : -half_pi);
  const float z = y / x;
  float r;
  if (fabsf(z) < 1.0f) {
    r = z / (1.0f + 0.28f * z * z);
    if (x < 0.0f) return y < 0.0f ? r - pi : r + pi;
  } else {
    r = half_pi - z / (z * z + 0.28f);
    if (y < 0.0f) return r - pi;
  }
  return r;
}

MH_HD inline float lo_of(float a, float b) { return a < b ? a : b; }  // Core/Math/MinMax.h:88-91
MH_HD inline float hi_of(float a, float b) { return a > b ? a : b; }  // Core/Math/MinMax.h:83-86

MH_HD inline float area(const Aabb& a) {  // BBox::computeArea: float arithmetic
  const float ex = a.hi[0] - a.lo[0], ey = a.hi[1] - a.lo[1], ez = a.hi[2] - a.lo[2];
  return 2 * (ex * ey + ey * ez + ex * ez);
}
MH_HD inline float inv_area(const Aabb& a) { return (float)(1.0f / (double)area(a)); }

MH_HD inline float sgn(int i) { return i % 2 == 1 ? -1.0f : 1.0f; }

MH_HD inline float perp_term(const float* tab, float v, float z, float x, float y, float z0,
                             float x1) {
  const float a = y - v, b = z0 - z, c = x - x1;
  const float aa = a * a, cc = c * c, bb = b * b;
  const float root = sqrtf(cc + bb);
  return a * root * approx_atan2(a, root) + .25f * (aa - bb - cc) * fast_log(tab, aa + bb + cc);
}

MH_HD inline float perpendicular_ff(const float* tab, const Aabb& b0, const Aabb& b1, int d1,
                                    int d2) {
  const int ax = (~(unsigned)(d1 + d2)) & 3;
  const float v[2] = {b1.lo[ax], b1.hi[ax]}, y[2] = {b0.lo[ax], b0.hi[ax]};
  const float x[2] = {b0.lo[d2], b0.hi[d2]}, z[2] = {b1.lo[d1], b1.hi[d1]};
  const float x1 = b1.lo[d2], z0 = b0.lo[d1];
  float F = 0;
  for (int i = 1; i <= 2; ++i)
    for (int j = 1; j <= 2; ++j)
      for (int k = 1; k <= 2; ++k)
        for (int m = 1; m <= 2; ++m)
          F += perp_term(tab, v[i - 1], z[j - 1], x[k - 1], y[m - 1], z0, x1) * sgn(i + j + k + m);
  F *= (float)M_1_PI;
  return fabsf(F);
}

MH_HD inline float par_term(const float* tab, float u, float v, float x, float y, float zi) {
  const float a = (x - u) * zi, b = (y - v) * zi;
  const float aa = a * a, bb = b * b;
  const float ra = sqrtf(1 + aa), rb = sqrtf(1 + bb);
  return b * ra * approx_atan2(b, ra) + a * rb * approx_atan2(a, rb) -
         0.5f * fast_log(tab, 1 + aa + bb);
}

MH_HD inline float parallel_ff(const float* tab, const Aabb& b0, const Aabb& b1, int flat) {
  const int d1 = (flat + 1) % 3, d2 = (flat + 2) % 3;
  const float y[2] = {b0.lo[d1], b0.hi[d1]}, v[2] = {b1.lo[d1], b1.hi[d1]};
  const float x[2] = {b0.lo[d2], b0.hi[d2]}, u[2] = {b1.lo[d2], b1.hi[d2]};
  const float z = b0.lo[flat] - b1.lo[flat];
  const float zi = 1.0f / z;
  float F = 0;
  for (int i = 1; i <= 2; ++i)
    for (int j = 1; j <= 2; ++j)
      for (int k = 1; k <= 2; ++k)
        for (int m = 1; m <= 2; ++m)
          F += par_term(tab, u[i - 1], v[j - 1], x[k - 1], y[m - 1], zi) * sgn(i + j + k + m);
  F *= z * z * (float)M_1_PI;
  return fabsf(F);
}

MH_HD inline float overlap_area(const Aabb& a, const Aabb& b) {
  Aabb o;
  for (int k = 0; k < 3; ++k) {
    o.lo[k] = hi_of(a.lo[k], b.lo[k]);
    o.hi[k] = lo_of(a.hi[k], b.hi[k]);
  }
  if (o.hi[0] - o.lo[0] < 0 || o.hi[1] - o.lo[1] < 0 || o.hi[2] - o.lo[2] < 0) return 0;
  return area(o);
}

// Probabilities that a ray through `parent` crosses both children (lr), only the left
// (jl) or only the right (jr); TreeTraversalProbability.cc:3-95.
MH_HD inline void probabilities(const float* tab, const Aabb& parent, const Aabb& L,
                                const Aabb& R, float& lr, float& jl, float& jr) {
  float dl[3], dr[3];
  for (int k = 0; k < 3; ++k) {
    dl[k] = L.hi[k] - L.lo[k];
    dr[k] = R.hi[k] - R.lo[k];
  }
  const bool thinL[3] = {dl[1] == 0 || dl[2] == 0, dl[0] == 0 || dl[2] == 0,
                         dl[0] == 0 || dl[1] == 0};
  const bool thinR[3] = {dr[1] == 0 || dr[2] == 0, dr[0] == 0 || dr[2] == 0,
                         dr[0] == 0 || dr[1] == 0};
  lr = 0;
  for (int i = 0; i < 6; ++i) {
    const int fi = i % 3;
    if (thinL[fi]) continue;
    if (i < 3 ? (L.hi[fi] >= R.hi[fi]) : (L.lo[fi] <= R.lo[fi])) continue;
    for (int j = 0; j < 6; ++j) {
      if (i == j) continue;
      const int fj = j % 3;
      if (thinR[fj]) continue;
      Aabb pl = L, pr = R;
      if (i < 3)
        pl.lo[fi] = pl.hi[fi];
      else
        pl.hi[fi] = pl.lo[fi];
      if (j < 3)
        pr.lo[fj] = pr.hi[fj];
      else
        pr.hi[fj] = pr.lo[fj];
      if (fi == fj) {
        if (i < 3 ? (pl.hi[fi] >= pr.lo[fi]) : (pl.lo[fi] <= pr.hi[fi])) continue;
        lr += parallel_ff(tab, pl, pr, fi);
      } else {
        if (j < 3) {
          if (pl.hi[fj] <= pr.hi[fj]) continue;
          pl.lo[fj] = hi_of(pl.lo[fj], pr.hi[fj]);
        } else {
          if (pl.lo[fj] >= pr.lo[fj]) continue;
          pl.hi[fj] = lo_of(pl.hi[fj], pr.lo[fj]);
        }
        if (i < 3)
          pr.lo[fi] = hi_of(pr.lo[fi], pl.hi[fi]);
        else
          pr.hi[fi] = lo_of(pr.hi[fi], pl.lo[fi]);
        lr += perpendicular_ff(tab, pl, pr, fi, fj);
      }
    }
  }
  const float pinv = inv_area(parent);
  lr += overlap_area(L, R);
  lr *= pinv;
  jl = (float)((double)area(L) * (double)pinv - (double)lr);
  jr = (float)((double)area(R) * (double)pinv - (double)lr);
}

// Expected traversal cost (c0) and pass-through probability (c1) of a subtree
// (DynBVH.cc:1594-1678).
struct CostPair {
  float c0, c1;
};

MH_HD inline CostPair leaf_cost(int children) { return CostPair{children * 3.0f + 1.0f, 0.0f}; }

// An inner node's pair from its children's; left_cheaper receives isLeftCheaper.
MH_HD inline CostPair combine(const float* tab, const Aabb& P, const Aabb& Lb, const Aabb& Rb,
                              const CostPair& L, const CostPair& R, bool& left_cheaper) {
  float Plr, Pjl, Pjr;
  probabilities(tab, P, Lb, Rb, Plr, Pjl, Pjr);
  const float Pl = Pjl + Plr, Pr = Pjr + Plr;
  const float Pe = 1 - Pjr - Pjl - Plr;
  const float lc = Pl * L.c0 + (Pjr + Plr * L.c1) * R.c0 + (1 + Pl * L.c1 + (1 - Pl)) * 1.0f;
  const float rc = Pr * R.c0 + (Pjl + Plr * R.c1) * L.c0 + (1 + Pr * R.c1 + (1 - Pr)) * 1.0f;
  float c1 = Pjl * L.c1 + Pjr * R.c1 + Plr * L.c1 * R.c1 + Pe;
  c1 = c1 <= 0.0f ? 0.0f : (c1 >= 1.0f ? 1.0f : c1);
  left_cheaper = lc < rc;
  return CostPair{lo_of(lc, rc), c1};
}

}  // namespace rtsah
}  // namespace mh
