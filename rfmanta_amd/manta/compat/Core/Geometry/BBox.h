// Subset of Manta's BBox (Core/Geometry/BBox.h:12-148).
#ifndef Manta_Core_BBox_h
#define Manta_Core_BBox_h

#include <Core/Geometry/Vector.h>
#include <limits>

namespace Manta {
class BBox {
 public:
  BBox() { reset(); }
  BBox(const Vector& lo, const Vector& hi) {
    bounds[0] = lo;
    bounds[1] = hi;
  }
  void reset() {
    const Real m = std::numeric_limits<Real>::max();
    bounds[0] = Vector(m, m, m);
    bounds[1] = Vector(-m, -m, -m);
  }
  void extendByPoint(const Vector& p) {
    for (int i = 0; i < 3; ++i) {
      bounds[0][i] = bounds[0][i] < p[i] ? bounds[0][i] : p[i];
      bounds[1][i] = bounds[1][i] > p[i] ? bounds[1][i] : p[i];
    }
  }
  void extendByBox(const BBox& b) {
    extendByPoint(b.bounds[0]);
    extendByPoint(b.bounds[1]);
  }
  const Vector& getMin() const { return bounds[0]; }
  const Vector& getMax() const { return bounds[1]; }
  Vector& operator[](int i) { return bounds[i]; }
  const Vector& operator[](int i) const { return bounds[i]; }

 private:
  Vector bounds[2];
};
}  // namespace Manta
#endif
