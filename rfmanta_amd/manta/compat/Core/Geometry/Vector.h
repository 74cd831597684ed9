// Subset of Manta's Vector (Core/Geometry/Vector.h:51-485): three Reals, indexable.
#ifndef Manta_Core_Vector_h
#define Manta_Core_Vector_h

namespace Manta {
typedef float Real;

class Vector {
 public:
  Vector() {}
  Vector(Real x, Real y, Real z) {
    data[0] = x;
    data[1] = y;
    data[2] = z;
  }
  Real operator[](int i) const { return data[i]; }
  Real& operator[](int i) { return data[i]; }
  Real x() const { return data[0]; }
  Real y() const { return data[1]; }
  Real z() const { return data[2]; }
  Real data[3];
};
}  // namespace Manta
#endif
