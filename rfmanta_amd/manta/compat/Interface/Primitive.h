// Subset of Manta's Primitive (Interface/Primitive.h:12-41).
#ifndef Manta_Interface_Primitive_h
#define Manta_Interface_Primitive_h

#include <Interface/Object.h>

namespace Manta {
class Primitive : public Object {
 public:
  Primitive() {}
  virtual ~Primitive() {}
};
}  // namespace Manta
#endif
