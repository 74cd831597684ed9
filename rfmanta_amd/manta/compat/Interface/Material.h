// Subset of Manta's Material (Interface/Material.h:43-66).
#ifndef Manta_Interface_Material_h
#define Manta_Interface_Material_h

namespace Manta {
class Material {
 public:
  virtual ~Material() {}
  virtual bool canAttenuateShadows() const { return false; }
};
}  // namespace Manta
#endif
