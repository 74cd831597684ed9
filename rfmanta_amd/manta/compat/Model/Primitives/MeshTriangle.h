// Subset of Manta's MeshTriangle, KenslerShirleyTriangle and WaldTriangle
// (Model/Primitives/MeshTriangle.h:12-65, KenslerShirleyTriangle.h, WaldTriangle.h).
// Each triangle is its own TexCoordMapper, as in the reference. The per-triangle
// intersect is the CPU path; under HipDynBVH the GPU performs the same test.
#ifndef Manta_Model_MeshTriangle_h
#define Manta_Model_MeshTriangle_h

#include <Core/Exceptions/InternalError.h>
#include <Interface/Primitive.h>
#include <Interface/TexCoordMapper.h>

namespace Manta {
class Mesh;

class MeshTriangle : public Primitive {
 public:
  MeshTriangle() : myID(0), mesh(0) {}
  virtual void attachMesh(Mesh* m, unsigned int id) {
    mesh = m;
    myID = id;
  }
  unsigned int getTriangleID() const { return myID; }
  void computeBounds(const PreprocessContext&, BBox&) const {
    throw InternalError("MeshTriangle::computeBounds is provided by the Manta tree");
  }
  void intersect(const RenderContext&, RayPacket&) const {
    throw InternalError("MeshTriangle::intersect is provided by the Manta tree");
  }

 protected:
  unsigned int myID;
  Mesh* mesh;
};

class KenslerShirleyTriangle : public MeshTriangle, public TexCoordMapper {
 public:
  KenslerShirleyTriangle() {}
  KenslerShirleyTriangle(Mesh* m, unsigned int id) { attachMesh(m, id); }
};

class WaldTriangle : public MeshTriangle, public TexCoordMapper {
 public:
  WaldTriangle() {}
  WaldTriangle(Mesh* m, unsigned int id) { attachMesh(m, id); }
};
}  // namespace Manta
#endif
