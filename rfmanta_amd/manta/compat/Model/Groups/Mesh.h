// Subset of Manta's Mesh (Model/Groups/Mesh.h:17-118): indexed triangles whose
// primitives are MeshTriangle objects numbered in insertion order.
#ifndef Manta_Model_Mesh_h
#define Manta_Model_Mesh_h

#include <Core/Geometry/BBox.h>
#include <Model/Groups/Group.h>

#include <vector>

namespace Manta {
class Material;
class MeshTriangle;

class Mesh : public Group {
 public:
  std::vector<Vector> vertices;
  std::vector<Material*> materials;
  std::vector<unsigned int> vertex_indices;  // 3 per triangle
  std::vector<unsigned int> face_material;   // 1 per triangle

  inline MeshTriangle* get(size_t i);
  inline const MeshTriangle* get(size_t i) const;
  inline void addTriangle(MeshTriangle* tri);
  size_t numTriangles() const { return objs.size(); }
  Vector getVertex(size_t tri, size_t which) const { return vertices[vertex_indices[3 * tri + which]]; }
  // Mesh::computeBounds is the box of every vertex (Model/Groups/Mesh.cc:195-225).
  void computeBounds(const PreprocessContext&, BBox& bbox) const {
    for (size_t i = 0; i < vertices.size(); ++i) bbox.extendByPoint(vertices[i]);
  }
};
}  // namespace Manta

#include <Model/Primitives/MeshTriangle.h>

namespace Manta {
inline MeshTriangle* Mesh::get(size_t i) { return static_cast<MeshTriangle*>(objs[i]); }
inline const MeshTriangle* Mesh::get(size_t i) const {
  return static_cast<const MeshTriangle*>(objs[i]);
}
inline void Mesh::addTriangle(MeshTriangle* tri) {
  tri->attachMesh(this, (unsigned)objs.size());
  objs.push_back(tri);
}
}  // namespace Manta
#endif
