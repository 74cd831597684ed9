// Subset of Manta's Group (Model/Groups/Group.h:47-115): a list of objects intersected
// in turn.
#ifndef Manta_Model_Group_h
#define Manta_Model_Group_h

#include <Core/Geometry/BBox.h>
#include <Interface/Object.h>

#include <cstddef>
#include <vector>

namespace Manta {
class Group : public Object {
 public:
  Group() {}
  virtual ~Group() {}
  virtual void add(Object* obj) { objs.push_back(obj); }
  Object* get(size_t i) { return objs[i]; }
  const Object* get(size_t i) const { return objs[i]; }
  size_t size() const { return objs.size(); }
  void preprocess(const PreprocessContext& c) {
    for (size_t i = 0; i < objs.size(); ++i) objs[i]->preprocess(c);
  }
  void computeBounds(const PreprocessContext& c, BBox& bbox) const {
    for (size_t i = 0; i < objs.size(); ++i) objs[i]->computeBounds(c, bbox);
  }
  void intersect(const RenderContext& c, RayPacket& rays) const {
    for (size_t i = 0; i < objs.size(); ++i) objs[i]->intersect(c, rays);
  }

 protected:
  std::vector<Object*> objs;
};
}  // namespace Manta
#endif
