// Subset of Manta's Object (Interface/Object.h:22-41).
#ifndef Manta_Interface_Object_h
#define Manta_Interface_Object_h

namespace Manta {
class BBox;
class PreprocessContext;
class RenderContext;
class RayPacket;
template <class KeyType> class UpdateGraph;
template <class KeyType> class UpdateGraphNode;
class Object;
typedef UpdateGraph<Object*> ObjectUpdateGraph;
typedef UpdateGraphNode<Object*> ObjectUpdateGraphNode;

class Object {
 public:
  Object() {}
  virtual ~Object() {}
  virtual void preprocess(const PreprocessContext&) {}
  virtual void computeBounds(const PreprocessContext& context, BBox& bbox) const = 0;
  virtual void intersect(const RenderContext& context, RayPacket& rays) const = 0;
  virtual void addToUpdateGraph(ObjectUpdateGraph*, ObjectUpdateGraphNode*) {}
};
}  // namespace Manta
#endif
