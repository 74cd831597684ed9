// HipDynBVH: Manta's DynBVH acceleration structure with traversal, triangle tests and
// refit on an MI355X, through the C ABI of libmh.so (include/mh.h).
//
// Drop-in for Model/Groups/DynBVH (Model/Groups/DynBVH.h:22-391): the same
// AccelerationStructure interface (Interface/AccelerationStructure.h:11-66), the same tree
// (identical SAH build and DynBVH file format), and the same per-ray results as
// DynBVH::intersect on a Mesh of KenslerShirleyTriangle or WaldTriangle primitives.
//
// intersect() keeps Manta's per-packet contract: it is const, re-entrant, and called by
// every render thread with its own RayPacket. Packets from concurrent callers are
// gathered into one device launch (one 64-ray packet per wave, the reference's packet
// traversal) and each caller returns with its results written back into its packet.
#ifndef Manta_Model_HipDynBVH_h
#define Manta_Model_HipDynBVH_h

#include <Core/Geometry/BBox.h>
#include <Interface/AccelerationStructure.h>
#include <Interface/RayPacket.h>

#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>

#include "mh.h"

namespace Manta {
class Group;
class Mesh;

class HipDynBVH : public AccelerationStructure {
 public:
  HipDynBVH();
  virtual ~HipDynBVH();

  // AccelerationStructure
  void setGroup(Group* new_group);
  void groupDirty();
  Group* getGroup() const;
  void rebuild(int proc = 0, int numProcs = 1);
  void update(int proc = 0, int numProcs = 1);
  void addToUpdateGraph(ObjectUpdateGraph* graph, ObjectUpdateGraphNode* parent);
  bool buildFromFile(const std::string& file);
  bool saveToFile(const std::string& file);

  // Object
  void preprocess(const PreprocessContext& context);
  void computeBounds(const PreprocessContext& context, BBox& bbox) const;
  void intersect(const RenderContext& context, RayPacket& rays) const;

  // Packets gathered per device launch (at most; a launch also starts when every render
  // thread has a packet waiting, or after `wait_us` microseconds).
  void setBatching(int max_packets, int wait_us);

 private:
  class PacketBatcher;

  void rebuildSerial(int numProcs);
  void ensureScene();
  void rendezvous(int numProcs);
  [[noreturn]] void fail(const char* what) const;

  Group* group_;
  Mesh* mesh_;
  bool group_changed_;
  mh_scene* scene_;
  int prim_kind_;
  int workers_;  // numProcs the scene's isLargeSubtree bits were set for (0: none yet)
  std::unique_ptr<PacketBatcher> batcher_;
  int max_packets_, wait_us_;

  std::mutex barrier_mu_;
  std::condition_variable barrier_cv_;
  int barrier_count_;
  unsigned barrier_generation_;
};
}  // namespace Manta

#endif
