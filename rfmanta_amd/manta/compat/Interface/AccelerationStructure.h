// Subset of Manta's AccelerationStructure (Interface/AccelerationStructure.h:11-66).
#ifndef Manta_Interface_AccelerationStructure_h
#define Manta_Interface_AccelerationStructure_h

#include <Interface/Object.h>
#include <string>

namespace Manta {
class Group;

class AccelerationStructure : public Object {
 public:
  AccelerationStructure() : needToSaveFile(false) {}
  virtual void setGroup(Group* new_group) = 0;
  virtual void groupDirty() = 0;
  virtual Group* getGroup() const = 0;
  virtual void update(int proc = 0, int numProcs = 1) { rebuild(proc, numProcs); }
  virtual void rebuild(int proc = 0, int numProcs = 1) = 0;
  virtual void addToUpdateGraph(ObjectUpdateGraph* graph, ObjectUpdateGraphNode* parent) = 0;
  virtual bool buildFromFile(const std::string&) { return false; }
  virtual bool saveToFile(const std::string&) { return false; }
  virtual bool isParallel() const { return true; }

 protected:
  bool needToSaveFile;
  std::string saveFileName;
};
}  // namespace Manta
#endif
