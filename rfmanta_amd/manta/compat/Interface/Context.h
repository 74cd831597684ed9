// Subset of Manta's contexts (Interface/Context.h:157-232): the fields DynBVH reads.
#ifndef Manta_Interface_Context_h
#define Manta_Interface_Context_h

namespace Manta {
class MantaInterface;

class PreprocessContext {
 public:
  PreprocessContext() : manta_interface(0), proc(0), numProcs(1) {}
  PreprocessContext(MantaInterface* mi, int p, int n) : manta_interface(mi), proc(p), numProcs(n) {}
  void done() const {}
  bool isInitialized() const { return manta_interface != 0; }
  MantaInterface* manta_interface;
  int proc;
  int numProcs;
};

class RenderContext {
 public:
  RenderContext(int p, int n) : proc(p), numProcs(n) {}
  int proc;
  int numProcs;
};
}  // namespace Manta
#endif
