// Batch-kernel launchers for Wald triangles over the PAIR node layout.
#include "batch_kernel.h"

namespace mh {

int launch_batch_wald_pair(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  return launch_b<1, kLayoutPair>(sc, bl, s);
}

}  // namespace mh
