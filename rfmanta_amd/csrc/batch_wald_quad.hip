// Batch-kernel launchers for Wald triangles over the QUAD node layout.
#include "batch_kernel.h"

namespace mh {

int launch_batch_wald_quad(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  return launch_b<1, kLayoutQuad>(sc, bl, s);
}

}  // namespace mh
