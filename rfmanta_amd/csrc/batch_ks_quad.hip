// Batch-kernel launchers for Kensler-Shirley triangles over the QUAD node layout.
#include "batch_kernel.h"

namespace mh {

int launch_batch_ks_quad(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  return launch_b<0, kLayoutQuad>(sc, bl, s);
}

}  // namespace mh
