// Batch-kernel launchers for Wald triangles over the REF node layout.
#include "batch_kernel.h"

namespace mh {

int launch_batch_wald_ref(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  return launch_b<1, kLayoutRef>(sc, bl, s);
}

}  // namespace mh
