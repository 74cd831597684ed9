// Batch-kernel launchers for Wald triangles over the SOA node layout.
#include "batch_kernel.h"

namespace mh {

int launch_batch_wald_soa(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  return launch_b<1, kLayoutSoa>(sc, bl, s);
}

}  // namespace mh
