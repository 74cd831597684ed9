// Batch-kernel launchers for Kensler-Shirley triangles over the SOA node layout.
#include "batch_kernel.h"

namespace mh {

int launch_batch_ks_soa(const DevScene& sc, const BatchLaunch& bl, hipStream_t s) {
  return launch_b<0, kLayoutSoa>(sc, bl, s);
}

}  // namespace mh
