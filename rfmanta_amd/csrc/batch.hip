// Entry of the persistent multi-batch trace (batch_kernel.h): argument checks and the
// choice of the launcher for the scene's triangle test and the node layout.
#include "batch_kernel.h"

namespace mh {

int launch_trace_batch(const DevScene& sc, const BatchLaunch& bl, void* stream) {
  if (bl.nb <= 0 || bl.nb > kMaxBatches) return (int)hipErrorInvalidValue;
  if (bl.chunk < 0 || bl.chunk % 64 != 0 || (bl.refill < 64 && bl.chunk == 0))
    return (int)hipErrorInvalidValue;
  if (bl.kinds < kKindMixed || bl.kinds > kKindAny) return (int)hipErrorInvalidValue;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (bl.layout == kLayoutSoa && !sc.pairs_soa) return (int)hipErrorInvalidValue;
  if (sc.prim_kind == 1)
    return bl.layout == kLayoutRef   ? launch_batch_wald_ref(sc, bl, s)
           : bl.layout == kLayoutSoa ? launch_batch_wald_soa(sc, bl, s)
                                     : launch_batch_wald_pair(sc, bl, s);
  return bl.layout == kLayoutRef   ? launch_batch_ks_ref(sc, bl, s)
         : bl.layout == kLayoutSoa ? launch_batch_ks_soa(sc, bl, s)
                                   : launch_batch_ks_pair(sc, bl, s);
}

}  // namespace mh
