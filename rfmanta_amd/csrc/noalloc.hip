// mh_trace_noalloc (include/mh.h): mh_trace for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops that forbid hipMalloc). The
// single-ray kernel's only scratch buffer, its per-lane traversal stack spill, comes from
// a fixed device array instead of the scene's lazily grown per-stream buffer. Nothing
// here calls hipMalloc, hipFree or an allocating library routine.
//
// Every call on one device shares that device's array and leaves its last contents; the
// spill is written before it is read, so nothing needs clearing. Calls that overlap on
// one device (two streams, two threads) race on the array.

#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"
#include "mh.h"

namespace {

// 32 MiB: the spill of a tree 64 levels deeper than the LDS ring at one 256-lane
// workgroup per compute unit of a 256-CU device, or of the 1M-triangle heightfield's
// tree (depth 22, ring 8) at more than four workgroups per compute unit.
constexpr size_t kScratchBytes = size_t(32) << 20;
static_assert(kScratchBytes == MH_NOALLOC_SCRATCH_BYTES, "header and source disagree");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
struct ScratchSlot {
  void* ptr = nullptr;          // device address of `scratch` on this device
  hipError_t err = hipSuccess;  // result of the lookup
  bool looked_up = false;
};
std::mutex g_mu;
ScratchSlot g_slots[kMaxDevices];

}  // namespace

void mh_internal::scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  ScratchSlot& s = g_slots[device];
  if (s.looked_up) return;
  void* p = nullptr;
  s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  s.ptr = s.err == hipSuccess ? p : nullptr;
  s.looked_up = true;
}

extern "C" int mh_trace_noalloc(const mh_scene* scene, const mh_rays* rays, const mh_hits* hits,
                                int64_t n, uint32_t flags, const mh_trace_opts* opts,
                                void* stream) {
  int rc = mh_internal::ensure_init();
  if (rc) return rc;
  int device = -1;
  HIP_OK(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices)
    return mh_internal::fail(MH_E_UNSUPPORTED, "device %d beyond the scratch table", device);
  void* ptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    const ScratchSlot& s = g_slots[device];
    if (!s.looked_up) return mh_internal::fail(MH_E_STATE, "scratch area of device %d not set up", device);
    if (s.err != hipSuccess)
      return mh_internal::fail(MH_E_HIP, "hipGetSymbolAddress(scratch): %s",
                               hipGetErrorString(s.err));
    ptr = s.ptr;
  }
  return mh_internal::trace_fixed_spill(scene, rays, hits, n, flags, opts, stream, ptr,
                                        kScratchBytes);
}
