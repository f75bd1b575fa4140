// The fixed device scratch of the allocation-free entry point nrgk_broaden_noalloc (no counterpart in
// the reference): for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  The partial sums of the broadening live in a
// file-scope device array of NOALLOC_SCRATCH_DOUBLES doubles (4 MiB), which every GPU the library
// runs on holds from module load until the process ends.  Its device address is looked up once per
// device, with that device current, when a context is created on it (nrgk_ctx_create); host code
// uses only that address, never the array's name.  Every allocation-free call on a GPU shares its
// array: calls that overlap on one GPU (different streams, or different host threads) race.
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace nrgk {

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mutex;
static double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// with `device` current: the device address of its copy of the array, looked up on the first call
double* noalloc_scratch_setup(int device) {
  NRGK_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, NRGK_ERR_INVALID,
               "nrgk: device ordinal beyond the allocation-free scratch table");
  std::lock_guard<std::mutex> lk(g_noalloc_mutex);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    NRGK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NRGK_REQUIRE(p != nullptr, NRGK_ERR_HIP, "nrgk: no device address for the allocation-free scratch");
    g_noalloc_ptr[device] = static_cast<double*>(p);
  }
  return g_noalloc_ptr[device];
}

// the address looked up by noalloc_scratch_setup (nullptr before it ran on that device)
double* noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mutex);
  return g_noalloc_ptr[device];
}

}  // namespace nrgk
