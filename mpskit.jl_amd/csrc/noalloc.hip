// Allocation-free apply: mk_apply with the scratch taken from a fixed device
// array instead of the context's arena, for callers that cannot allocate.
#include <mutex>

#include "ops.h"

namespace {

constexpr size_t kNoallocBytes = (size_t)MK_NOALLOC_SCRATCH_MB << 20;
constexpr int kMaxDevices = 64;

}  // namespace

// The scratch array (one copy per GPU, held until the process ends).  In host
// code `scratch` only names the array; its device address is resolved once per
// device by noalloc_setup and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace mk {
namespace {
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

mk_status noalloc_setup(int device) {
  MK_REQUIRE(device >= 0 && device < kMaxDevices, MK_BAD_ARG, "noalloc: device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return MK_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    set_error("noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s", device,
              hipGetErrorString(e));
    return MK_HIP_ERROR;
  }
  g_ptr[device] = ptr;
  return MK_OK;
}

}  // namespace mk

using namespace mk;

extern "C" mk_status mk_apply_noalloc(mk_op h, const double* x, double* y, double alpha, double beta,
                                      int precision) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "apply_noalloc: null operator");
  Op* op = reinterpret_cast<Op*>(h);
  MK_REQUIRE(op->blocks.empty(), MK_NOT_SUPPORTED,
             "apply_noalloc: an operator with charge blocks builds its launch plan on first use (call mk_apply)");
  const int dev = op->ctx->device;
  MK_CHECK_HIP(hipSetDevice(dev));
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ptr = (dev >= 0 && dev < kMaxDevices) ? g_ptr[dev] : nullptr;
  }
  MK_REQUIRE(ptr != nullptr, MK_HIP_ERROR, "apply_noalloc: scratch array not set up on device %d", dev);
  FixedScratchGuard guard(ptr, kNoallocBytes);
  return op_apply(op, x, y, alpha, beta, precision);
}
