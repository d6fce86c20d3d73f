// chip_bp_noalloc: chip_bp with the part scratch of the split single-slice
// adjoint bp in a static device array instead of plan memory, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The array lives in the library's code object: every GPU the library runs on
// holds its own 16 MiB copy until the process ends. Host code reaches it only
// through the per-device address hipGetSymbolAddress returns (the array's
// name in host code is a stand-in, not a device address).
#include "common.hpp"

namespace chip {
namespace {

constexpr int64_t kNoAllocScratchFloats = int64_t(4) << 20;  // 16 MiB
constexpr int kMaxDevices = 64;

__device__ __align__(256) float scratch[kNoAllocScratchFloats];

std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
int scratch_address(int dev, float** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      set_error(std::string("chip_bp_noalloc: hipGetSymbolAddress(scratch) -> ") + hipGetErrorString(e));
      return CHIP_EHIP;
    }
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return CHIP_OK;
}

}  // namespace
}  // namespace chip

using namespace chip;

extern "C" int chip_bp_noalloc(const chip_plan* plan, const float* sino, const float* maps, int64_t maps_slice_stride,
                               float* vol, int batch, int mode, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && sino != nullptr, "vol/sino is NULL");
  CHIP_CHECK_ARG(batch >= 1 && batch <= 65535, "batch must be in [1, 65535]");
  CHIP_CHECK_ARG(mode >= CHIP_BP_SKIMAGE && mode <= CHIP_BP_ADJOINT, "unknown bp mode");
  CHIP_CHECK_ARG(maps_slice_stride >= 0, "maps_slice_stride must be >= 0");
  const Plan& p = plan->p;
  int dev = -1;
  CHIP_HIP(hipGetDevice(&dev));
  CHIP_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  CHIP_CHECK_ARG(dev == p.device, "the plan was created on another device");
  BpScratch scr;
  scr.tune = p.bp_parts_tune;
  if (bp_may_split(p, batch, mode, p.n_angles)) {
    // the part count chip_bp would use; the call is refused when those parts do not fit
    const int parts = bp_adjoint_parts(p, maps != nullptr, kBpMaxParts);
    const int64_t nn = static_cast<int64_t>(p.n) * p.n;
    if (parts > 0) {
      if (parts * nn > kNoAllocScratchFloats) {
        set_error("chip_bp_noalloc: " + std::to_string(parts) + " parts of " + std::to_string(p.n) + "^2 floats exceed " +
                  "the static scratch of " + std::to_string(kNoAllocScratchFloats) +
                  " floats (16 MiB); use chip_bp or pin fewer parts with chip_plan_tune(\"bp_parts\")");
        return CHIP_EINVAL;
      }
      int rc = scratch_address(dev, &scr.part);
      if (rc != CHIP_OK) return rc;
      scr.cap_elems = kNoAllocScratchFloats;
    }
  }
  launch_bp_sinogram(p, sino, maps, maps_slice_stride, vol, batch, mode, as_stream(stream), scr);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}
