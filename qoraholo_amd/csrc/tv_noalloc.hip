// The TV prox on device memory with no allocation: the FGP dual lives in a fixed array of
// device memory that the library owns for the life of the process, so the call makes no
// hipMalloc or hipFree (for callers that cannot allocate: code under graph capture, serving
// loops, harnesses that forbid it).
#include <mutex>

#include "../../include/qoraholo_hip.h"
#include "runtime.hpp"
#include "status.hpp"
#include "tv.hpp"

namespace qh {

namespace {

// Four dual buffers of three components for volumes of up to QH_TV_NOALLOC_MAX_VOXELS voxels
// (96 MiB). Host code must not pass `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): it uses the address hipGetSymbolAddress returns for the
// current device.
constexpr std::size_t kScratchFloats = 12 * static_cast<std::size_t>(QH_TV_NOALLOC_MAX_VOXELS);
static_assert(kScratchFloats >= 4 * ((3 * static_cast<std::size_t>(QH_TV_NOALLOC_MAX_VOXELS) + 63) / 64 * 64),
              "four 256-byte aligned dual buffers fit");

__device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float* scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
float* device_scratch() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  require(dev >= 0 && dev < kMaxDevices, "tv_prox_dev_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratch_ptr[dev] = static_cast<float*>(p);
  }
  return scratch_ptr[dev];
}

}  // namespace

}  // namespace qh

extern "C" {

int qh_hip_tv_prox_dev_noalloc(const float* b, int nz, int ny, int nx, float lambda, int inner,
                               int fused, float* x, void* stream) {
  using namespace qh;
  return guarded([&] {
    require(b && x && nz >= 1 && ny >= 1 && nx >= 1 && lambda > 0 && inner >= 0,
            "tv_prox_dev_noalloc: bad arguments");
    require(static_cast<long long>(nz) * ny * nx <= QH_TV_NOALLOC_MAX_VOXELS,
            "tv_prox_dev_noalloc: volume of " + std::to_string(static_cast<long long>(nz) * ny * nx) +
                " voxels exceeds the fixed dual buffer of " +
                std::to_string(QH_TV_NOALLOC_MAX_VOXELS) + " voxels");
    require(nx <= 512, "tv_prox_dev_noalloc: rows up to 512 wide");
    require(TvState::workspace_floats(nz, ny, nx) <= kScratchFloats,
            "tv_prox_dev_noalloc: dual buffers exceed the scratch array");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    TvState state(nz, ny, nx, device_scratch());
    state.set_fused(fused != 0);
    state.reset(s);  // the dual starts at zero: the array keeps the last call's contents
    state.prox(b, x, lambda, inner, s);
  });
}

}  // extern "C"
