// hsba_scratch.hip — the fixed device scratch of the allocation-free entry
// point (hsba_covariance_static).
//
// One file-scope device array per GPU, held for the life of the process.  In
// host code the array's name is only a host-side stand-in, so its device
// address is looked up with hipGetSymbolAddress once per device (with that
// device current) and cached; every use goes through that pointer.  Every
// call on one GPU shares the array: calls that overlap on one GPU race.
#include <mutex>

#include "hsba_kernels.h"

namespace hsba {

namespace {

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double *scratch_dev[kMaxDevices] = {};

}  // namespace

int scratch_region(double **ptr) {
  *ptr = nullptr;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    scratch_dev[dev] = static_cast<double *>(p);
  }
  *ptr = scratch_dev[dev];
  return 0;
}

}  // namespace hsba
