// The FP64 direct sum with its chunk partial sums in a static device array instead of an
// allocated buffer, for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops).  The kernels are those of direct_sum (kernels.hip); only the home
// of the partials differs, so the result equals direct_sum's bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

namespace pvfmm_amd {

namespace {

// One copy per device, kept until the process ends.  Host code never passes `scratch` itself
// to a kernel: on the host the name stands for a host-side shadow, not a device address; the
// device address comes from hipGetSymbolAddress, once per device.
__device__ __align__(256) double scratch[kDirectScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

double* scratch_on_current_device() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw Error(kErrDevice, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) throw Error(kErrDevice, "device ordinal out of range for the static scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw Error(kErrDevice, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}

}  // namespace

void direct_scratch_init() { (void)scratch_on_current_device(); }

void direct_sum_static(hipStream_t s, int kind, const double* src, int64_t n_src, const double* den, const double* trg,
                       int64_t n_trg, double* out) {
  const DirectSplit d = direct_split(kind, n_src, n_trg);
  if (d.scratch_bytes > kDirectScratchBytes)
    throw Error(kErrInvalidArgument, "direct sum needs " + std::to_string(d.scratch_bytes) +
                                         " bytes of partial sums, more than the " + std::to_string(kDirectScratchBytes) +
                                         "-byte static scratch");
  double* part = d.scratch_bytes ? scratch_on_current_device() : nullptr;
  direct_sum(s, kind, src, n_src, den, trg, n_trg, out, part);
}

}  // namespace pvfmm_amd
