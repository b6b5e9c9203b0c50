// Shared definitions of the MI355X particle-FMM library (host side).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace pvfmm_amd {

constexpr int kMaxDepth = 30;     // tree depth limit of the reference (include/pvfmm_common.hpp:29)
constexpr double kRad0 = 1.05;    // inner surface radius factor (include/pvfmm_common.hpp:33)
constexpr double kRad1 = 2.95;    // outer surface radius factor (include/pvfmm_common.hpp:34)

// Error codes returned by every internal entry point and reported by PVFMMGetLastError().
enum ErrorCode : int {
  kOk = 0,
  kErrInvalidArgument = 1,
  kErrUnsupported = 2,
  kErrDevice = 3,
  kErrOutOfMemory = 4,
  kErrState = 5,
  kErrInternal = 6,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// Kernel descriptor: the point kernel (source->target) and the kernel used for the
// equivalent-density expansions.  Values of `id` follow enum PVFMMKernel (include/pvfmm.h:21-28).
struct KernelInfo {
  int id;            // 0 laplace_poten, 1 laplace_grad, 3 stokes_vel
  int k0, k1;        // source / target dims of the point kernel (s2t, m2t, l2t)
  int e0, e1;        // source / target dims of the expansion kernel (s2m, s2l, m2m, m2l, l2l)
  int pt_kind;       // device kernel of the point evaluation: 0 poten, 1 grad, 2 stokes vel
  int exp_kind;      // device kernel of the expansions
  double src_scal;   // scale-invariance exponents of the expansion kernel (source)
  double trg_scal;   //   ... (target):  K(2x) = 2^-(src+trg) K(x)
  double out_scal;   // exponent of the point kernel output (C-ABI rescaling, src/pvfmm-wrapper.cpp:522-541)
  const char* name;
};

inline KernelInfo kernel_info(int id) {
  switch (id) {
    // include/kernel.txx:1333-1344 (laplace_poten, laplace_grad), 1469-1473 (stokes_vel)
    case 0: return {0, 1, 1, 1, 1, 0, 0, 0.0, 1.0, 1.0, "laplace"};
    case 1: return {1, 1, 3, 1, 1, 1, 0, 0.0, 1.0, 2.0, "laplace_grad"};
    case 3: return {3, 3, 3, 3, 3, 2, 2, 0.0, 1.0, 1.0, "stokes_vel"};
    default: throw Error(kErrUnsupported, "kernel id " + std::to_string(id) + " is not supported (0, 1 and 3 are)");
  }
}

inline int n_surf(int m) { return 6 * (m - 1) * (m - 1) + 2; }

}  // namespace pvfmm_amd
