// Device kernels of the operator precompute (implemented in precompute.hip).
#pragma once
#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace pvfmm_amd {

// Kernel::BuildMatrix (include/kernel.txx:1059-1135) on the device for the expansion kernels
// (kind 0 laplace potential, 1 laplace gradient, 2 stokes velocity): out[(s*k0 + i) *
// (n_trg*k1) + t*k1 + j] = K_ij(trg_t - src_s), FP64, device arrays (AoS coordinates).
void build_matrix_device(hipStream_t s, int kind, const double* src, int n_src, const double* trg, int n_trg, double* out);

// Sweeps of the parallel one-sided Jacobi SVD until no rotation happens (or max_sweeps): G holds
// the c columns of an r-row matrix as rows (cp x r, cp = c rounded up to even, padding rows
// zero), W the c x c right rotations (cp x c, starting from the identity), flag one device int.
// Returns the number of sweeps run.
int jacobi_sweeps_device(hipStream_t s, double* G, double* W, int r, int c, int cp, int* flag, int max_sweeps);
// G = A^T (A r x c row-major) and W = I in the layouts above.
void jacobi_init_device(hipStream_t s, const double* A, int r, int c, int cp, double* G, double* W);

}  // namespace pvfmm_amd
