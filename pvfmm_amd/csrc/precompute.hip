// Device side of the operator precompute (the work of FMM_Pts::Initialize,
// include/fmm_pts.txx:223-331 and 453-764): kernel matrices (Kernel::BuildMatrix,
// include/kernel.txx:1059-1135) and the one-sided Jacobi SVD behind the check-to-equivalent
// pseudo-inverses (include/fmm_pts.txx:478-581), all FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_fmm.hpp"
#include "precompute.hpp"

namespace pvfmm_amd {

#define PRE_LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

// K(r) for r = x_trg - x_src with the kernel's 1/(4 pi) or 1/(8 pi) scale, in the formulas of
// the host precompute (correctly rounded 1 / sqrt in FP64).
template <int KIND>
__device__ __forceinline__ void pair_block(double r0, double r1, double r2, double* out) {
  constexpr double pi = 3.14159265358979323846;
  const double rr = r0 * r0 + r1 * r1 + r2 * r2;
  if (!(rr > 0)) {
#pragma unroll
    for (int i = 0; i < (KIND == 2 ? 9 : (KIND == 1 ? 3 : 1)); i++) out[i] = 0.0;
    return;
  }
  const double rinv = 1.0 / sqrt(rr);
  if constexpr (KIND == 0) {
    out[0] = rinv / (4 * pi);
  } else if constexpr (KIND == 1) {
    const double f = rinv * rinv * rinv / (4 * pi);
    out[0] = r0 * f;
    out[1] = r1 * f;
    out[2] = r2 * f;
  } else {
    const double r[3] = {r0, r1, r2};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) out[3 * i + j] = rinv * ((i == j) + r[i] * r[j] * rinv * rinv) / (8 * pi);
  }
}

// out[(s k0 + i) (n_trg k1) + t k1 + j] = K_ij(trg_t - src_s); one thread per (s, t)
template <int KIND>
__global__ __launch_bounds__(256) void k_build_matrix(const double* __restrict__ src, int n_src, const double* __restrict__ trg,
                                                      int n_trg, double* __restrict__ out) {
  constexpr int K0 = KIND == 2 ? 3 : 1, K1 = KIND == 0 ? 1 : 3;
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_src * n_trg) return;
  const int s = (int)(idx / n_trg), t = (int)(idx % n_trg);
  double blk[K0 * K1];
  pair_block<KIND>(trg[3 * t] - src[3 * s], trg[3 * t + 1] - src[3 * s + 1], trg[3 * t + 2] - src[3 * s + 2], blk);
  const int64_t row_len = (int64_t)n_trg * K1;
#pragma unroll
  for (int i = 0; i < K0; i++)
#pragma unroll
    for (int j = 0; j < K1; j++) out[((int64_t)s * K0 + i) * row_len + (int64_t)t * K1 + j] = blk[i * K1 + j];
}

void build_matrix_device(hipStream_t s, int kind, const double* src, int n_src, const double* trg, int n_trg, double* out) {
  const int64_t n = (int64_t)n_src * n_trg;
  if (n == 0) return;
  const dim3 g((unsigned)((n + 255) / 256));
  if (kind == 0)
    hipLaunchKernelGGL(k_build_matrix<0>, g, dim3(256), 0, s, src, n_src, trg, n_trg, out);
  else if (kind == 1)
    hipLaunchKernelGGL(k_build_matrix<1>, g, dim3(256), 0, s, src, n_src, trg, n_trg, out);
  else
    hipLaunchKernelGGL(k_build_matrix<2>, g, dim3(256), 0, s, src, n_src, trg, n_trg, out);
  PRE_LAUNCH_CHECK();
}

// One round of the parallel one-sided Jacobi sweep: block k orthogonalises the row pair at
// ring positions k and cp-1-k of the round-robin schedule (position 0 fixed, positions
// 1..cp-1 rotated by `round`), the schedule of the host pinv_jacobi.  G holds the columns of A
// as rows (cp x r), W the accumulated right rotations (cp x c).  A block that rotates stores
// 1 to *flag (plain stores of one value; the host clears it before each sweep).
__global__ __launch_bounds__(256) void k_jacobi_round(double* __restrict__ G, double* __restrict__ W, int r, int c, int cp,
                                                      int round, int* __restrict__ flag) {
  __shared__ double red[3][256];
  const int k = blockIdx.x, tid = threadIdx.x;
  auto at = [&](int pos) { return pos == 0 ? 0 : 1 + (((pos - 1 - round) % (cp - 1)) + (cp - 1)) % (cp - 1); };
  int p = at(k), q = at(cp - 1 - k);
  if (p >= c || q >= c) return;
  if (p > q) {
    const int tmp = p;
    p = q;
    q = tmp;
  }
  double* gp = G + (int64_t)p * r;
  double* gq = G + (int64_t)q * r;
  double a = 0, b = 0, g = 0;
  for (int i = tid; i < r; i += 256) {
    const double x = gp[i], y = gq[i];
    a = fma(x, x, a);
    b = fma(y, y, b);
    g = fma(x, y, g);
  }
  red[0][tid] = a;
  red[1][tid] = b;
  red[2][tid] = g;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      red[0][tid] += red[0][tid + w];
      red[1][tid] += red[1][tid + w];
      red[2][tid] += red[2][tid + w];
    }
    __syncthreads();
  }
  a = red[0][0];
  b = red[1][0];
  g = red[2][0];
  const double eps = 2.220446049250313080847e-16;
  if (g == 0 || fabs(g) <= eps * sqrt(a * b)) return;
  const double zeta = (b - a) / (2 * g);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
  const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
  for (int i = tid; i < r; i += 256) {
    const double x = gp[i], y = gq[i];
    gp[i] = cs * x - sn * y;
    gq[i] = sn * x + cs * y;
  }
  double* wp = W + (int64_t)p * c;
  double* wq = W + (int64_t)q * c;
  for (int i = tid; i < c; i += 256) {
    const double x = wp[i], y = wq[i];
    wp[i] = cs * x - sn * y;
    wq[i] = sn * x + cs * y;
  }
  if (tid == 0) flag[0] = 1;
}

__global__ void k_clear_flag(int* flag) {
  if (threadIdx.x == 0) flag[0] = 0;
}

int jacobi_sweeps_device(hipStream_t s, double* G, double* W, int r, int c, int cp, int* flag, int max_sweeps) {
  if (c < 2) return 0;
  int sweep = 0;
  for (; sweep < max_sweeps; sweep++) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag);
    PRE_LAUNCH_CHECK();
    for (int round = 0; round < cp - 1; round++) {
      hipLaunchKernelGGL(k_jacobi_round, dim3(cp / 2), dim3(256), 0, s, G, W, r, c, cp, round, flag);
      PRE_LAUNCH_CHECK();
    }
    int h = 0;
    HIP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (!h) return sweep + 1;
  }
  return sweep;
}

}  // namespace pvfmm_amd

namespace pvfmm_amd {

// G (cp x r) = A^T with zero padding rows, W (cp x c) = identity with zero padding rows
__global__ void k_jacobi_init(const double* __restrict__ A, int r, int c, int cp, double* __restrict__ G,
                              double* __restrict__ W) {
  const int64_t nG = (int64_t)cp * r, nW = (int64_t)cp * c;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nG + nW; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < nG) {
      const int j = (int)(i / r), k = (int)(i % r);
      G[i] = j < c ? A[(int64_t)k * c + j] : 0.0;
    } else {
      const int64_t w = i - nG;
      const int j = (int)(w / c), k = (int)(w % c);
      W[w] = (j == k) ? 1.0 : 0.0;
    }
  }
}

void jacobi_init_device(hipStream_t s, const double* A, int r, int c, int cp, double* G, double* W) {
  const int64_t n = (int64_t)cp * (r + c);
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_jacobi_init, dim3(blocks), dim3(256), 0, s, A, r, c, cp, G, W);
  PRE_LAUNCH_CHECK();
}

}  // namespace pvfmm_amd
