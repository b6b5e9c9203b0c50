// gfx950 kernels of the particle FMM evaluation.
//
//  * P2P (U list, S2M/X onto check surfaces, L2T/W from equivalent surfaces): one wave64 per
//    target box; source tiles of 64 points staged in LDS as SoA; when a box has fewer than 64
//    targets the wave splits the source loop over 64/next_pow2(n_targets) lane groups and
//    combines them with a butterfly of __shfl_xor (wavefront reduction).  Per source box the
//    partial sum is added to an FP64 accumulator (the FP64 merge of the mixed-precision path).
//  * Row-gathered GEMM for S2M/L2T (check -> equivalent), M2M and L2L (8 children stored
//    contiguously so a parent row is one 8*K row of the child block).
//  * V list: pruned separable DFT of the m^3 surface lattice embedded in the (2m)^3 grid,
//    processed one x-frequency slice at a time in LDS; a target-stationary complex Hadamard
//    sum over the 26 parent offsets with the transfer blocks broadcast from the scalar cache;
//    the inverse DFT evaluates only the check-surface points.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace pvfmm_amd {

#define HIP_LAUNCH_CHECK()                                                       \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) throw_device_error(e_, __FILE__, __LINE__);            \
  } while (0)

[[noreturn]] void throw_device_error(hipError_t e, const char* file, int line);

template <typename Real>
__device__ __forceinline__ Real rsqrt_masked(Real r2);
template <>
__device__ __forceinline__ float rsqrt_masked<float>(float r2) {
  float y = __builtin_amdgcn_rsqf(r2);
  return r2 > 0.0f ? y : 0.0f;
}
template <>
__device__ __forceinline__ double rsqrt_masked<double>(double r2) {
  double y = rsqrt(r2);
  return r2 > 0.0 ? y : 0.0;
}

template <int KIND> struct KDims;
template <> struct KDims<0> { static constexpr int k0 = 1, k1 = 1; };
template <> struct KDims<1> { static constexpr int k0 = 1, k1 = 3; };
template <> struct KDims<2> { static constexpr int k0 = 3, k1 = 3; };

template <int KIND>
__host__ __device__ constexpr double kscale() { return KIND == 2 ? 0.039788735772973833942 : 0.079577471545947667884; }  // 1/(8pi), 1/(4pi)

// u += K(r) f with r = x_trg - x_src (include/kernel.txx:1276-1364)
template <int KIND, typename Real>
__device__ __forceinline__ void interact(const Real r0, const Real r1, const Real r2_, const Real* f, Real* u) {
  const Real rr = r0 * r0 + r1 * r1 + r2_ * r2_;
  const Real rinv = rsqrt_masked<Real>(rr);
  if constexpr (KIND == 0) {
    u[0] = fma(rinv, f[0], u[0]);
  } else if constexpr (KIND == 1) {
    const Real f3 = rinv * rinv * rinv * f[0];
    u[0] = fma(r0, f3, u[0]);
    u[1] = fma(r1, f3, u[1]);
    u[2] = fma(r2_, f3, u[2]);
  } else {
    const Real ip = (f[0] * r0 + f[1] * r1 + f[2] * r2_) * rinv * rinv;
    u[0] = fma(rinv, f[0] + r0 * ip, u[0]);
    u[1] = fma(rinv, f[1] + r1 * ip, u[1]);
    u[2] = fma(rinv, f[2] + r2_ * ip, u[2]);
  }
}

// ----------------------------------------------------------------------------- small kernels
template <typename Real>
__global__ void k_gather_density(int64_t n, int k0, const double* __restrict__ in, const int64_t* __restrict__ perm,
                                 double scal, Real* __restrict__ den, int64_t in_stride) {
  const int v = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = perm[i];
    for (int c = 0; c < k0; c++) den[(int64_t)v * k0 * n + c * n + i] = (Real)(in[v * in_stride + k0 * p + c] * scal);
  }
}

__global__ void k_scatter_output(int64_t n, int k1, const double* __restrict__ val, const int64_t* __restrict__ perm,
                                 double scal, double* __restrict__ out, int64_t out_stride) {
  const int v = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = perm[i];
    for (int c = 0; c < k1; c++) out[v * out_stride + k1 * p + c] = val[(int64_t)v * k1 * n + c * n + i] * scal;
  }
}

template <typename T>
__global__ void k_fill(T* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = T(0);
}

// ----------------------------------------------------------------------------- P2P: points -> points
// Coordinates are FP64 leaf-local offsets; the pair displacement is formed in FP64 and then
// rounded to Real, so that near-coincident pairs keep their digits in the FP32 evaluator.
template <typename Real, int KIND>
__global__ __launch_bounds__(64) void k_p2p_points(const int32_t* __restrict__ tleaf, const int32_t* __restrict__ item_ptr,
                                                   const int32_t* __restrict__ items, const int32_t* __restrict__ trg_begin,
                                                   const int32_t* __restrict__ trg_end, const int32_t* __restrict__ src_begin,
                                                   const int32_t* __restrict__ src_end, const double* __restrict__ center,
                                                   const double* __restrict__ tx, int64_t n_trg, const double* __restrict__ sx,
                                                   const Real* __restrict__ sden, int64_t n_src, double* __restrict__ out,
                                                   bool accumulate) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1;
  __shared__ double lc[3 * 64];
  __shared__ Real ld[K0 * 64];
  const int row = blockIdx.x, lane = threadIdx.x;
  const int T = tleaf[row];
  const int tb = trg_begin[T], nt = trg_end[T] - tb;
  const double ct0 = center[3 * T], ct1 = center[3 * T + 1], ct2 = center[3 * T + 2];
  const int ib = item_ptr[row], ie = item_ptr[row + 1];
  for (int c0 = 0; c0 < nt; c0 += 64) {
    const int ntc = min(64, nt - c0);
    int tper = 1;
    while (tper < ntc) tper <<= 1;
    const int nsplit = 64 / tper;
    const int t = lane & (tper - 1), j = lane / tper;
    const bool valid = t < ntc;
    double p0 = 0, p1 = 0, p2 = 0;
    if (valid) {
      p0 = tx[tb + c0 + t];
      p1 = tx[n_trg + tb + c0 + t];
      p2 = tx[2 * n_trg + tb + c0 + t];
    }
    double acc[K1];
#pragma unroll
    for (int k = 0; k < K1; k++) acc[k] = 0;
    for (int it = ib; it < ie; it++) {
      const int S = items[it];
      const int sb = src_begin[S], ns = src_end[S] - sb;
      const double o0 = center[3 * S] - ct0, o1 = center[3 * S + 1] - ct1, o2 = center[3 * S + 2] - ct2;
      Real part[K1];
#pragma unroll
      for (int k = 0; k < K1; k++) part[k] = 0;
      for (int s0 = 0; s0 < ns; s0 += 64) {
        const int nsc = min(64, ns - s0);
        __syncthreads();
        if (lane < nsc) {
          const int64_t g = sb + s0 + lane;
          lc[lane] = sx[g] + o0;
          lc[64 + lane] = sx[n_src + g] + o1;
          lc[128 + lane] = sx[2 * n_src + g] + o2;
#pragma unroll
          for (int c = 0; c < K0; c++) ld[c * 64 + lane] = sden[c * n_src + g];
        }
        __syncthreads();
        for (int s = j; s < nsc; s += nsplit) {
          Real f[K0];
#pragma unroll
          for (int c = 0; c < K0; c++) f[c] = ld[c * 64 + s];
          interact<KIND, Real>((Real)(p0 - lc[s]), (Real)(p1 - lc[64 + s]), (Real)(p2 - lc[128 + s]), f, part);
        }
      }
#pragma unroll
      for (int k = 0; k < K1; k++) acc[k] += (double)part[k];
    }
    for (int o = tper; o < 64; o <<= 1)
#pragma unroll
      for (int k = 0; k < K1; k++) acc[k] += __shfl_xor(acc[k], o);
    if (j == 0 && valid) {
#pragma unroll
      for (int k = 0; k < K1; k++) {
        double* d = &out[k * n_trg + tb + c0 + t];
        const double v = acc[k] * kscale<KIND>();
        *d = accumulate ? *d + v : v;
      }
    }
  }
}

// ----------------------------------------------------------------------------- P2P: surfaces -> points
template <typename Real, int KIND>
__global__ __launch_bounds__(64) void k_far_points(const int32_t* __restrict__ tleaf, const int32_t* __restrict__ item_ptr,
                                                   const int32_t* __restrict__ item_node, const int64_t* __restrict__ item_den,
                                                   const double* __restrict__ item_rad, const int32_t* __restrict__ trg_begin,
                                                   const int32_t* __restrict__ trg_end, const double* __restrict__ center,
                                                   const float* __restrict__ node_size, const Real* __restrict__ unit_surf,
                                                   int ns, const Real* __restrict__ tx, int64_t n_trg,
                                                   const Real* __restrict__ dens, double* __restrict__ out, bool accumulate) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1;
  __shared__ Real ls[(3 + K0) * 64];
  const int row = blockIdx.x, lane = threadIdx.x;
  const int T = tleaf[row];
  const int tb = trg_begin[T], nt = trg_end[T] - tb;
  const double ct0 = center[3 * T], ct1 = center[3 * T + 1], ct2 = center[3 * T + 2];
  const int ib = item_ptr[row], ie = item_ptr[row + 1];
  for (int c0 = 0; c0 < nt; c0 += 64) {
    const int ntc = min(64, nt - c0);
    int tper = 1;
    while (tper < ntc) tper <<= 1;
    const int nsplit = 64 / tper;
    const int t = lane & (tper - 1), j = lane / tper;
    const bool valid = t < ntc;
    Real p0 = 0, p1 = 0, p2 = 0;
    if (valid) {
      p0 = tx[tb + c0 + t];
      p1 = tx[n_trg + tb + c0 + t];
      p2 = tx[2 * n_trg + tb + c0 + t];
    }
    double acc[K1];
#pragma unroll
    for (int k = 0; k < K1; k++) acc[k] = 0;
    for (int it = ib; it < ie; it++) {
      const int Q = item_node[it];
      const Real* den = dens + item_den[it];
      const Real sc = (Real)(item_rad[it] * node_size[Q]);
      const Real o0 = (Real)(center[3 * Q] - ct0), o1 = (Real)(center[3 * Q + 1] - ct1), o2 = (Real)(center[3 * Q + 2] - ct2);
      Real part[K1];
#pragma unroll
      for (int k = 0; k < K1; k++) part[k] = 0;
      for (int s0 = 0; s0 < ns; s0 += 64) {
        const int nsc = min(64, ns - s0);
        __syncthreads();
        if (lane < nsc) {
          const int p = s0 + lane;
          ls[lane] = o0 + sc * unit_surf[3 * p];
          ls[64 + lane] = o1 + sc * unit_surf[3 * p + 1];
          ls[128 + lane] = o2 + sc * unit_surf[3 * p + 2];
#pragma unroll
          for (int c = 0; c < K0; c++) ls[(3 + c) * 64 + lane] = den[p * K0 + c];
        }
        __syncthreads();
        for (int s = j; s < nsc; s += nsplit) {
          Real f[K0];
#pragma unroll
          for (int c = 0; c < K0; c++) f[c] = ls[(3 + c) * 64 + s];
          interact<KIND, Real>(p0 - ls[s], p1 - ls[64 + s], p2 - ls[128 + s], f, part);
        }
      }
#pragma unroll
      for (int k = 0; k < K1; k++) acc[k] += (double)part[k];
    }
    for (int o = tper; o < 64; o <<= 1)
#pragma unroll
      for (int k = 0; k < K1; k++) acc[k] += __shfl_xor(acc[k], o);
    if (j == 0 && valid) {
#pragma unroll
      for (int k = 0; k < K1; k++) {
        double* d = &out[k * n_trg + tb + c0 + t];
        const double v = acc[k] * kscale<KIND>();
        *d = accumulate ? *d + v : v;
      }
    }
  }
}

// ----------------------------------------------------------------------------- P2P: points -> surfaces
template <typename Real, int KIND>
__global__ __launch_bounds__(64) void k_p2p_surface(const int32_t* __restrict__ tnode, const int32_t* __restrict__ item_ptr,
                                                    const int32_t* __restrict__ items, const int32_t* __restrict__ src_begin,
                                                    const int32_t* __restrict__ src_end, const double* __restrict__ center,
                                                    const float* __restrict__ node_size, Real rad,
                                                    const Real* __restrict__ unit_surf, int ns, const Real* __restrict__ sx,
                                                    const Real* __restrict__ sden, int64_t n_src, Real* __restrict__ out,
                                                    const int32_t* __restrict__ orow, int64_t ld, bool accumulate) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1;
  __shared__ Real ls[(3 + K0) * 64];
  const int row = blockIdx.x, lane = threadIdx.x;
  const int p = blockIdx.y * 64 + lane;
  const int T = tnode[row];
  const double ct0 = center[3 * T], ct1 = center[3 * T + 1], ct2 = center[3 * T + 2];
  const Real sc = rad * (Real)node_size[T];
  Real p0 = 0, p1 = 0, p2 = 0;
  if (p < ns) {
    p0 = sc * unit_surf[3 * p];
    p1 = sc * unit_surf[3 * p + 1];
    p2 = sc * unit_surf[3 * p + 2];
  }
  double acc[K1];
#pragma unroll
  for (int k = 0; k < K1; k++) acc[k] = 0;
  const int ib = item_ptr[row], ie = item_ptr[row + 1];
  for (int it = ib; it < ie; it++) {
    const int S = items[it];
    const int sb = src_begin[S], nsrc = src_end[S] - sb;
    const Real o0 = (Real)(center[3 * S] - ct0), o1 = (Real)(center[3 * S + 1] - ct1), o2 = (Real)(center[3 * S + 2] - ct2);
    Real part[K1];
#pragma unroll
    for (int k = 0; k < K1; k++) part[k] = 0;
    for (int s0 = 0; s0 < nsrc; s0 += 64) {
      const int nsc = min(64, nsrc - s0);
      __syncthreads();
      if (lane < nsc) {
        const int64_t g = sb + s0 + lane;
        ls[lane] = sx[g] + o0;
        ls[64 + lane] = sx[n_src + g] + o1;
        ls[128 + lane] = sx[2 * n_src + g] + o2;
#pragma unroll
        for (int c = 0; c < K0; c++) ls[(3 + c) * 64 + lane] = sden[c * n_src + g];
      }
      __syncthreads();
      for (int s = 0; s < nsc; s++) {
        Real f[K0];
#pragma unroll
        for (int c = 0; c < K0; c++) f[c] = ls[(3 + c) * 64 + s];
        interact<KIND, Real>(p0 - ls[s], p1 - ls[64 + s], p2 - ls[128 + s], f, part);
      }
    }
#pragma unroll
    for (int k = 0; k < K1; k++) acc[k] += (double)part[k];
  }
  if (p < ns) {
    Real* o = out + (int64_t)orow[row] * ld + p * K1;
#pragma unroll
    for (int k = 0; k < K1; k++) {
      const Real v = (Real)(acc[k] * kscale<KIND>());
      o[k] = accumulate ? o[k] + v : v;
    }
  }
}

// ----------------------------------------------------------------------------- row-gathered GEMM
template <typename Real, bool ACC>
__global__ __launch_bounds__(256) void k_gemm_rows(int M, int N, int K, const Real* __restrict__ A, const int32_t* __restrict__ ra,
                                                   int64_t lda, const Real* __restrict__ B, Real* __restrict__ C,
                                                   const int32_t* __restrict__ rc, int64_t ldc, const Real* __restrict__ alpha) {
  constexpr int BM = 64, BN = 64, BK = 16;
  __shared__ Real As[BK][BM + 1];
  __shared__ Real Bs[BK][BN + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  int64_t abase[4];
  bool aval[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int e = tid + 256 * q, r = e >> 4, gm = m0 + r;
    aval[q] = gm < M;
    abase[q] = aval[q] ? (int64_t)ra[gm] * lda : 0;
  }
  Real acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q, r = e >> 4, c = e & 15, gk = k0 + c;
      As[c][r] = (aval[q] && gk < K) ? A[abase[q] + gk] : Real(0);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q, r = e >> 6, c = e & 63, gk = k0 + r, gn = n0 + c;
      Bs[r][c] = (gk < K && gn < N) ? B[(int64_t)gk * N + gn] : Real(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      Real a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
    Real* crow = C + (int64_t)rc[gm] * ldc;
    const Real al = alpha ? alpha[gm] : Real(1);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int gn = n0 + tx * 4 + j;
      if (gn < N) crow[gn] = ACC ? fma(al, acc[i][j], crow[gn]) : al * acc[i][j];
    }
  }
}

// ----------------------------------------------------------------------------- V list: DFTs
// Twiddles w[k] = exp(sign * 2 pi i k / n1), kept in LDS
template <typename Real>
__device__ void init_twiddles(Real* twr, Real* twi, int n1, double sign) {
  for (int k = threadIdx.x; k < n1; k += blockDim.x) {
    double s, c;
    sincospi(2.0 * k / n1, &s, &c);
    twr[k] = (Real)c;
    twi[k] = (Real)(sign * s);
  }
}

// grid: n_slots * 8 * e0 blocks of 256 threads.  Per block one child grid: m^3 lattice
// (surface values only) -> n1 x n1 x nh spectrum written to X[slot][f][c*e0+i].
template <typename Real>
__global__ __launch_bounds__(256) void k_fft_up(const int32_t* __restrict__ first_child, const Real* __restrict__ up, int K0, int e0,
                                                int m, const int32_t* __restrict__ lattice, const Real* __restrict__ scal,
                                                Real* __restrict__ X) {
  extern __shared__ unsigned char smem_raw[];
  Real* sm = reinterpret_cast<Real*>(smem_raw);
  const int n1 = 2 * m, nh = m + 1, M_dim = n1 * n1 * nh, ns = 6 * (m - 1) * (m - 1) + 2;
  Real* cube = sm;                       // m^3
  Real* twr = cube + m * m * m;          // n1
  Real* twi = twr + n1;                  // n1
  Real* s1r = twi + n1;                  // m*m   [z][y]
  Real* s1i = s1r + m * m;
  Real* s2r = s1i + m * m;               // m*n1  [z][ky]
  Real* s2i = s2r + m * n1;
  const int b = blockIdx.x;
  const int slot = b / (8 * e0), rem = b % (8 * e0), c = rem / e0, i = rem % e0;
  const int child = first_child[slot] + c;
  const Real sc = scal[slot];
  init_twiddles(twr, twi, n1, -1.0);
  for (int q = threadIdx.x; q < m * m * m; q += blockDim.x) cube[q] = 0;
  __syncthreads();
  const Real* u = up + (int64_t)child * K0;
  for (int p = threadIdx.x; p < ns; p += blockDim.x) {
    const int x = lattice[3 * p], y = lattice[3 * p + 1], z = lattice[3 * p + 2];
    cube[(z * m + y) * m + x] = u[p * e0 + i] * sc;
  }
  __syncthreads();
  Real* Xs = X + (int64_t)slot * M_dim * (16 * e0) + 2 * (c * e0 + i);
  for (int kx = 0; kx < nh; kx++) {
    for (int o = threadIdx.x; o < m * m; o += blockDim.x) {  // x: real -> complex
      const Real* line = &cube[o * m];
      Real ar = 0, ai = 0;
      for (int x = 0; x < m; x++) {
        const int e = (kx * x) % n1;
        ar = fma(line[x], twr[e], ar);
        ai = fma(line[x], twi[e], ai);
      }
      s1r[o] = ar;
      s1i[o] = ai;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < m * n1; o += blockDim.x) {  // y
      const int z = o / n1, ky = o % n1;
      Real ar = 0, ai = 0;
      for (int y = 0; y < m; y++) {
        const int e = (ky * y) % n1;
        const Real vr = s1r[z * m + y], vi = s1i[z * m + y];
        ar += vr * twr[e] - vi * twi[e];
        ai += vr * twi[e] + vi * twr[e];
      }
      s2r[o] = ar;
      s2i[o] = ai;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < n1 * n1; o += blockDim.x) {  // z
      const int ky = o % n1, kz = o / n1;
      Real ar = 0, ai = 0;
      for (int z = 0; z < m; z++) {
        const int e = (kz * z) % n1;
        const Real vr = s2r[z * n1 + ky], vi = s2i[z * n1 + ky];
        ar += vr * twr[e] - vi * twi[e];
        ai += vr * twi[e] + vi * twr[e];
      }
      const int f = kx + nh * (ky + n1 * kz);
      Xs[(int64_t)f * (16 * e0)] = ar;
      Xs[(int64_t)f * (16 * e0) + 1] = ai;
    }
    __syncthreads();
  }
}

// grid: n_slots * 8 * e1 blocks of 256 threads.  Inverse DFT of one child spectrum,
// evaluated at the ns check-surface lattice points only; dn[child] += scal * value.
template <typename Real>
__global__ __launch_bounds__(256) void k_ifft_down(const int32_t* __restrict__ first_child, const Real* __restrict__ Y,
                                                   Real* __restrict__ dn, int K1, int e1, int m, const int32_t* __restrict__ lattice,
                                                   int ns, const Real* __restrict__ scal) {
  extern __shared__ unsigned char smem_raw[];
  Real* sm = reinterpret_cast<Real*>(smem_raw);
  const int n1 = 2 * m, nh = m + 1, M_dim = n1 * n1 * nh;
  Real* twr = sm;                 // n1
  Real* twi = twr + n1;
  Real* yr = twi + n1;            // n1*n1 [kz][ky]
  Real* yi = yr + n1 * n1;
  Real* s1r = yi + n1 * n1;       // n1*m  [ky][z]
  Real* s1i = s1r + n1 * m;
  Real* s2r = s1i + n1 * m;       // m*m   [y][z]
  Real* s2i = s2r + m * m;
  constexpr int MAXP = 8;
  const int b = blockIdx.x;
  const int slot = b / (8 * e1), rem = b % (8 * e1), c = rem / e1, j = rem % e1;
  const int child = first_child[slot] + c;
  init_twiddles(twr, twi, n1, 1.0);
  Real acc[MAXP];
#pragma unroll
  for (int q = 0; q < MAXP; q++) acc[q] = 0;
  const Real* Ys = Y + (int64_t)slot * M_dim * (16 * e1) + 2 * (c * e1 + j);
  __syncthreads();
  for (int kx = 0; kx < nh; kx++) {
    for (int o = threadIdx.x; o < n1 * n1; o += blockDim.x) {  // o = ky + n1*kz
      const int64_t f = kx + (int64_t)nh * o;
      yr[o] = Ys[f * (16 * e1)];
      yi[o] = Ys[f * (16 * e1) + 1];
    }
    __syncthreads();
    for (int o = threadIdx.x; o < n1 * m; o += blockDim.x) {  // z
      const int ky = o / m, z = o % m;
      Real ar = 0, ai = 0;
      for (int kz = 0; kz < n1; kz++) {
        const int e = (kz * z) % n1;
        const Real vr = yr[ky + n1 * kz], vi = yi[ky + n1 * kz];
        ar += vr * twr[e] - vi * twi[e];
        ai += vr * twi[e] + vi * twr[e];
      }
      s1r[o] = ar;
      s1i[o] = ai;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < m * m; o += blockDim.x) {  // y
      const int y = o / m, z = o % m;
      Real ar = 0, ai = 0;
      for (int ky = 0; ky < n1; ky++) {
        const int e = (ky * y) % n1;
        const Real vr = s1r[ky * m + z], vi = s1i[ky * m + z];
        ar += vr * twr[e] - vi * twi[e];
        ai += vr * twi[e] + vi * twr[e];
      }
      s2r[o] = ar;
      s2i[o] = ai;
    }
    __syncthreads();
    const Real wk = (kx == 0 || 2 * kx == n1) ? Real(1) : Real(2);
#pragma unroll
    for (int q = 0; q < MAXP; q++) {
      const int p = threadIdx.x + q * blockDim.x;
      if (p < ns) {
        const int x = lattice[3 * p], y = lattice[3 * p + 1], z = lattice[3 * p + 2];
        const int e = (kx * x) % n1;
        acc[q] += wk * (s2r[y * m + z] * twr[e] - s2i[y * m + z] * twi[e]);
      }
    }
    __syncthreads();
  }
  const Real sc = scal[slot];
  Real* d = dn + (int64_t)child * K1;
#pragma unroll
  for (int q = 0; q < MAXP; q++) {
    const int p = threadIdx.x + q * blockDim.x;
    if (p < ns) d[p * e1 + j] += acc[q] * sc;
  }
}

// ----------------------------------------------------------------------------- V list: Hadamard
// One wave = 64 target slots at one frequency; blocks are assigned so that all blocks of a
// frequency run on the same XCD (block b lands on XCD b % 8), keeping X[.][f] in that L2.
template <typename Real, int E0, int E1>
__global__ __launch_bounds__(64) void k_hadamard(int n_slots, int M_dim, int n_tblk, const int32_t* __restrict__ partner,
                                                 const int32_t* __restrict__ offsets, int n_off, const Real* __restrict__ V,
                                                 const Real* __restrict__ X, Real* __restrict__ Y) {
  const int b = blockIdx.x;
  const int xcd = b & 7, local = b >> 3;
  const int f = (local / n_tblk) * 8 + xcd;
  const int tblk = local % n_tblk;
  if (f >= M_dim) return;
  const int slot = tblk * 64 + threadIdx.x;
  if (slot >= n_slots) return;
  constexpr int NR = 8 * E1, NS = 8 * E0;
  Real yr[NR], yi[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) yr[r] = yi[r] = 0;
  for (int oi = 0; oi < n_off; oi++) {
    const int p = partner[(int64_t)slot * n_off + oi];
    if (p < 0) continue;
    const int R = offsets[oi];
    const Real* x = X + ((int64_t)p * M_dim + f) * (2 * NS);
    Real xr[NS], xi[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      xr[s] = x[2 * s];
      xi[s] = x[2 * s + 1];
    }
    const Real* v = V + ((int64_t)R * M_dim + f) * (2 * NR * NS);
#pragma unroll
    for (int r = 0; r < NR; r++) {
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const Real vr = v[2 * (r * NS + s)], vi = v[2 * (r * NS + s) + 1];
        yr[r] = fma(vr, xr[s], fma(-vi, xi[s], yr[r]));
        yi[r] = fma(vr, xi[s], fma(vi, xr[s], yi[r]));
      }
    }
  }
  Real* y = Y + ((int64_t)slot * M_dim + f) * (2 * NR);
#pragma unroll
  for (int r = 0; r < NR; r++) {
    y[2 * r] = yr[r];
    y[2 * r + 1] = yi[r];
  }
}

// ----------------------------------------------------------------------------- launchers
static inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 65535 * 16) g = 65535 * 16;
  if (g < 1) g = 1;
  return (int)g;
}

template <typename Real>
void DeviceKernels<Real>::gather_density(hipStream_t s, int64_t n, int k0, const double* in, const int64_t* perm,
                                         double scal, Real* den, int n_vec, int64_t in_stride) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_density<Real>, dim3(grid_for(n, 256), n_vec), dim3(256), 0, s, n, k0, in, perm, scal, den, in_stride);
  HIP_LAUNCH_CHECK();
}
template <typename Real>
void DeviceKernels<Real>::scatter_output(hipStream_t s, int64_t n, int k1, const double* val, const int64_t* perm,
                                         double scal, double* out, int n_vec, int64_t out_stride) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scatter_output, dim3(grid_for(n, 256), n_vec), dim3(256), 0, s, n, k1, val, perm, scal, out, out_stride);
  HIP_LAUNCH_CHECK();
}
template <typename Real>
void DeviceKernels<Real>::fill(hipStream_t s, Real* p, int64_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill<Real>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
  HIP_LAUNCH_CHECK();
}
template <typename Real>
void DeviceKernels<Real>::fill64(hipStream_t s, double* p, int64_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::p2p_surface(hipStream_t s, int kind, int n_rows, const int32_t* tnode, const int32_t* item_ptr,
                                      const int32_t* items, const int32_t* src_begin, const int32_t* src_end,
                                      const double* center, const float* node_size, double rad, const Real* unit_surf, int ns,
                                      const Real* sx, const Real* sden, int64_t n_src, Real* out, const int32_t* orow,
                                      int64_t ld, bool accumulate) {
  if (n_rows == 0) return;
  dim3 g(n_rows, (ns + 63) / 64);
  if (kind == 0)
    hipLaunchKernelGGL((k_p2p_surface<Real, 0>), g, dim3(64), 0, s, tnode, item_ptr, items, src_begin, src_end, center, node_size,
                       (Real)rad, unit_surf, ns, sx, sden, n_src, out, orow, ld, accumulate);
  else if (kind == 2)
    hipLaunchKernelGGL((k_p2p_surface<Real, 2>), g, dim3(64), 0, s, tnode, item_ptr, items, src_begin, src_end, center, node_size,
                       (Real)rad, unit_surf, ns, sx, sden, n_src, out, orow, ld, accumulate);
  else
    throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::p2p_points(hipStream_t s, int kind, int n_rows, const int32_t* tleaf, const int32_t* item_ptr,
                                     const int32_t* items, const int32_t* trg_begin, const int32_t* trg_end,
                                     const int32_t* src_begin, const int32_t* src_end, const double* center, const double* tx,
                                     int64_t n_trg, const double* sx, const Real* sden, int64_t n_src, double* out,
                                     bool accumulate) {
  if (n_rows == 0) return;
#define P2P_ARGS tleaf, item_ptr, items, trg_begin, trg_end, src_begin, src_end, center, tx, n_trg, sx, sden, n_src, out, accumulate
  if (kind == 0) hipLaunchKernelGGL((k_p2p_points<Real, 0>), dim3(n_rows), dim3(64), 0, s, P2P_ARGS);
  else if (kind == 1) hipLaunchKernelGGL((k_p2p_points<Real, 1>), dim3(n_rows), dim3(64), 0, s, P2P_ARGS);
  else hipLaunchKernelGGL((k_p2p_points<Real, 2>), dim3(n_rows), dim3(64), 0, s, P2P_ARGS);
#undef P2P_ARGS
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::far_points(hipStream_t s, int kind, int n_rows, const int32_t* tleaf, const int32_t* item_ptr,
                                     const int32_t* item_node, const int64_t* item_den, const double* item_rad,
                                     const int32_t* trg_begin, const int32_t* trg_end, const double* center,
                                     const float* node_size, const Real* unit_surf, int ns, const Real* tx, int64_t n_trg,
                                     const Real* dens, double* out, bool accumulate) {
  if (n_rows == 0) return;
#define FAR_ARGS tleaf, item_ptr, item_node, item_den, item_rad, trg_begin, trg_end, center, node_size, unit_surf, ns, tx, n_trg, dens, out, accumulate
  if (kind == 0) hipLaunchKernelGGL((k_far_points<Real, 0>), dim3(n_rows), dim3(64), 0, s, FAR_ARGS);
  else if (kind == 1) hipLaunchKernelGGL((k_far_points<Real, 1>), dim3(n_rows), dim3(64), 0, s, FAR_ARGS);
  else hipLaunchKernelGGL((k_far_points<Real, 2>), dim3(n_rows), dim3(64), 0, s, FAR_ARGS);
#undef FAR_ARGS
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::gemm_rows(hipStream_t s, int M, int N, int K, const Real* A, const int32_t* ra, int64_t lda,
                                    const Real* B, Real* C, const int32_t* rc, int64_t ldc, const Real* alpha, bool accumulate) {
  if (M == 0) return;
  dim3 g((N + 63) / 64, (M + 63) / 64);
  if (accumulate) hipLaunchKernelGGL((k_gemm_rows<Real, true>), g, dim3(256), 0, s, M, N, K, A, ra, lda, B, C, rc, ldc, alpha);
  else hipLaunchKernelGGL((k_gemm_rows<Real, false>), g, dim3(256), 0, s, M, N, K, A, ra, lda, B, C, rc, ldc, alpha);
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::fft_up(hipStream_t s, int n_slots, const int32_t* first_child, const Real* up, int K0, int e0, int m,
                                 const int32_t* lattice, const Real* scal, Real* X) {
  if (n_slots == 0) return;
  const int n1 = 2 * m;
  size_t lds = sizeof(Real) * (m * m * m + 2 * n1 + 2 * m * m + 2 * m * n1);
  hipLaunchKernelGGL(k_fft_up<Real>, dim3(n_slots * 8 * e0), dim3(256), lds, s, first_child, up, K0, e0, m, lattice, scal, X);
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::hadamard(hipStream_t s, int n_slots, int M_dim, int e0, int e1, const int32_t* partner,
                                   const int32_t* offsets, int n_off, const Real* V, const Real* X, Real* Y) {
  if (n_slots == 0) return;
  const int n_tblk = (n_slots + 63) / 64;
  const int64_t nb = (int64_t)n_tblk * ((M_dim + 7) / 8) * 8;
  if (e0 == 1 && e1 == 1)
    hipLaunchKernelGGL((k_hadamard<Real, 1, 1>), dim3((unsigned)nb), dim3(64), 0, s, n_slots, M_dim, n_tblk, partner, offsets, n_off, V, X, Y);
  else if (e0 == 3 && e1 == 3)
    hipLaunchKernelGGL((k_hadamard<Real, 3, 3>), dim3((unsigned)nb), dim3(64), 0, s, n_slots, M_dim, n_tblk, partner, offsets, n_off, V, X, Y);
  else
    throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::ifft_down(hipStream_t s, int n_slots, const int32_t* first_child, const Real* Y, Real* dn, int K1, int e1,
                                    int m, const int32_t* lattice, int ns, const Real* scal) {
  if (n_slots == 0) return;
  if (ns > 8 * 256) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  const int n1 = 2 * m;
  size_t lds = sizeof(Real) * (2 * n1 + 2 * n1 * n1 + 2 * n1 * m + 2 * m * m);
  hipLaunchKernelGGL(k_ifft_down<Real>, dim3(n_slots * 8 * e1), dim3(256), lds, s, first_child, Y, dn, K1, e1, m, lattice, ns, scal);
  HIP_LAUNCH_CHECK();
}

template struct DeviceKernels<float>;
template struct DeviceKernels<double>;

}  // namespace pvfmm_amd
