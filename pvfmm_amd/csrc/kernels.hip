// gfx950 kernels of the particle FMM evaluation.
//
//  * P2P (U list, S2M/X onto check surfaces, L2T/W from equivalent surfaces): one 256-thread
//    block per target leaf or node; sources staged in LDS as SoA chunks; threads are
//    (target, split) pairs so that small leaves still fill the block; partial sums over at
//    most 64 sources are added into an FP64 accumulator (the FP64 merge of the
//    mixed-precision path), the splits of a target in a fixed order.
//  * Row-gathered GEMM for S2M/L2T (check -> equivalent), M2M and L2L (8 children stored
//    contiguously so a parent row is one 8*K row of the child block), on MFMA.
//  * V list: pruned separable DFT of the m^3 surface lattice embedded in the (2m)^3 grid,
//    one x-frequency slice at a time in LDS; a tiled frequency-domain Hadamard sum over the 26
//    parent offsets on the matrix cores (FP32 scalar kernels on the bf16 cores with exact
//    three-way operand splits); the inverse DFT evaluates only the check-surface points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

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

// The same for VB density vectors sharing the geometry (one rsqrt per pair); with VB = 1 it
// rounds exactly like interact.
template <int KIND, int VB, typename Real>
__device__ __forceinline__ void interact_v(const Real r0, const Real r1, const Real r2_, const Real (*f)[KDims<KIND>::k0],
                                           Real (*u)[KDims<KIND>::k1]) {
  const Real rr = r0 * r0 + r1 * r1 + r2_ * r2_;
  const Real rinv = rsqrt_masked<Real>(rr);
  if constexpr (KIND == 0) {
#pragma unroll
    for (int v = 0; v < VB; v++) u[v][0] = fma(rinv, f[v][0], u[v][0]);
  } else if constexpr (KIND == 1) {
    const Real rinv3 = rinv * rinv * rinv;
#pragma unroll
    for (int v = 0; v < VB; v++) {
      const Real f3 = rinv3 * f[v][0];
      u[v][0] = fma(r0, f3, u[v][0]);
      u[v][1] = fma(r1, f3, u[v][1]);
      u[v][2] = fma(r2_, f3, u[v][2]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < VB; v++) {
      const Real ip = (f[v][0] * r0 + f[v][1] * r1 + f[v][2] * r2_) * rinv * rinv;
      u[v][0] = fma(rinv, f[v][0] + r0 * ip, u[v][0]);
      u[v][1] = fma(rinv, f[v][1] + r1 * ip, u[v][1]);
      u[v][2] = fma(rinv, f[v][2] + r2_ * ip, u[v][2]);
    }
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
// the same with 16-byte stores: p 16-byte aligned, n16 = the number of 16-byte words
__global__ void k_fill16(uint4* p, int64_t n16) {
  const uint4 z = {0u, 0u, 0u, 0u};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x) p[i] = z;
}

// Periodic volume-potential correction: each block first sums the root's equivalent values per
// component (K0 of them, at most a few thousand, read through L2), then corrects its targets.
// Grid (blocks, nv); val SoA [j n + i] per vector, x SoA absolute coordinates.
template <typename Real>
__global__ __launch_bounds__(256) void k_vol_correction(int kind, int64_t n, const double* __restrict__ x,
                                                        const Real* __restrict__ eq_root, int K0, int e0,
                                                        double* __restrict__ val, int64_t eq_vs, int64_t val_vs) {
  __shared__ double part[3][256];
  const int v = blockIdx.y;
  const Real* eq = eq_root + v * eq_vs;
  double acc[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < K0; i += 256) acc[i % e0] += (double)eq[i];
  for (int c = 0; c < 3; c++) part[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; c++) part[c][threadIdx.x] += part[c][threadIdx.x + w];
    __syncthreads();
  }
  const double a0 = part[0][0], a1 = part[1][0], a2 = part[2][0];
  double* out = val + v * val_vs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double x0 = x[i], x1 = x[n + i], x2 = x[2 * n + i];
    const double r2 = x0 * x0 + x1 * x1 + x2 * x2;
    if (kind == 0) {
      out[i] += a0 * r2 * (1.0 / 6.0);
    } else {
      out[i] += a0 * (r2 - x0 * x0) * 0.25;
      out[n + i] += a1 * (r2 - x1 * x1) * 0.25;
      out[2 * n + i] += a2 * (r2 - x2 * x2) * 0.25;
    }
  }
}

// Transpose of k_vol_correction, first step: block b of vector v sums g_j(x_i) W_cj(x_i) over
// its targets (i = 256 b + t, stepping by the grid) into part[(v nb + b) 3 + c].
constexpr int kVolAdjBlocks = 1024;
template <typename Real>
__global__ __launch_bounds__(256) void k_vol_adjoint_partial(int kind, int64_t n, const double* __restrict__ x,
                                                             const Real* __restrict__ g, int64_t g_vs,
                                                             double* __restrict__ part) {
  __shared__ double red[3][256];
  const int v = blockIdx.y, nb = gridDim.x;
  const Real* gv = g + v * g_vs;
  double acc[3] = {0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x) {
    const double x0 = x[i], x1 = x[n + i], x2 = x[2 * n + i];
    const double r2 = x0 * x0 + x1 * x1 + x2 * x2;
    if (kind == 0) {
      acc[0] += (double)gv[i] * r2 * (1.0 / 6.0);
    } else {
      acc[0] += (double)gv[i] * (r2 - x0 * x0) * 0.25;
      acc[1] += (double)gv[n + i] * (r2 - x1 * x1) * 0.25;
      acc[2] += (double)gv[2 * n + i] * (r2 - x2 * x2) * 0.25;
    }
  }
  for (int c = 0; c < 3; c++) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[((int64_t)v * nb + blockIdx.x) * 3 + threadIdx.x] = red[threadIdx.x][0];
}
// second step (one block per vector): the block sums in block order, added to every surface
// point's component c of the root's equivalent density
template <typename Real>
__global__ __launch_bounds__(256) void k_vol_adjoint_apply(int nb, const double* __restrict__ part, int K0, int e0,
                                                           Real* __restrict__ eq, int64_t eq_vs) {
  __shared__ double tot[3];
  const int v = blockIdx.x;
  if (threadIdx.x < 3) {
    double t = 0;
    for (int b = 0; b < nb; b++) t += part[((int64_t)v * nb + b) * 3 + threadIdx.x];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K0; i += 256) eq[v * eq_vs + i] += (Real)tot[i % e0];
}

// ----------------------------------------------------------------------------- direct sum
// FP64 all-pairs evaluation (the reference's Kernel<T>::ker_poten, include/kernel.hpp:84).
// Block (b, c) takes targets [256 b, 256 b + 256), one per thread, and the sources of chunk c,
// [c chunk, (c + 1) chunk), staged through LDS 256 at a time.  With one chunk the block writes
// the scaled result; with several it writes its unscaled partial sums to
// part[c][k1 n_trg] and k_direct_reduce adds the chunks in chunk order.  AoS inputs and outputs.
template <int KIND>
__global__ __launch_bounds__(256) void k_direct(const double* __restrict__ src, int64_t n_src, const double* __restrict__ den,
                                                const double* __restrict__ trg, int64_t n_trg, int64_t chunk,
                                                double* __restrict__ out, double* __restrict__ part) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1;
  __shared__ double sc[3 * 256];
  __shared__ double sd[K0 * 256];
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  const bool valid = i < n_trg;
  const double tx = valid ? trg[3 * i] : 0.0, ty = valid ? trg[3 * i + 1] : 0.0, tz = valid ? trg[3 * i + 2] : 0.0;
  double u[K1];
#pragma unroll
  for (int c = 0; c < K1; c++) u[c] = 0.0;
  const int64_t s_begin = blockIdx.y * chunk, s_end = min(n_src, s_begin + chunk);
  for (int64_t j0 = s_begin; j0 < s_end; j0 += 256) {
    const int64_t j = j0 + threadIdx.x;
    const int nj = (int)min((int64_t)256, s_end - j0);
    __syncthreads();
    if (j < s_end) {
      for (int d = 0; d < 3; d++) sc[3 * threadIdx.x + d] = src[3 * j + d];
      for (int c = 0; c < K0; c++) sd[K0 * threadIdx.x + c] = den[K0 * j + c];
    }
    __syncthreads();
    for (int jj = 0; jj < nj; jj++)
      interact<KIND, double>(tx - sc[3 * jj], ty - sc[3 * jj + 1], tz - sc[3 * jj + 2], &sd[K0 * jj], u);
  }
  if (!valid) return;
  if (gridDim.y == 1) {
#pragma unroll
    for (int c = 0; c < K1; c++) out[K1 * i + c] = u[c] * kscale<KIND>();
  } else {
    double* p = part + (int64_t)blockIdx.y * K1 * n_trg + K1 * i;
#pragma unroll
    for (int c = 0; c < K1; c++) p[c] = u[c];
  }
}

// out[i] = scale * sum over chunks c = 0, 1, ... of part[c][i], i < nv
template <int KIND>
__global__ __launch_bounds__(256) void k_direct_reduce(int64_t nv, int n_chunks, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < n_chunks; c++) s += part[(int64_t)c * nv + i];
    out[i] = s * kscale<KIND>();
  }
}

DirectSplit direct_split(int kind, int64_t n_src, int64_t n_trg) {
  DirectSplit d;
  if (n_trg <= 0 || n_src <= 0) return d;
  const int k1 = kind == 0 ? 1 : 3;
  const int64_t tblocks = (n_trg + 255) / 256;
  // about 2048 blocks (8 per CU) when the targets alone give fewer; chunks of whole 256-source
  // stages
  const int64_t want = std::min<int64_t>((2048 + tblocks - 1) / tblocks, (n_src + 255) / 256);
  if (want <= 1) {
    d.chunk = n_src;
    return d;
  }
  d.chunk = ((n_src + want - 1) / want + 255) / 256 * 256;
  d.n_chunks = (int)((n_src + d.chunk - 1) / d.chunk);
  d.scratch_bytes = d.n_chunks > 1 ? (size_t)d.n_chunks * k1 * n_trg * sizeof(double) : 0;
  if (d.n_chunks == 1) d.chunk = n_src;
  return d;
}

void direct_sum(hipStream_t s, int kind, const double* src, int64_t n_src, const double* den, const double* trg, int64_t n_trg,
                double* out, double* part) {
  if (n_trg == 0) return;
  const DirectSplit d = direct_split(kind, n_src, n_trg);
  if (d.n_chunks > 1 && !part) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  const dim3 g((unsigned)((n_trg + 255) / 256), (unsigned)d.n_chunks);
  const int64_t nv = (int64_t)(kind == 0 ? 1 : 3) * n_trg;
  const unsigned gr = (unsigned)std::min<int64_t>((nv + 255) / 256, 4096);
#define PVFMM_DIRECT(KK)                                                                                     \
  hipLaunchKernelGGL(k_direct<KK>, g, dim3(256), 0, s, src, n_src, den, trg, n_trg, d.chunk, out, part);     \
  HIP_LAUNCH_CHECK();                                                                                        \
  if (d.n_chunks > 1) {                                                                                      \
    hipLaunchKernelGGL(k_direct_reduce<KK>, dim3(gr), dim3(256), 0, s, nv, d.n_chunks, (const double*)part, out); \
    HIP_LAUNCH_CHECK();                                                                                      \
  }
  if (kind == 0) {
    PVFMM_DIRECT(0)
  } else if (kind == 1) {
    PVFMM_DIRECT(1)
  } else {
    PVFMM_DIRECT(2)
  }
#undef PVFMM_DIRECT
}

// ----------------------------------------------------------------------------- P2P, 256-thread leaf blocks
// One target leaf per block of 256 threads.  Threads are (target t, split j) with
// tid = t + ntc * j, j < nsplit = 256 / ntc (ntc targets in the current chunk of at most 256);
// thread (t, j) sums the staged sources s = j, j + nsplit, ...  Sources are packed into LDS
// chunks of up to SC points across consecutive items (so a chunk usually holds several
// small leaves and one barrier pair serves them all); the partial sums of a chunk are added
// into an FP64 accumulator, and the nsplit partials of a target are added in split order at
// the end (deterministic).
template <typename Real, int KIND, int VB>
__global__ __launch_bounds__(256) void k_p2p_leaf(const int32_t* __restrict__ tleaf, const int32_t* __restrict__ item_ptr,
                                                  const int32_t* __restrict__ items, const int32_t* __restrict__ trg_begin,
                                                  const int32_t* __restrict__ trg_end, const int32_t* __restrict__ src_begin,
                                                  const int32_t* __restrict__ src_end, const double* __restrict__ center,
                                                  const double* __restrict__ tx, int64_t n_trg, const double* __restrict__ sx,
                                                  const Real* __restrict__ sden, int64_t n_src, int64_t den_vs,
                                                  double* __restrict__ out, int64_t out_vs, bool accumulate) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1, NT = 256, SC = VB * K0 >= 8 ? 256 : 512;
  __shared__ __align__(16) double lc[3 * SC];  // FP64 source coordinates in the target leaf's frame
  __shared__ __align__(16) Real ld[VB * K0 * SC];
  __shared__ double red[K1 * NT];
  const int row = blockIdx.x, tid = threadIdx.x;
  const int T = tleaf[row];
  const int tb = trg_begin[T], nt = trg_end[T] - tb;
  const double ct0 = center[3 * T], ct1 = center[3 * T + 1], ct2 = center[3 * T + 2];
  const int ib = item_ptr[row], ie = item_ptr[row + 1];
  for (int c0 = 0; c0 < nt; c0 += NT) {
    const int ntc = min(NT, nt - c0), nsplit = NT / ntc;
    const int t = tid % ntc, j = tid / ntc;
    const bool active = j < nsplit;
    double p0 = 0, p1 = 0, p2 = 0;
    if (active) {
      p0 = tx[tb + c0 + t];
      p1 = tx[n_trg + tb + c0 + t];
      p2 = tx[2 * n_trg + tb + c0 + t];
    }
    double acc[VB][K1];
#pragma unroll
    for (int v = 0; v < VB; v++)
#pragma unroll
      for (int k = 0; k < K1; k++) acc[v][k] = 0;
    int it = ib, off = 0;
    while (it < ie) {
      __syncthreads();  // the previous chunk is consumed
      int pos = 0;
      while (it < ie && pos < SC) {
        const int S = items[it];
        const int sb = src_begin[S], ns = src_end[S] - sb;
        const int take = min(ns - off, SC - pos);
        const double o0 = center[3 * S] - ct0, o1 = center[3 * S + 1] - ct1, o2 = center[3 * S + 2] - ct2;
        for (int q = tid; q < take; q += NT) {
          const int64_t g = sb + off + q;
          lc[pos + q] = sx[g] + o0;
          lc[SC + pos + q] = sx[n_src + g] + o1;
          lc[2 * SC + pos + q] = sx[2 * n_src + g] + o2;
#pragma unroll
          for (int v = 0; v < VB; v++)
#pragma unroll
            for (int c = 0; c < K0; c++) ld[(v * K0 + c) * SC + pos + q] = sden[v * den_vs + c * n_src + g];
        }
        pos += take;
        off += take;
        if (off == ns) {
          it++;
          off = 0;
        }
      }
      __syncthreads();
      if (active) {
        if constexpr (sizeof(Real) == 4 && VB == 1) {
          // split j takes the contiguous run [jb, je) of the chunk's sources (a multiple of 4
          // long, so that 4 sources are read with 16-byte LDS reads at immediate offsets); Real
          // partial sums over at most 64 sources, added into the FP64 accumulator
          const int L = ((pos + 4 * nsplit - 1) / (4 * nsplit)) * 4;
          const int jb = min(pos, j * L), je = min(pos, jb + L);
          for (int s0 = jb; s0 < je; s0 += 64) {
            Real part[VB][K1];
#pragma unroll
            for (int v = 0; v < VB; v++)
#pragma unroll
              for (int k = 0; k < K1; k++) part[v][k] = 0;
            const int s1 = min(je, s0 + 64);
            int s = s0;
            // s is a multiple of 4 here: the 4 sources' coordinates and densities are 16-byte
            // LDS words at immediate offsets from one base per array
            const double2* lc2 = reinterpret_cast<const double2*>(lc) + (s >> 1);
            const float4* ld4 = reinterpret_cast<const float4*>(ld) + (s >> 2);
            for (; s + 4 <= s1; s += 4, lc2 += 2, ld4 += 1) {
              double xs[3][4];
#pragma unroll
              for (int d = 0; d < 3; d++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                  const double2 w2 = lc2[(d * SC >> 1) + h];
                  xs[d][2 * h] = w2.x;
                  xs[d][2 * h + 1] = w2.y;
                }
              Real fs[VB][K0][4];
#pragma unroll
              for (int v = 0; v < VB; v++)
#pragma unroll
                for (int c = 0; c < K0; c++) {
                  const float4 q4 = ld4[(v * K0 + c) * SC >> 2];
                  fs[v][c][0] = q4.x;
                  fs[v][c][1] = q4.y;
                  fs[v][c][2] = q4.z;
                  fs[v][c][3] = q4.w;
                }
#pragma unroll
              for (int u = 0; u < 4; u++) {
                Real f[VB][K0];
#pragma unroll
                for (int v = 0; v < VB; v++)
#pragma unroll
                  for (int c = 0; c < K0; c++) f[v][c] = fs[v][c][u];
                interact_v<KIND, VB, Real>((Real)(p0 - xs[0][u]), (Real)(p1 - xs[1][u]), (Real)(p2 - xs[2][u]), f, part);
              }
            }
            for (; s < s1; s++) {
              Real f[VB][K0];
#pragma unroll
              for (int v = 0; v < VB; v++)
#pragma unroll
                for (int c = 0; c < K0; c++) f[v][c] = ld[(v * K0 + c) * SC + s];
              interact_v<KIND, VB, Real>((Real)(p0 - lc[s]), (Real)(p1 - lc[SC + s]), (Real)(p2 - lc[2 * SC + s]), f, part);
            }
#pragma unroll
            for (int v = 0; v < VB; v++)
#pragma unroll
              for (int k = 0; k < K1; k++) acc[v][k] += (double)part[v][k];
          }
        } else {
          // FP64 and several density vectors: split j takes sources j, j + nsplit, ... (the
          // unrolled form above needs more registers than this kernel's occupancy allows
          // there); partial sums over at most 64 sources, added into the accumulator
          for (int s0 = j; s0 < pos; s0 += 64 * nsplit) {
            Real part[VB][K1];
#pragma unroll
            for (int v = 0; v < VB; v++)
#pragma unroll
              for (int k = 0; k < K1; k++) part[v][k] = 0;
            const int s1 = min(pos, s0 + 64 * nsplit);
            for (int s = s0; s < s1; s += nsplit) {
              Real f[VB][K0];
#pragma unroll
              for (int v = 0; v < VB; v++)
#pragma unroll
                for (int c = 0; c < K0; c++) f[v][c] = ld[(v * K0 + c) * SC + s];
              interact_v<KIND, VB, Real>((Real)(p0 - lc[s]), (Real)(p1 - lc[SC + s]), (Real)(p2 - lc[2 * SC + s]), f, part);
            }
#pragma unroll
            for (int v = 0; v < VB; v++)
#pragma unroll
              for (int k = 0; k < K1; k++) acc[v][k] += (double)part[v][k];
          }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < VB; v++) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K1; k++) red[k * NT + tid] = active ? acc[v][k] : 0.0;
      __syncthreads();
      if (tid < ntc) {
#pragma unroll
        for (int k = 0; k < K1; k++) {
          double sum = 0;
          for (int jj = 0; jj < nsplit; jj++) sum += red[k * NT + tid + jj * ntc];
          double* d = &out[v * out_vs + k * n_trg + tb + c0 + tid];
          sum *= kscale<KIND>();
          *d = accumulate ? *d + sum : sum;
        }
      }
    }
  }
}

// Equivalent surfaces -> targets (own L2T surface and W-list M2T) with the leaf-block layout
// of k_p2p_leaf; a staged source is the surface point o + r * size * unit_surf[p] of its node.
template <typename Real, int KIND, int VB>
__global__ __launch_bounds__(256) void k_far_leaf(const int32_t* __restrict__ tleaf, const int32_t* __restrict__ item_ptr,
                                                  const int32_t* __restrict__ item_node, const int64_t* __restrict__ item_den,
                                                  const double* __restrict__ item_rad, const int32_t* __restrict__ trg_begin,
                                                  const int32_t* __restrict__ trg_end, const double* __restrict__ center,
                                                  const float* __restrict__ node_size, const Real* __restrict__ unit_surf,
                                                  int ns, const Real* __restrict__ tx, int64_t n_trg,
                                                  const Real* __restrict__ dens, int64_t den_vs, double* __restrict__ out,
                                                  int64_t out_vs, bool accumulate) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1, NT = 256, SC = VB * K0 >= 8 ? 256 : 512;
  __shared__ Real lc[3 * SC];
  __shared__ Real ld[VB * K0 * SC];
  __shared__ double red[K1 * NT];
  const int row = blockIdx.x, tid = threadIdx.x;
  const int T = tleaf[row];
  const int tb = trg_begin[T], nt = trg_end[T] - tb;
  const double ct0 = center[3 * T], ct1 = center[3 * T + 1], ct2 = center[3 * T + 2];
  const int ib = item_ptr[row], ie = item_ptr[row + 1];
  for (int c0 = 0; c0 < nt; c0 += NT) {
    const int ntc = min(NT, nt - c0), nsplit = NT / ntc;
    const int t = tid % ntc, j = tid / ntc;
    const bool active = j < nsplit;
    Real p0 = 0, p1 = 0, p2 = 0;
    if (active) {
      p0 = tx[tb + c0 + t];
      p1 = tx[n_trg + tb + c0 + t];
      p2 = tx[2 * n_trg + tb + c0 + t];
    }
    double acc[VB][K1];
#pragma unroll
    for (int v = 0; v < VB; v++)
#pragma unroll
      for (int k = 0; k < K1; k++) acc[v][k] = 0;
    int it = ib, off = 0;
    while (it < ie) {
      __syncthreads();
      int pos = 0;
      while (it < ie && pos < SC) {
        const int Q = item_node[it];
        const Real* den = dens + item_den[it];
        const int take = min(ns - off, SC - pos);
        const Real sc = (Real)(item_rad[it] * node_size[Q]);
        const Real o0 = (Real)(center[3 * Q] - ct0), o1 = (Real)(center[3 * Q + 1] - ct1), o2 = (Real)(center[3 * Q + 2] - ct2);
        for (int q = tid; q < take; q += NT) {
          const int pp = off + q;
          lc[pos + q] = o0 + sc * unit_surf[3 * pp];
          lc[SC + pos + q] = o1 + sc * unit_surf[3 * pp + 1];
          lc[2 * SC + pos + q] = o2 + sc * unit_surf[3 * pp + 2];
#pragma unroll
          for (int v = 0; v < VB; v++)
#pragma unroll
            for (int c = 0; c < K0; c++) ld[(v * K0 + c) * SC + pos + q] = den[v * den_vs + pp * K0 + c];
        }
        pos += take;
        off += take;
        if (off == ns) {
          it++;
          off = 0;
        }
      }
      __syncthreads();
      if (active) {
        for (int s0 = j; s0 < pos; s0 += 64 * nsplit) {
          Real part[VB][K1];
#pragma unroll
          for (int v = 0; v < VB; v++)
#pragma unroll
            for (int k = 0; k < K1; k++) part[v][k] = 0;
          const int s1 = min(pos, s0 + 64 * nsplit);
          for (int s = s0; s < s1; s += nsplit) {
            Real f[VB][K0];
#pragma unroll
            for (int v = 0; v < VB; v++)
#pragma unroll
              for (int c = 0; c < K0; c++) f[v][c] = ld[(v * K0 + c) * SC + s];
            interact_v<KIND, VB, Real>(p0 - lc[s], p1 - lc[SC + s], p2 - lc[2 * SC + s], f, part);
          }
#pragma unroll
          for (int v = 0; v < VB; v++)
#pragma unroll
            for (int k = 0; k < K1; k++) acc[v][k] += (double)part[v][k];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < VB; v++) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K1; k++) red[k * NT + tid] = active ? acc[v][k] : 0.0;
      __syncthreads();
      if (tid < ntc) {
#pragma unroll
        for (int k = 0; k < K1; k++) {
          double sum = 0;
          for (int jj = 0; jj < nsplit; jj++) sum += red[k * NT + tid + jj * ntc];
          double* d = &out[v * out_vs + k * n_trg + tb + c0 + tid];
          sum *= kscale<KIND>();
          *d = accumulate ? *d + sum : sum;
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------- P2P: points -> surfaces
// Sources (point ranges of the row's items, leaf-local coordinates shifted into the target
// node's frame) onto the ns surface points of the target node.  One 256-thread block per row
// (and per 256 PPT surface points); thread tid takes the points tid, tid + 256, ... of its
// block's range, PPT of them, so every staged source is read from LDS once for PPT points.
// Sources are packed into LDS chunks of up to 256 points across consecutive items, one
// barrier pair per chunk; partial sums over at most 64 sources are kept in Real and added
// into an FP64 accumulator.
template <typename Real, int KIND, int VB, int PPT>
__global__ __launch_bounds__(256) void k_p2p_surface(const int32_t* __restrict__ tnode, const int32_t* __restrict__ item_ptr,
                                                     const int32_t* __restrict__ items, const int32_t* __restrict__ src_begin,
                                                     const int32_t* __restrict__ src_end, const double* __restrict__ center,
                                                     const float* __restrict__ node_size, Real rad,
                                                     const Real* __restrict__ unit_surf, int ns, const Real* __restrict__ sx,
                                                     const Real* __restrict__ sden, int64_t n_src, int64_t den_vs,
                                                     Real* __restrict__ out, int64_t out_vs, const int32_t* __restrict__ orow,
                                                     int64_t ld, bool accumulate) {
  constexpr int K0 = KDims<KIND>::k0, K1 = KDims<KIND>::k1, NT = 256, SC = 256;
  __shared__ __align__(16) Real ls[(3 + VB * K0) * SC];
  const int row = blockIdx.x, tid = threadIdx.x;
  const int T = tnode[row];
  const double ct0 = center[3 * T], ct1 = center[3 * T + 1], ct2 = center[3 * T + 2];
  const Real sc = rad * (Real)node_size[T];
  Real px[PPT], py[PPT], pz[PPT];
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    const int p = (blockIdx.y * PPT + q) * NT + tid;
    const bool ok = p < ns;
    px[q] = ok ? sc * unit_surf[3 * p] : Real(0);
    py[q] = ok ? sc * unit_surf[3 * p + 1] : Real(0);
    pz[q] = ok ? sc * unit_surf[3 * p + 2] : Real(0);
  }
  double acc[PPT][VB][K1];
#pragma unroll
  for (int q = 0; q < PPT; q++)
#pragma unroll
    for (int v = 0; v < VB; v++)
#pragma unroll
      for (int k = 0; k < K1; k++) acc[q][v][k] = 0;
  const int ib = item_ptr[row], ie = item_ptr[row + 1];
  int it = ib, off = 0;
  while (it < ie) {
    __syncthreads();  // the previous chunk is consumed
    int pos = 0;
    while (it < ie && pos < SC) {
      const int S = items[it];
      const int sb = src_begin[S], nsrc = src_end[S] - sb;
      const int take = min(nsrc - off, SC - pos);
      const Real o0 = (Real)(center[3 * S] - ct0), o1 = (Real)(center[3 * S + 1] - ct1), o2 = (Real)(center[3 * S + 2] - ct2);
      for (int q = tid; q < take; q += NT) {
        const int64_t g = sb + off + q;
        ls[pos + q] = sx[g] + o0;
        ls[SC + pos + q] = sx[n_src + g] + o1;
        ls[2 * SC + pos + q] = sx[2 * n_src + g] + o2;
#pragma unroll
        for (int v = 0; v < VB; v++)
#pragma unroll
          for (int c = 0; c < K0; c++) ls[(3 + v * K0 + c) * SC + pos + q] = sden[v * den_vs + c * n_src + g];
      }
      pos += take;
      off += take;
      if (off == nsrc) {
        it++;
        off = 0;
      }
    }
    __syncthreads();
    for (int s0 = 0; s0 < pos; s0 += 64) {
      Real part[PPT][VB][K1];
#pragma unroll
      for (int q = 0; q < PPT; q++)
#pragma unroll
        for (int v = 0; v < VB; v++)
#pragma unroll
          for (int k = 0; k < K1; k++) part[q][v][k] = 0;
      const int s1 = min(pos, s0 + 64);
      for (int s = s0; s < s1; s++) {
        const Real x = ls[s], y = ls[SC + s], z = ls[2 * SC + s];
        Real f[VB][K0];
#pragma unroll
        for (int v = 0; v < VB; v++)
#pragma unroll
          for (int c = 0; c < K0; c++) f[v][c] = ls[(3 + v * K0 + c) * SC + s];
#pragma unroll
        for (int q = 0; q < PPT; q++) interact_v<KIND, VB, Real>(px[q] - x, py[q] - y, pz[q] - z, f, part[q]);
      }
#pragma unroll
      for (int q = 0; q < PPT; q++)
#pragma unroll
        for (int v = 0; v < VB; v++)
#pragma unroll
          for (int k = 0; k < K1; k++) acc[q][v][k] += (double)part[q][v][k];
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    const int p = (blockIdx.y * PPT + q) * NT + tid;
    if (p >= ns) continue;
#pragma unroll
    for (int v = 0; v < VB; v++) {
      Real* o = out + v * out_vs + (int64_t)orow[row] * ld + p * K1;
#pragma unroll
      for (int k = 0; k < K1; k++) {
        const Real val = (Real)(acc[q][v][k] * kscale<KIND>());
        o[k] = accumulate ? o[k] + val : val;
      }
    }
  }
}

// MFMA operand helpers.  gfx950 lane layouts (checked on the device): A 16x4 lane l holds
// (row l%16, k l/16); B 4x16 lane l holds (k l/16, col l%16); the 16x16 result holds
// rows 4*(l/16)+r (FP32) or l/16+4r (FP64), r = 0..3, at column l%16.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
template <typename Real> struct Mfma;
template <> struct Mfma<float> {
  using V4 = floatx4;
  static __device__ __forceinline__ V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <> struct Mfma<double> {
  using V4 = doublex4;
  static __device__ __forceinline__ V4 mma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};

// ----------------------------------------------------------------------------- row-gathered GEMM
// C[rc[i]*ldc + n] (+)= alpha[i] * sum_k A[ra[i]*lda + k] * B[k*N + n] on MFMA, for nv
// density vectors at once: the M = nv * Mv rows are (vector v, index row i), with A and C of
// vector v offset by v * avs and v * cvs.
// Block tile 64x64 (4 waves, each a 32x32 quadrant = 2x2 MFMA 16x16 tiles), k staged 16 at a
// time in LDS with k contiguous for both operands so each lane feeds its four k values of an
// MFMA group from one 4-vector read.  With ksplit > 1 the block handles one k range and writes
// its partial tile to work[z][i][n]; k_gemm_reduce then sums the partials in a fixed order.
template <typename Real, bool ACC, bool SPLIT, bool VEC>
__global__ __launch_bounds__(256, (sizeof(Real) == 4 ? 8 : 4)) void k_gemm_rows(int M, int N, int K, int kchunk, const Real* __restrict__ A,
                                                   const int32_t* __restrict__ ra, int64_t lda, const Real* __restrict__ B,
                                                   Real* __restrict__ C, const int32_t* __restrict__ rc, int64_t ldc,
                                                   const Real* __restrict__ alpha, Real* __restrict__ work, int Mv,
                                                   int64_t avs, int64_t cvs) {
  using MF = Mfma<Real>;
  using V4 = typename MF::V4;
  constexpr int BM = 64, BN = 64, BK = 32, LD = BK + 4;
  __shared__ __align__(16) Real As[BM * LD];
  __shared__ __align__(16) Real Bt[BN * LD];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int wm = w >> 1, wn = w & 1;
  // tiles in row-major order, cut into 8 contiguous ranges, one per XCD (block b runs on XCD
  // b mod 8): the column tiles of a row tile run together on one XCD and share its A rows in L2
  const int gx = (N + BN - 1) / BN, n_tiles = gx * ((M + BM - 1) / BM);
  const int per = (n_tiles + 7) / 8;
  const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (tile >= n_tiles) return;  // whole block
  const int m0 = (tile / gx) * BM, n0 = (tile % gx) * BN;
  const int kb0 = blockIdx.z * kchunk, kb1 = min(K, kb0 + kchunk);
  // global -> register -> LDS assignments: A: 8 consecutive k of one row; B: 8 consecutive n of one k
  const int ar = tid >> 2, ak = (tid & 3) * 8;
  const int bk = tid >> 3, bn = (tid & 7) * 8;
  const bool arow_ok = m0 + ar < M;
  const int av = arow_ok ? (m0 + ar) / Mv : 0, ai = arow_ok ? m0 + ar - av * Mv : 0;
  const Real* arow = A + (arow_ok ? av * avs + (int64_t)ra[ai] * lda : 0);
  Real ra_[8], rb_[8];
  // load the k-step starting at k0 into registers (zero outside the matrix)
  auto load = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = k0 + ak + 4 * h;
      if (VEC && arow_ok && k + 3 < kb1) {
        const V4 v = *reinterpret_cast<const V4*>(arow + k);
#pragma unroll
        for (int q = 0; q < 4; q++) ra_[4 * h + q] = v[q];
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) ra_[4 * h + q] = (arow_ok && k + q < kb1) ? arow[k + q] : Real(0);
      }
    }
    const int k = k0 + bk;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int n = n0 + bn + 4 * h;
      if (VEC && k < kb1 && n + 3 < N) {
        const V4 v = *reinterpret_cast<const V4*>(B + (int64_t)k * N + n);
#pragma unroll
        for (int q = 0; q < 4; q++) rb_[4 * h + q] = v[q];
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) rb_[4 * h + q] = (k < kb1 && n + q < N) ? B[(int64_t)k * N + n + q] : Real(0);
      }
    }
  };
  V4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = V4{0, 0, 0, 0};
  if (kb0 < kb1) load(kb0);
  for (int k0 = kb0; k0 < kb1; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 8; q++) As[ar * LD + ak + q] = ra_[q];
    // B transposed into LDS with its 4-wide k groups XOR-swizzled by n/8 (the 8 n-blocks of a
    // store hit different banks; a 4-vector read stays contiguous and aligned)
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int n = bn + q;
      Bt[n * LD + 4 * ((bk >> 2) ^ ((n >> 3) & 7)) + (bk & 3)] = rb_[q];
    }
    __syncthreads();
    if (k0 + BK < kb1) load(k0 + BK);  // in flight during the MFMAs of this step
#pragma unroll
    for (int kk = 0; kk < BK / 16; kk++) {
      V4 av[2], bv[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        av[t] = *reinterpret_cast<const V4*>(&As[(wm * 32 + t * 16 + (l & 15)) * LD + 16 * kk + 4 * (l >> 4)]);
        const int n = wn * 32 + t * 16 + (l & 15);
        bv[t] = *reinterpret_cast<const V4*>(&Bt[n * LD + 4 * ((4 * kk + (l >> 4)) ^ ((n >> 3) & 7))]);
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          acc[i][j] = MF::mma(av[i][0], bv[j][0], acc[i][j]);
          acc[i][j] = MF::mma(av[i][1], bv[j][1], acc[i][j]);
          acc[i][j] = MF::mma(av[i][2], bv[j][2], acc[i][j]);
          acc[i][j] = MF::mma(av[i][3], bv[j][3], acc[i][j]);
        }
    }
    __syncthreads();
  }
  if constexpr (!SPLIT && ACC) {
    // read-modify-write: every C value of the lane is loaded before any is stored, so the
    // loads are in flight together
    Real* cp[2][4];
    Real al[2][4], cold[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gm = m0 + wm * 32 + i * 16 + MF::row(l, r);
        const int cv = gm < M ? gm / Mv : 0, ci = gm < M ? gm - cv * Mv : 0;
        cp[i][r] = gm < M ? C + cv * cvs + (int64_t)rc[ci] * ldc : nullptr;
        al[i][r] = gm < M && alpha ? alpha[ci] : Real(1);
      }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gn = n0 + wn * 32 + j * 16 + (l & 15);
          cold[i][j][r] = (cp[i][r] && gn < N) ? cp[i][r][gn] : Real(0);
        }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gn = n0 + wn * 32 + j * 16 + (l & 15);
          if (cp[i][r] && gn < N) cp[i][r][gn] = fma(al[i][r], acc[i][j][r], cold[i][j][r]);
        }
    return;
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gm = m0 + wm * 32 + i * 16 + MF::row(l, r);
        const int gn = n0 + wn * 32 + j * 16 + (l & 15);
        if (gm >= M || gn >= N) continue;
        const Real v = acc[i][j][r];
        if (SPLIT) {
          work[((int64_t)blockIdx.z * M + gm) * N + gn] = v;
        } else {
          const int cv = gm / Mv, ci = gm - cv * Mv;
          Real* crow = C + cv * cvs + (int64_t)rc[ci] * ldc;
          const Real al = alpha ? alpha[ci] : Real(1);
          crow[gn] = al * v;
        }
      }
}

template <typename Real, bool ACC>
__global__ void k_gemm_reduce(int M, int N, int ksplit, const Real* __restrict__ work, Real* __restrict__ C,
                              const int32_t* __restrict__ rc, int64_t ldc, const Real* __restrict__ alpha, int Mv,
                              int64_t cvs) {
  const int64_t total = (int64_t)M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / N), n = (int)(e % N);
    // partials added in split order; 8 loads in flight at a time
    const Real* wp = work + (int64_t)i * N + n;
    const int64_t zs = (int64_t)M * N;
    Real v = 0;
    int z = 0;
    for (; z + 8 <= ksplit; z += 8) {
      Real t[8];
#pragma unroll
      for (int q = 0; q < 8; q++) t[q] = wp[(z + q) * zs];
#pragma unroll
      for (int q = 0; q < 8; q++) v += t[q];
    }
    for (; z < ksplit; z++) v += wp[z * zs];
    const int cv = i / Mv, ci = i - cv * Mv;
    Real* crow = C + cv * cvs + (int64_t)rc[ci] * ldc;
    const Real al = alpha ? alpha[ci] : Real(1);
    crow[n] = ACC ? fma(al, v, crow[n]) : al * v;
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

// Forward DFT of CB children of one source parent per block (grid: n_slots * e0 * (8/CB)
// blocks of 512 threads).  The m^3 lattice cube of each child (surface values only) sits in
// LDS; the transform runs one x-frequency slice at a time: x (real->complex, m taps), y and z
// (m taps each, the pruned inputs).  Spectra are written as X[slot][f][8i+c] complex so
// that the children of one frequency form one contiguous row, component-major.
template <typename Real>
__global__ __launch_bounds__(512) void k_fft_up(const int32_t* __restrict__ first_child, const Real* __restrict__ up, int K0, int e0,
                                                int m, int CB, const int32_t* __restrict__ lattice, const Real* __restrict__ scal,
                                                Real* __restrict__ X, int vg, int vi) {
  extern __shared__ unsigned char smem_raw[];
  Real* sm = reinterpret_cast<Real*>(smem_raw);
  const int n1 = 2 * m, nh = m + 1, M_dim = n1 * n1 * nh, ns = 6 * (m - 1) * (m - 1) + 2;
  const int m2 = m * m, m3 = m2 * m, mn = m * n1;
  Real* cube = sm;                 // CB * m^3
  Real* twr = cube + CB * m3;      // n1
  Real* twi = twr + n1;
  Real* s1r = twi + n1;            // CB * m*m   [c][z][y]
  Real* s1i = s1r + CB * m2;
  Real* s2r = s1i + CB * m2;       // CB * m*n1  [c][z][ky]
  Real* s2i = s2r + CB * mn;
  const int nblk = 8 / CB;
  const int slot = blockIdx.x / (e0 * nblk), rem = blockIdx.x % (e0 * nblk), i = rem / nblk, c0 = (rem % nblk) * CB;
  const int fc = first_child[slot] + c0;
  const Real sc = scal[slot];
  const int tid = threadIdx.x, nt = blockDim.x;
  init_twiddles(twr, twi, n1, -1.0);
  for (int q = tid; q < CB * m3; q += nt) cube[q] = 0;
  __syncthreads();
  for (int q = tid; q < CB * ns; q += nt) {
    const int c = q / ns, p = q % ns;
    const int x = lattice[3 * p], y = lattice[3 * p + 1], z = lattice[3 * p + 2];
    cube[c * m3 + (z * m + y) * m + x] = up[(int64_t)(fc + c) * K0 + p * e0 + i] * sc;
  }
  __syncthreads();
  Real* Xs = X + (((int64_t)slot * M_dim * vg + vi) * 8 * e0 + 8 * i + c0) * 2;
  for (int kx = 0; kx < nh; kx++) {
    for (int o = tid; o < CB * m2; o += nt) {  // x: real -> complex
      const Real* line = &cube[(o / m2) * m3 + (o % m2) * m];
      Real ar = 0, ai = 0;
      for (int x = 0, e = 0; x < m; x++, e = (e + kx >= n1) ? e + kx - n1 : e + kx) {
        ar = fma(line[x], twr[e], ar);
        ai = fma(line[x], twi[e], ai);
      }
      s1r[o] = ar;
      s1i[o] = ai;
    }
    __syncthreads();
    for (int o = tid; o < CB * mn; o += nt) {  // y
      const int c = o / mn, r = o % mn, z = r / n1, ky = r % n1;
      const Real* vr = &s1r[c * m2 + z * m];
      const Real* vi = &s1i[c * m2 + z * m];
      Real ar = 0, ai = 0;
      for (int y = 0, e = 0; y < m; y++, e = (e + ky >= n1) ? e + ky - n1 : e + ky) {
        ar += vr[y] * twr[e] - vi[y] * twi[e];
        ai += vr[y] * twi[e] + vi[y] * twr[e];
      }
      s2r[o] = ar;
      s2i[o] = ai;
    }
    __syncthreads();
    for (int o = tid; o < CB * n1 * n1; o += nt) {  // z; children fastest so rows are written contiguously
      const int c = o % CB, q = o / CB, ky = q % n1, kz = q / n1;
      const Real* vr = &s2r[c * mn + ky];
      const Real* vi = &s2i[c * mn + ky];
      Real ar = 0, ai = 0;
      for (int z = 0, e = 0; z < m; z++, e = (e + kz >= n1) ? e + kz - n1 : e + kz) {
        ar += vr[z * n1] * twr[e] - vi[z * n1] * twi[e];
        ai += vr[z * n1] * twi[e] + vi[z * n1] * twr[e];
      }
      const int64_t f = kx + nh * (ky + n1 * kz);
      Real* d = Xs + (f * vg * 8 * e0 + c) * 2;
      d[0] = ar;
      d[1] = ai;
    }
    __syncthreads();
  }
}

// Inverse DFT of CB children of one target parent per block (grid: n_slots * e1 * (8/CB)
// blocks of 512 threads; Y rows component-major, Y[slot][f][8j+c] complex), evaluated only at
// the check-surface lattice points; per x-frequency
// slice the n1 x n1 spectrum plane of each child is read once (contiguous child rows), then z,
// y (n1 taps, m outputs) and the c2r x-sum is accumulated in registers.
template <typename Real>
__global__ __launch_bounds__(512) void k_ifft_down(const int32_t* __restrict__ first_child, const Real* __restrict__ Y,
                                                   Real* __restrict__ dn, int K1, int e1, int m, int CB,
                                                   const int32_t* __restrict__ lattice, int ns, const Real* __restrict__ scal,
                                                   int vg, int vi) {
  extern __shared__ unsigned char smem_raw[];
  Real* sm = reinterpret_cast<Real*>(smem_raw);
  const int n1 = 2 * m, nh = m + 1, M_dim = n1 * n1 * nh;
  const int nn = n1 * n1, nm = n1 * m, m2 = m * m;
  Real* twr = sm;                  // n1
  Real* twi = twr + n1;
  Real* yr = twi + n1;             // CB * n1*n1  [c][kz][ky]
  Real* yi = yr + CB * nn;
  Real* s1r = yi + CB * nn;        // CB * n1*m   [c][ky][z]
  Real* s1i = s1r + CB * nm;
  Real* s2r = s1i + CB * nm;       // CB * m*m    [c][y][z]
  Real* s2i = s2r + CB * m2;
  constexpr int MAXQ = 24;
  const int nblk = 8 / CB;
  const int slot = blockIdx.x / (e1 * nblk), rem = blockIdx.x % (e1 * nblk), j = rem / nblk, c0 = (rem % nblk) * CB;
  const int fc = first_child[slot] + c0;
  const int tid = threadIdx.x, nt = blockDim.x;
  init_twiddles(twr, twi, n1, 1.0);
  Real acc[MAXQ];
#pragma unroll
  for (int q = 0; q < MAXQ; q++) acc[q] = 0;
  const Real* Ys = Y + (((int64_t)slot * M_dim * vg + vi) * 8 * e1 + 8 * j + c0) * 2;
  __syncthreads();
  for (int kx = 0; kx < nh; kx++) {
    for (int o = tid; o < CB * nn; o += nt) {  // children fastest: contiguous rows of Y
      const int c = o % CB, q = o / CB;         // q = ky + n1*kz
      const int64_t f = kx + (int64_t)nh * q;
      const Real* src = Ys + (f * vg * 8 * e1 + c) * 2;
      yr[c * nn + q] = src[0];
      yi[c * nn + q] = src[1];
    }
    __syncthreads();
    for (int o = tid; o < CB * nm; o += nt) {  // z
      const int c = o / nm, r = o % nm, ky = r / m, z = r % m;
      const Real* vr = &yr[c * nn + ky];
      const Real* vi = &yi[c * nn + ky];
      Real ar = 0, ai = 0;
      for (int kz = 0, e = 0; kz < n1; kz++, e = (e + z >= n1) ? e + z - n1 : e + z) {
        ar += vr[kz * n1] * twr[e] - vi[kz * n1] * twi[e];
        ai += vr[kz * n1] * twi[e] + vi[kz * n1] * twr[e];
      }
      s1r[o] = ar;
      s1i[o] = ai;
    }
    __syncthreads();
    for (int o = tid; o < CB * m2; o += nt) {  // y
      const int c = o / m2, r = o % m2, y = r / m, z = r % m;
      const Real* vr = &s1r[c * nm + z];
      const Real* vi = &s1i[c * nm + z];
      Real ar = 0, ai = 0;
      for (int ky = 0, e = 0; ky < n1; ky++, e = (e + y >= n1) ? e + y - n1 : e + y) {
        ar += vr[ky * m] * twr[e] - vi[ky * m] * twi[e];
        ai += vr[ky * m] * twi[e] + vi[ky * m] * twr[e];
      }
      s2r[o] = ar;
      s2i[o] = ai;
    }
    __syncthreads();
    const Real wk = (kx == 0 || 2 * kx == n1) ? Real(1) : Real(2);
#pragma unroll
    for (int q = 0; q < MAXQ; q++) {
      const int o = tid + q * nt;
      if (o < CB * ns) {
        const int c = o / ns, p = o % ns;
        const int x = lattice[3 * p], y = lattice[3 * p + 1], z = lattice[3 * p + 2];
        const int kxx = kx * x, e = kxx - n1 * (kxx / n1);
        const int si = c * m2 + y * m + z;
        acc[q] += wk * (s2r[si] * twr[e] - s2i[si] * twi[e]);
      }
    }
    __syncthreads();
  }
  const Real sc = scal[slot];
#pragma unroll
  for (int q = 0; q < MAXQ; q++) {
    const int o = tid + q * nt;
    if (o < CB * ns) {
      const int c = o / ns, p = o % ns;
      dn[(int64_t)(fc + c) * K1 + p * e1 + j] += acc[q] * sc;
    }
  }
}

// ----------------------------------------------------------------------------- V list: DFTs, fixed m
// Twiddles w^k = exp(2 pi i k / N) as compile-time constants (Taylor series on the angle
// reduced to [0, pi/4], 14 terms: below double rounding), so fully unrolled tap loops carry
// them as literals.
constexpr double ct_sin_small(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; k++) { term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0)); sum += term; }
  return sum;
}
constexpr double ct_cos_small(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; k++) { term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k)); sum += term; }
  return sum;
}
// cos or sin of 2 pi k / n: quadrant q, remainder angle (pi/2) r/n folded into [0, pi/4]
constexpr double ct_cs(int k, int n, bool want_sin) {
  k %= n;
  if (k < 0) k += n;
  const int q = (4 * k) / n, r = 4 * k - q * n;
  const double half_pi = 1.57079632679489661923;
  double c = 0, s = 0;
  if (2 * r <= n) {
    c = ct_cos_small(half_pi * r / n);
    s = ct_sin_small(half_pi * r / n);
  } else {
    c = ct_sin_small(half_pi * (n - r) / n);
    s = ct_cos_small(half_pi * (n - r) / n);
  }
  if (q == 1) return want_sin ? c : -s;
  if (q == 2) return want_sin ? -s : -c;
  if (q == 3) return want_sin ? -c : s;
  return want_sin ? s : c;
}
template <int N>
struct TwTable {
  double c[N], s[N];
  constexpr TwTable() : c(), s() {
    for (int k = 0; k < N; k++) { c[k] = ct_cs(k, N, false); s[k] = ct_cs(k, N, true); }
  }
};

// Compile-time loop: f(std::integral_constant<int, i>) for i < N.
template <int I, int N, typename F>
__device__ __forceinline__ void sfor_(F& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor_<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F f) { sfor_<0, N>(f); }

// s += g * w^{SIGN K}, w = exp(2 pi i / N), K compile-time: twiddles +-1 and +-i become adds,
// real or imaginary twiddles one FMA per part.
template <int N, int K, int SIGN, typename Real>
__device__ __forceinline__ void cmac_tw(Real gr, Real gi, Real& sr, Real& si) {
  constexpr double c = ct_cs(K, N, false), s = SIGN * ct_cs(K, N, true);
  if constexpr (s == 0.0) {
    if constexpr (c == 1.0) { sr += gr; si += gi; }
    else if constexpr (c == -1.0) { sr -= gr; si -= gi; }
    else { sr = fma(gr, (Real)c, sr); si = fma(gi, (Real)c, si); }
  } else if constexpr (c == 0.0) {
    if constexpr (s == 1.0) { sr -= gi; si += gr; }
    else if constexpr (s == -1.0) { sr += gi; si -= gr; }
    else { sr = fma(-gi, (Real)s, sr); si = fma(gr, (Real)s, si); }
  } else {
    sr = fma(gr, (Real)c, sr);
    sr = fma(-gi, (Real)s, sr);
    si = fma(gr, (Real)s, si);
    si = fma(gi, (Real)c, si);
  }
}

// The spectrum bound of k_fft_up_m for the generic-m transform: bound[slot * e0 + i] =
// |scal[slot]| * sum of |up| over the 8 children's ns values of component i, in a fixed order.
template <typename Real>
__global__ __launch_bounds__(256) void k_spectrum_bound(int n_slots, const int32_t* __restrict__ first_child,
                                                        const Real* __restrict__ up, int K0, int e0, int ns,
                                                        const Real* __restrict__ scal, float* __restrict__ bound) {
  __shared__ Real red[256];
  const int slot = blockIdx.x, t = threadIdx.x;
  if (slot >= n_slots) return;  // whole block
  const int fc = first_child[slot];
  for (int i = 0; i < e0; i++) {
    Real a = 0;
    for (int q = t; q < 8 * ns; q += 256) a += fabs(up[(int64_t)(fc + q / ns) * K0 + (q % ns) * e0 + i]);
    red[t] = a;
    __syncthreads();
    if (t < 64) {
      Real b = red[t] + red[t + 64] + red[t + 128] + red[t + 192];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o);
      if (t == 0) bound[(int64_t)slot * e0 + i] = (float)(b * fabs(scal[slot]));
    }
    __syncthreads();
  }
}

// Scale of each run of tpb consecutive V-list tiles for the fp16 Hadamard: the power of two s
// with s * b < 2^14 <= 2 s * b, b the largest spectrum bound of the sources the run stages (its
// first tile's partners and every later tile's new ones, as (slot << 9 | pool slot)); 1 when b
// is 0 (or not finite).
__global__ __launch_bounds__(256) void k_run_scale(int n_tiles, int tpb, const int32_t* __restrict__ tile_u0,
                                                   const int32_t* __restrict__ usrc, const int32_t* __restrict__ tile_d0,
                                                   const int32_t* __restrict__ dent, const float* __restrict__ bound,
                                                   float* __restrict__ run_scale) {
  __shared__ float red[256];
  const int r = blockIdx.x, t = threadIdx.x, k0 = r * tpb, k1 = min(n_tiles, k0 + tpb);
  float mx = 0.f;
  for (int u = tile_u0[k0] + t; u < tile_u0[k0 + 1]; u += 256) mx = fmaxf(mx, bound[usrc[u] >> 9]);
  if (k0 + 1 < k1)
    for (int d = tile_d0[k0 + 1] + t; d < tile_d0[k1]; d += 256) mx = fmaxf(mx, bound[dent[d] >> 9]);
  red[t] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] = fmaxf(red[t], red[t + o]);
    __syncthreads();
  }
  if (t == 0) {
    const float b = red[0];
    float sc = 1.f;
    if (b > 0.f && b <= 3.4e38f) {
      int e = 0;
      (void)frexpf(b, &e);  // b < 2^e
      sc = ldexpf(1.f, min(100, max(-100, 14 - e)));
    }
    run_scale[r] = sc;
  }
}

// Shape of the fixed-m V-list transforms.  Lattice lines (c, y, z) are owned by threads:
// NT = m^2 * CG threads, thread t -> (y, z) = t % m^2 and children cg + CG*q (q < LPT).
// KC x-frequency planes of the intermediate are staged in LDS at a time, sized so that two
// blocks share one CU's 160 KiB.
template <typename Real, int M>
struct DftShape {
  static constexpr int N1 = 2 * M, NH = M + 1, MD = N1 * N1 * NH, M2 = M * M;
  static constexpr int CG = (8 * M2 <= 512) ? 8 : (4 * M2 <= 512) ? 4 : (2 * M2 <= 512) ? 2 : 1;
  static constexpr int NT = M2 * CG;
  static constexpr int LPT = 8 / CG;
  static constexpr int kc_for(int per_plane) {
    const int budget = 78 * 1024;
    const int kmax = budget / per_plane < 1 ? 1 : (budget / per_plane > NH ? NH : budget / per_plane);
    const int nch = (NH + kmax - 1) / kmax;
    return (NH + nch - 1) / nch;
  }
  static constexpr int KC_UP = kc_for(8 * M2 * 2 * (int)sizeof(Real));       // s1[c][kx][z][y]
  static constexpr int KC_DN = kc_for(8 * N1 * M * 2 * (int)sizeof(Real));   // s1[c][kx][ky][z]
  // FP64 recomputes the y stage per half (ky, ky+m) instead of holding both in registers
  static constexpr bool SPLIT_H = sizeof(Real) == 8;
};

// Forward DFT of the 8 children of one source parent, component i (grid n_slots * e0).
// x stage (real -> complex, kx < m+1) on thread-owned lattice lines into LDS; then one task
// per (c, kx, ky < m) runs the y stage for ky and ky+m (w^{-(ky+m)y} = (-1)^y w^{-ky y}) and the
// z stage of both (X[kz] = E + O, X[kz+m] = E - O over even/odd z), writing the spectra as
// X[slot][f][8i+c] complex (component-major, so each component's block writes its own 64-byte
// segment of a row), f = kx + (m+1)(ky + 2m kz).
// NTS: non-temporal stores of X, which is read back only by the Hadamard, long after its lines
// would have left the caches (1.20 -> 1.04 ms for the 2M adaptive workload).  Only where each
// block writes whole lines: with several components (or vectors in one group) the blocks of one
// parent interleave their stores within a line, which must merge in L2 first.
template <typename Real, int M, bool NTS>
__global__ __launch_bounds__((DftShape<Real, M>::NT), 4) void k_fft_up_m(const int32_t* __restrict__ first_child,
                                                                         const Real* __restrict__ up, int K0, int e0,
                                                                         const int32_t* __restrict__ lat_inv,
                                                                         const Real* __restrict__ scal, Real* __restrict__ X,
                                                                         int n_slots, int vg, int vi, float* __restrict__ bound) {
  using S = DftShape<Real, M>;
  constexpr int N1 = S::N1, NH = S::NH, MD = S::MD, M2 = S::M2, NT = S::NT, LPT = S::LPT, CG = S::CG, KC = S::KC_UP;
  constexpr int CST = KC * M2 + 1;  // child stride (odd, spreads the children over LDS banks)
  constexpr TwTable<N1> TW{};
  __shared__ Real twc[N1], tws[N1];
  __shared__ Real s1r[8 * CST], s1i[8 * CST];
  // the e0 components of a source parent run back to back on one XCD (blocks b, b + 8, ...:
  // block b runs on XCD b mod 8), so their interleaved stores to X merge in that XCD's L2
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int i = local % e0, slot = (local / e0) * 8 + xcd;
  if (slot >= n_slots) return;  // whole block
  const int fc = first_child[slot];
  const Real sc = scal[slot];
  const int t = threadIdx.x, yz = t % M2, cg = t / M2, y = yz % M, z = yz / M;
  if (t < N1) { twc[t] = (Real)TW.c[t]; tws[t] = (Real)TW.s[t]; }
  Real v[LPT][M];  // surface values on the owned lines (zero inside the cube)
#pragma unroll
  for (int x = 0; x < M; x++) {
    const int p = lat_inv[(z * M + y) * M + x];
#pragma unroll
    for (int q = 0; q < LPT; q++) v[q][x] = p >= 0 ? up[(int64_t)(fc + cg + CG * q) * K0 + p * e0 + i] * sc : Real(0);
  }
  if (bound) {
    // sum of |v| over the slot's lattice values of component i: no frequency of that
    // component's spectrum exceeds it (unit twiddles); added in a fixed order (per thread, then
    // the first wave, NT >= 64)
    Real a = 0;
#pragma unroll
    for (int q = 0; q < LPT; q++)
#pragma unroll
      for (int x = 0; x < M; x++) a += fabs(v[q][x]);
    s1r[t] = a;  // s1r is free until the x stage
    __syncthreads();
    if (t < 64) {
      Real b = 0;
      for (int u = t; u < NT; u += 64) b += s1r[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o);
      if (t == 0) bound[(int64_t)slot * e0 + i] = (float)b;
    }
  }
  __syncthreads();
  Real* Xs = X + (((int64_t)slot * MD * vg + vi) * 8 * e0 + 8 * i) * 2;
  for (int k0 = 0; k0 < NH; k0 += KC) {
    const int kn = (NH - k0 < KC) ? NH - k0 : KC;
#pragma unroll
    for (int q = 0; q < LPT; q++) {  // x stage
      const int c = cg + CG * q;
      for (int kl = 0; kl < kn; kl++) {
        const int kx = k0 + kl;
        Real ar = 0, ai = 0;
#pragma unroll
        for (int x = 0, e = 0; x < M; x++) {
          ar = fma(v[q][x], twc[e], ar);
          ai = fma(v[q][x], -tws[e], ai);
          e += kx;
          if (e >= N1) e -= N1;
        }
        s1r[c * CST + kl * M2 + yz] = ar;
        s1i[c * CST + kl * M2 + yz] = ai;
      }
    }
    __syncthreads();
    const int ntask = 8 * M * kn;
    for (int T = t; T < ntask; T += NT) {  // y and z stages
      // child fastest, then kx: a wave's stores of one (ky, kz) cover consecutive frequencies
      const int c = T % 8, kl = (T / 8) % kn, ky = T / (8 * kn), kx = k0 + kl;
      const Real* pr = &s1r[c * CST + kl * M2];
      const Real* pi = &s1i[c * CST + kl * M2];
      // z stage of one y-stage output column g (z < m) for frequency row kyy
      auto zstage = [&](const Real* gr, const Real* gi, int kyy) __attribute__((always_inline)) {
        sfor<M>([&](auto ikz) __attribute__((always_inline)) {
          constexpr int kz = decltype(ikz)::value;
          Real er = 0, ei = 0, orr = 0, oi = 0;
          sfor<M>([&](auto izz) __attribute__((always_inline)) {
            constexpr int zz = decltype(izz)::value;
            if constexpr (zz % 2 == 0)
              cmac_tw<N1, (kz * zz) % N1, -1>(gr[zz], gi[zz], er, ei);
            else
              cmac_tw<N1, (kz * zz) % N1, -1>(gr[zz], gi[zz], orr, oi);
          });
          const int64_t f0 = kx + (int64_t)NH * (kyy + N1 * kz), f1 = kx + (int64_t)NH * (kyy + N1 * (kz + M));
          // an opaque copy of the base keeps the 64-bit store addresses from being hoisted
          // out of the unrolled loop (they would not fit in registers for FP64)
          Real* xb = Xs;
          asm volatile("" : "+s"(xb));
          Real* d0 = xb + (f0 * vg * 8 * e0 + c) * 2;
          Real* d1 = xb + (f1 * vg * 8 * e0 + c) * 2;
          if constexpr (NTS) {
            __builtin_nontemporal_store(er + orr, &d0[0]);
            __builtin_nontemporal_store(ei + oi, &d0[1]);
            __builtin_nontemporal_store(er - orr, &d1[0]);
            __builtin_nontemporal_store(ei - oi, &d1[1]);
          } else {
            d0[0] = er + orr; d0[1] = ei + oi;
            d1[0] = er - orr; d1[1] = ei - oi;
          }
        });
      };
      // y stage: column zz of ky (E + O) and of ky + m (E - O)
      auto ystage = [&](int zz, Real& er, Real& ei, Real& orr, Real& oi) __attribute__((always_inline)) {
        er = ei = orr = oi = 0;
#pragma unroll
        for (int yy = 0, e = 0; yy < M; yy++) {
          const Real dr = pr[zz * M + yy], di = pi[zz * M + yy];
          const Real wc = twc[e], ws = -tws[e];
          const Real tr = dr * wc - di * ws, ti = dr * ws + di * wc;
          if (yy % 2 == 0) { er += tr; ei += ti; } else { orr += tr; oi += ti; }
          e += ky;
          if (e >= N1) e -= N1;
        }
      };
      if constexpr (S::SPLIT_H) {
        Real gr[M], gi[M];
#pragma unroll
        for (int zz = 0; zz < M; zz++) {
          Real er, ei, orr, oi;
          ystage(zz, er, ei, orr, oi);
          gr[zz] = er + orr; gi[zz] = ei + oi;
        }
        zstage(gr, gi, ky);
#pragma unroll
        for (int zz = 0; zz < M; zz++) {
          Real er, ei, orr, oi;
          ystage(zz, er, ei, orr, oi);
          gr[zz] = er - orr; gi[zz] = ei - oi;
        }
        zstage(gr, gi, ky + M);
      } else {
        Real ar[M], ai[M], br[M], bi[M];
#pragma unroll
        for (int zz = 0; zz < M; zz++) {
          Real er, ei, orr, oi;
          ystage(zz, er, ei, orr, oi);
          ar[zz] = er + orr; ai[zz] = ei + oi;
          br[zz] = er - orr; bi[zz] = ei - oi;
        }
        zstage(ar, ai, ky);
        zstage(br, bi, ky + M);
      }
    }
    __syncthreads();
  }
}

// Inverse DFT of the 8 children of one target parent, component j (grid n_slots * e1),
// evaluated at the check-surface lattice points only; Y rows are component-major
// (Y[slot][f][8j+c] complex), so a component's 8 children are one 64-byte segment.  z stage
// per (c, kx, ky) task straight from Y (output z < m needs P = Y[a] + Y[a+m] for even z, Q = Y[a] - Y[a+m] for odd z) into
// LDS; then each thread runs the y stage
// of its own lines (c, y, z) and accumulates the c2r x sum in registers; the lattice values
// are added into dn at the end.
// FP64: one block per CU, so that the compiler keeps the y/z stage in registers instead of
// spilling them to fit four (measured at m = 10: 3.78 -> 2.82 ms; the FP32 form is fastest at 4)
template <typename Real, int M>
__global__ __launch_bounds__((DftShape<Real, M>::NT), (sizeof(Real) == 8 ? 1 : 4)) void k_ifft_down_m(const int32_t* __restrict__ first_child,
                                                                            const Real* __restrict__ Y, Real* __restrict__ dn,
                                                                            int K1, int e1, const int32_t* __restrict__ lat_inv,
                                                                            const Real* __restrict__ scal, int n_slots, int vg,
                                                                            int vi) {
  using S = DftShape<Real, M>;
  constexpr int N1 = S::N1, NH = S::NH, MD = S::MD, M2 = S::M2, NT = S::NT, LPT = S::LPT, CG = S::CG, KC = S::KC_DN;
  constexpr int PL = N1 * M;         // one plane of one child: [ky][z]
  constexpr int CST = KC * PL + 1;   // child stride (odd)
  constexpr TwTable<N1> TW{};
  __shared__ Real twc[N1], tws[N1];
  __shared__ Real s1r[8 * CST], s1i[8 * CST];
  // the e1 components of a target parent run back to back on one XCD, so their interleaved
  // reads of Y share that XCD's L2 lines
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int j = local % e1, slot = (local / e1) * 8 + xcd;
  if (slot >= n_slots) return;  // whole block
  const int fc = first_child[slot];
  const int t = threadIdx.x, yz = t % M2, cg = t / M2, y = yz % M, z = yz / M;
  if (t < N1) { twc[t] = (Real)TW.c[t]; tws[t] = (Real)TW.s[t]; }
  __syncthreads();
  const Real sgn = (y & 1) ? Real(-1) : Real(1);
  Real acc[LPT][M];
#pragma unroll
  for (int q = 0; q < LPT; q++)
#pragma unroll
    for (int x = 0; x < M; x++) acc[q][x] = 0;
  const Real* Ys = Y + (((int64_t)slot * MD * vg + vi) * 8 * e1 + 8 * j) * 2;
  for (int k0 = 0; k0 < NH; k0 += KC) {
    const int kn = (NH - k0 < KC) ? NH - k0 : KC;
    const int ntask = 8 * N1 * kn;
    for (int T = t; T < ntask; T += NT) {  // z stage
      const int c = T % 8, ky = (T / 8) % N1, kl = T / (8 * N1), kx = k0 + kl;
      Real pr_[M], pi_[M], qr_[M], qi_[M];
#pragma unroll
      for (int a = 0; a < M; a++) {
        const Real* s0 = Ys + ((kx + (int64_t)NH * (ky + N1 * a)) * vg * 8 * e1 + c) * 2;
        const Real* s1 = Ys + ((kx + (int64_t)NH * (ky + N1 * (a + M))) * vg * 8 * e1 + c) * 2;
        const Real ur = s0[0], ui = s0[1], vr = s1[0], vi = s1[1];
        pr_[a] = ur + vr; pi_[a] = ui + vi;
        qr_[a] = ur - vr; qi_[a] = ui - vi;
      }
      Real* dr = &s1r[c * CST + kl * PL + ky * M];
      Real* di = &s1i[c * CST + kl * PL + ky * M];
      sfor<M>([&](auto izz) __attribute__((always_inline)) {
        constexpr int zz = decltype(izz)::value;
        Real sr = 0, si = 0;
        sfor<M>([&](auto ia) __attribute__((always_inline)) {
          constexpr int a = decltype(ia)::value;
          if constexpr (zz % 2 == 0)
            cmac_tw<N1, (a * zz) % N1, 1>(pr_[a], pi_[a], sr, si);
          else
            cmac_tw<N1, (a * zz) % N1, 1>(qr_[a], qi_[a], sr, si);
        });
        dr[zz] = sr;
        di[zz] = si;
      });
    }
    __syncthreads();
    for (int kl = 0; kl < kn; kl++) {  // y stage and c2r x sum of the owned lines
      const int kx = k0 + kl;
      const Real wk = (kx == 0 || 2 * kx == N1) ? Real(1) : Real(2);
      Real sr[LPT], si[LPT];
#pragma unroll
      for (int q = 0; q < LPT; q++) {
        const int c = cg + CG * q;
        const Real* pr = &s1r[c * CST + kl * PL + z];
        const Real* pi = &s1i[c * CST + kl * PL + z];
        Real ar = 0, ai = 0;
#pragma unroll
        for (int a = 0, e = 0; a < M; a++) {
          const Real gr = pr[a * M] + sgn * pr[(a + M) * M];
          const Real gi = pi[a * M] + sgn * pi[(a + M) * M];
          const Real wc = twc[e], ws = tws[e];
          ar += gr * wc - gi * ws;
          ai += gr * ws + gi * wc;
          e += y;
          if (e >= N1) e -= N1;
        }
        sr[q] = wk * ar;
        si[q] = wk * ai;
      }
#pragma unroll
      for (int x = 0, e = 0; x < M; x++) {
        const Real wc = twc[e], ws = tws[e];
#pragma unroll
        for (int q = 0; q < LPT; q++) acc[q][x] += sr[q] * wc - si[q] * ws;
        e += kx;
        if (e >= N1) e -= N1;
      }
    }
    __syncthreads();
  }
  const Real sc = scal[slot];
#pragma unroll
  for (int q = 0; q < LPT; q++) {
    const int c = cg + CG * q;
#pragma unroll
    for (int x = 0; x < M; x++) {
      const int p = lat_inv[(z * M + y) * M + x];
      if (p >= 0) dn[(int64_t)(fc + c) * K1 + p * e1 + j] += acc[q][x] * sc;
    }
  }
}

// ----------------------------------------------------------------------------- V list: Hadamard
// Frequency-domain M2L as a real GEMM per frequency f:
//   Y[t-slots][f][16*E1 reals] = sum over the 26 parent offsets R of
//   X[partner_R(t)][f][16*E0 reals] x B_R(f)[16*E0 x 16*E1],
// the complex 8x8 (x E0,E1) transfer block in real form.  Vp holds B pre-permuted per lane:
// lane l's four values for k-step q are B[16kb + 4(l/16) + q][16nb + l%16], so both MFMA
// operands load as 4-vectors.
//
// Tiled frequency-domain M2L.  A block owns one tile of consecutive target parents (at most
// 64 = 4 groups of 16, with at most UMAX distinct V1 partners) and FB consecutive
// frequencies.  The partners' spectra for those frequencies are staged in LDS once (entry
// stride padded by one 4-vector against bank conflicts, plus one zero entry that stands for
// a missing partner).  The WPF = 4/FB waves of one frequency split the 26 parent offsets;
// each runs its offsets over all 4 target groups with the transfer block of the offset as the
// shared MFMA B operand, held in registers, and the outputs accumulated in registers.  The
// partial sums of the waves of one frequency are added through LDS in a fixed order.  The
// instruction stream has no data-dependent branches, so the accumulators stay in place.
// Blocks are mapped so that all tiles of one frequency group run on one XCD and share its L2.
// Blocks of the tiled kernel that fit a CU's 160 KiB of LDS (at most 4); the register
// budget is set to match, so the compiler keeps to it.
template <typename Real, int E0, int FB, int UMAX>
constexpr int hadamard_blocks_per_cu() {
  constexpr int lds = (UMAX + 1) * (FB * 16 * E0 + 4) * (int)sizeof(Real) + 64 * 26 * 2;
  return 163840 / lds > 4 ? 4 : 163840 / lds;
}

template <typename Real, int E0, int E1, int FB, int UMAX, int NACC>
__global__ __launch_bounds__(256, (hadamard_blocks_per_cu<Real, E0, FB, UMAX>())) void k_hadamard_tiled(int M_dim, int n_tiles, const int32_t* __restrict__ tile_t0,
                                                        const int32_t* __restrict__ tile_u0, const int32_t* __restrict__ usrc,
                                                        const uint16_t* __restrict__ uidx, const uint32_t* __restrict__ gmask,
                                                        const int32_t* __restrict__ offsets,
                                                        const Real* __restrict__ Vp, const Real* __restrict__ X,
                                                        Real* __restrict__ Y) {
  using MF = Mfma<Real>;
  using V4 = typename MF::V4;
  constexpr int NOFF = 26;                        // V1 parent offsets
  constexpr int RW = 16 * E0;                     // reals of one (slot, frequency) spectrum row
  constexpr int PAD = 4;                          // one 4-vector of pad per staged entry
  constexpr int EST = FB * RW + PAD;              // staged entry stride (reals)
  constexpr int WPF = 4 / FB;                     // waves per frequency (split the offsets)
  constexpr int NOW = NOFF / WPF;                 // offsets per wave
  static_assert(NOW * WPF == NOFF, "offsets split evenly");
  constexpr int NB = E0 * E1;                     // 4-vectors of one offset's transfer block per lane
  // transfer blocks held in registers, PD offsets ahead (all of the wave's offsets when they
  // fit in ~104 VGPRs); the first PD are loaded before the staging barrier
  constexpr int PD0 = 104 / (NB * (int)sizeof(V4) / 4);
  constexpr int PD = PD0 < 1 ? 1 : (PD0 > NOW ? NOW : PD0);
  __shared__ __align__(16) Real Xl[(UMAX + 1) * EST];  // entry UMAX stays zero
  __shared__ __align__(8) uint16_t ul[64 * NOFF];
  const int b = blockIdx.x, xcd = b & 7, local = b >> 3;
  const int fg = (local / n_tiles) * 8 + xcd, tile = local % n_tiles;
  const int f0 = fg * FB;
  if (f0 >= M_dim) return;  // whole block
  const int t0 = tile_t0[tile], nt = tile_t0[tile + 1] - t0;
  const int u0 = tile_u0[tile], nu = tile_u0[tile + 1] - u0;
  const int fn = min(FB, M_dim - f0);
  constexpr int V4PU = FB * RW / 4;  // 4-vectors per entry
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int fl = w / WPF, gsub = w % WPF;
  const bool live = fl < fn;                  // waves past the last frequency compute, not store
  const int f = f0 + (live ? fl : 0);
  const int ob = gsub * NOW;                  // this wave's first offset
  const Real* vf = Vp + (int64_t)f * (NB * 256) + l * 4;
  const int64_t vstride = (int64_t)M_dim * NB * 256;
  // The prologue keeps its global loads in flight together (one latency each for the index
  // and data reads instead of one per loop trip): the partner table, the staged spectra in
  // batches of SB passes, and the transfer blocks (first when the registers allow).
  constexpr bool V_FIRST = sizeof(Real) == 4 && E0 == 1;
  V4 bq[PD][NB];
  auto load_v = [&]() {
#pragma unroll
    for (int p = 0; p < PD; p++)
#pragma unroll
      for (int q = 0; q < NB; q++) bq[p][q] = *reinterpret_cast<const V4*>(vf + offsets[ob + p] * vstride + q * 256);
  };
  if constexpr (V_FIRST) load_v();
  constexpr int UW = 64 * NOFF / 4;  // partner table as 8-byte words
  constexpr int UWP = (UW + 255) / 256;
  uint2 uw[UWP];
  const uint2* ug = reinterpret_cast<const uint2*>(uidx + (int64_t)tile * 64 * NOFF);
#pragma unroll
  for (int i = 0; i < UWP; i++) {
    const int e = threadIdx.x + 256 * i;
    if (e < UW) uw[i] = ug[e];
  }
  constexpr int NPASS = (UMAX * V4PU + 255) / 256;
  constexpr int SB = sizeof(Real) == 4 ? 8 : 4;
#pragma unroll
  for (int p0 = 0; p0 < NPASS; p0 += SB) {
    if (256 * p0 >= nu * V4PU) break;  // block-uniform
    int src[SB];
#pragma unroll
    for (int i = 0; i < SB; i++) {
      const int u = (threadIdx.x + 256 * (p0 + i)) / V4PU;
      src[i] = (p0 + i < NPASS && u < nu) ? usrc[u0 + u] : -1;
    }
    V4 val[SB];
#pragma unroll
    for (int i = 0; i < SB; i++) {
      const int e = threadIdx.x + 256 * (p0 + i), u = e / V4PU, r = e - u * V4PU;
      val[i] = V4{0, 0, 0, 0};
      if (src[i] >= 0 && r / (RW / 4) < fn)
        val[i] = *reinterpret_cast<const V4*>(X + ((int64_t)src[i] * M_dim + f0) * RW + 4 * r);
    }
#pragma unroll
    for (int i = 0; i < SB; i++) {
      const int e = threadIdx.x + 256 * (p0 + i), u = e / V4PU, r = e - u * V4PU;
      if (p0 + i < NPASS && u < nu) *reinterpret_cast<V4*>(&Xl[u * EST + 4 * r]) = val[i];
    }
  }
  // partner table [offset][row][group]: one 8-byte read per lane and offset gives its 4 groups
#pragma unroll
  for (int i = 0; i < UWP; i++) {
    const int e = threadIdx.x + 256 * i;
    if (e < UW) reinterpret_cast<uint2*>(ul)[e] = uw[i];
  }
  for (int e = threadIdx.x; e < EST; e += 256) Xl[UMAX * EST + e] = Real(0);
  if constexpr (!V_FIRST) load_v();
  // pin the blocks in registers here: otherwise the loads are sunk to their first use past
  // the barrier and the wave stalls on each of them in the offset loop
#pragma unroll
  for (int p = 0; p < PD; p++)
#pragma unroll
    for (int q = 0; q < NB; q++) asm volatile("" : "+v"(bq[p][q]));
  __syncthreads();
  const int row = l & 15, kg = l >> 4;
  // offsets present per 16-target group of the tile: for the FP32 3x3 (Stokes) blocks a group's
  // 36 MFMAs at an offset where none of its targets has a partner are skipped (block-uniform
  // branches).  The scalar kernels keep the dense stream: there a branch guards only 4 MFMAs
  // and makes the compiler copy the accumulators through VGPRs around it; the FP64 3x3 form
  // already uses 252 of its 256 VGPRs and would spill.
  uint32_t gm[4];
#pragma unroll
  for (int g = 0; g < 4; g++) gm[g] = gmask[4 * tile + g];
  // NACC accumulator sets (offset j uses set j % NACC): independent MFMA chains
  V4 accs[NACC][4][E1];
#pragma unroll
  for (int a = 0; a < NACC; a++)
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int nb = 0; nb < E1; nb++) accs[a][g][nb] = V4{0, 0, 0, 0};
  // A operands (partner spectra of the 4 groups) of the wave's next offset are read one
  // offset ahead, and their partner indices two ahead, so that the LDS latency overlaps the
  // MFMAs; scheduling barriers keep the reads ahead of the MFMAs, which the scheduler would
  // otherwise sink next to their use (scalar kernels; the Stokes blocks would need 3x the
  // registers)
  V4 an[4];
  using U4 = unsigned short __attribute__((ext_vector_type(4)));
  U4 un;
  const int lane_off = fl * RW + 4 * kg;
  if constexpr (E0 == 1) {
    const U4 uu = *reinterpret_cast<const U4*>(&ul[(ob * 16 + row) * 4]);
#pragma unroll
    for (int g = 0; g < 4; g++) an[g] = *reinterpret_cast<const V4*>(&Xl[(int)uu[g] * EST + lane_off]);
    if (NOW > 1) un = *reinterpret_cast<const U4*>(&ul[((ob + 1) * 16 + row) * 4]);
  }
#pragma unroll
  for (int j = 0; j < NOW; j++) {
    const int oi = ob + j;
    auto& acc = accs[j % NACC];
    V4 cur[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) cur[q] = bq[j % PD][q];
    if (j + PD < NOW) {
#pragma unroll
      for (int q = 0; q < NB; q++) bq[j % PD][q] = *reinterpret_cast<const V4*>(vf + offsets[oi + PD] * vstride + q * 256);
    }
    if constexpr (E0 == 1) {
      V4 ac[4];
#pragma unroll
      for (int g = 0; g < 4; g++) ac[g] = an[g];
      if (j + 1 < NOW) {
#pragma unroll
        for (int g = 0; g < 4; g++) an[g] = *reinterpret_cast<const V4*>(&Xl[(int)un[g] * EST + lane_off]);
      }
      if (j + 2 < NOW) un = *reinterpret_cast<const U4*>(&ul[((oi + 2) * 16 + row) * 4]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int g = 0; g < 4; g++)
#pragma unroll
        for (int nb = 0; nb < E1; nb++) {
          acc[g][nb] = MF::mma(ac[g][0], cur[nb][0], acc[g][nb]);
          acc[g][nb] = MF::mma(ac[g][1], cur[nb][1], acc[g][nb]);
          acc[g][nb] = MF::mma(ac[g][2], cur[nb][2], acc[g][nb]);
          acc[g][nb] = MF::mma(ac[g][3], cur[nb][3], acc[g][nb]);
        }
      __builtin_amdgcn_sched_barrier(0);
      continue;
    }
#pragma unroll
    for (int g = 0; g < 4; g++) {
      if constexpr (sizeof(Real) == 4)
        if (!((gm[g] >> oi) & 1u)) continue;
      const int ue = ul[(oi * 16 + row) * 4 + g];  // missing partners point at the zero entry
#pragma unroll
      for (int kb = 0; kb < E0; kb++) {
        const V4 a = *reinterpret_cast<const V4*>(&Xl[ue * EST + fl * RW + 16 * kb + 4 * kg]);
#pragma unroll
        for (int nb = 0; nb < E1; nb++) {
          const V4 bv = cur[kb * E1 + nb];
          acc[g][nb] = MF::mma(a[0], bv[0], acc[g][nb]);
          acc[g][nb] = MF::mma(a[1], bv[1], acc[g][nb]);
          acc[g][nb] = MF::mma(a[2], bv[2], acc[g][nb]);
          acc[g][nb] = MF::mma(a[3], bv[3], acc[g][nb]);
        }
      }
    }
  }
  V4 acc[4][E1];
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int nb = 0; nb < E1; nb++) {
      acc[g][nb] = accs[0][g][nb];
#pragma unroll
      for (int a = 1; a < NACC; a++) acc[g][nb] += accs[a][g][nb];
    }
  if constexpr (WPF > 1) {
    // partial sums of the other waves of this frequency, added in wave order
    static_assert(WPF == 2, "two waves per frequency");
    static_assert((int)sizeof(Xl) >= 2 * 4 * E1 * 256 * (int)sizeof(Real), "exchange fits in the staging area");
    __syncthreads();  // every wave is done reading Xl
    Real* ex = Xl + fl * (4 * E1 * 256);
    if (gsub == 1) {
#pragma unroll
      for (int g = 0; g < 4; g++)
#pragma unroll
        for (int nb = 0; nb < E1; nb++) *reinterpret_cast<V4*>(&ex[((g * E1 + nb) * 64 + l) * 4]) = acc[g][nb];
    }
    __syncthreads();
    if (gsub == 1) return;
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int nb = 0; nb < E1; nb++) acc[g][nb] += *reinterpret_cast<const V4*>(&ex[((g * E1 + nb) * 64 + l) * 4]);
  }
  if (!live) return;
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int nb = 0; nb < E1; nb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int tg = g * 16 + MF::row(l, r);
        if (tg < nt) Y[((int64_t)(t0 + tg) * M_dim + f) * (16 * E1) + 16 * nb + row] = acc[g][nb][r];
      }
}

// The three bf16 parts of an FP32 value x = x0 + x1 + x2 (x0 the top 8 significant bits, x1
// the next 8, x2 the rest; each step exact): the operand split of k_hadamard_vec16, whose six
// largest products x0 t0 + (x0 t1 + x1 t0) + (x0 t2 + x1 t1 + x2 t0) leave out less than
// 2^-21 |x t| (transfer parts rounded to nearest: |t1| <= 2^-8 |t|, |t2| <= 2^-16 |t|).
//
// FP32 frequency-domain M2L for the scalar kernels (k_hadamard_split) on the fp16 matrix
// cores: operands are scaled by powers of two into fp16 range and split exactly into two fp16
// parts x = x0 + x1 (x0 = RN(x), x1 = RN(x - x0): 22 significant bits, |x - x0 - x1| <=
// 2^-22 |x|), and the three largest products are summed in FP32,
//   x.t ~ x0 t0 + (x0 t1 + x1 t0),
// each exact in FP32 (11 x 11 bits); the dropped x1 t1 stays below 2^-22 |x t|.  The spectra
// of the sources a block stages are scaled by its run's power of two s (k_run_scale: s times
// the largest bound of their spectra is below 2^14), the transfer blocks of frequency f by
// tau_f (host: tau_f times the block's largest entry below 2^14), and the result is divided by
// s tau_f (exact).  Spectra far below their run's largest bound lose relative precision in x1
// (fp16 subnormals) but stay within about 2^-24 of that bound, the FP32 rounding of the run's
// largest values.
// One v_mfma_f32_16x16x32_f16 covers two parent offsets (k = 16 a + real index, a the offset
// of the pair), so an offset pair of a 16-target group costs 3 MFMAs of K = 32 instead of the 8
// FP32 MFMAs of K = 4 per offset.  A block owns one frequency and a run of consecutive tiles;
// wave w computes target group w (targets 16w..16w+15 of a tile) over the 13 offset pairs, with
// the transfer-block parts of its frequency held in registers for the run.  The partners'
// spectra are staged in LDS as their two fp16 parts (scaled and split while staging): while a
// tile's MFMAs run, the next tile's partner table and spectra are read from global memory and
// written to LDS, one barrier per tile.  Pairs at which none of the group's targets has a
// partner are skipped.  Two blocks run per CU.
// The staged entries form one LDS pool of POOL entries for the block's run of tiles (entry
// POOL stays zero, entry POOL + 1 takes the writes of idle lanes): a partner that consecutive
// tiles list stays in its slot, so after the run's first tile only the partners new since the
// previous tile are read and staged, into slots that the previous tile does not read (the host
// assigns the slots).  Staging passes past the length of that list are skipped.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kSplitPool = 496;

// the three bf16 parts of an FP32 value, by truncation (each step exact), as the high halves
// of three words
__device__ __forceinline__ void split3(float x, uint32_t& h0, uint32_t& h1, uint32_t& h2) {
  const uint32_t u0 = __float_as_uint(x);
  h0 = u0 & 0xFFFF0000u;
  const float r1 = x - __uint_as_float(h0);
  const uint32_t u1 = __float_as_uint(r1);
  h1 = u1 & 0xFFFF0000u;
  const float r2 = r1 - __uint_as_float(h1);
  h2 = __float_as_uint(r2) & 0xFFFF0000u;  // at most 8 significant bits: exact
}
// two bf16 (the high halves of a and b, low halves zero) packed into one word, a in the low half
__device__ __forceinline__ uint32_t pack_hi(uint32_t a, uint32_t b) { return (a >> 16) | b; }

template <int UMAX, int POOL>
__global__ __launch_bounds__(256, 2) void k_hadamard_split(int M_dim, int n_tiles, int tiles_per_block, int n_chunks,
                                                          const int32_t* __restrict__ tile_t0,
                                                          const int32_t* __restrict__ tile_u0, const int32_t* __restrict__ usrc,
                                                          const int32_t* __restrict__ tile_d0, const int32_t* __restrict__ dent,
                                                          const uint16_t* __restrict__ uidx, const uint32_t* __restrict__ gmask,
                                                          const uint16_t* __restrict__ Vh, const float* __restrict__ tinv,
                                                          const float* __restrict__ run_scale, const float* __restrict__ X,
                                                          float* __restrict__ Y) {
  constexpr int NOFF = 26, NPAIR = NOFF / 2;
  constexpr int RW = 16;                  // reals per staged entry and frequency
  constexpr int EST = 2 * RW / 2 + 4;     // words per staged entry: 2 parts x 16 fp16, padded
  constexpr int XBUF = (POOL + 2) * EST;  // entry POOL stays zero, POOL + 1 takes unused lanes
  constexpr int UT = 4 * 16 * 2 * 16;     // partner table of a tile: [group][pair / 8][a][row][pair % 8] (pairs padded to 16)
  constexpr int UW = UT / 4;              // as 8-byte words
  constexpr int UWP = UW / 256;
  constexpr int EPP = 256 / (RW / 4);     // staged entries per pass (a thread stages 4 reals)
  constexpr int NPASS = (UMAX + EPP - 1) / EPP;
  constexpr int TMAX = 64;                // longest run of tiles (the launcher clamps to it)
  static_assert(NPASS < NPAIR && UW % 256 == 0 && POOL + 1 < 512, "staging shape");
  __shared__ __align__(16) uint32_t Xl[XBUF];
  __shared__ __align__(16) uint16_t ul[2][UT];
  __shared__ uint32_t m_g[4 * TMAX];  // the run's skip masks, delta starts and target starts
  __shared__ int32_t m_d[TMAX + 1], m_t[TMAX + 1];
  // blocks b and b + 8 run on one XCD one after the other and take the two frequencies of one
  // 128-byte line of X and Y (a frequency's 16 reals are 64 bytes) for the same run of tiles,
  // so the second block's reads hit that XCD's L2
  const int b = blockIdx.x, xcd = b & 7, local = b >> 3, lp = local >> 1;
  const int f = ((lp / n_chunks) * 8 + xcd) * 2 + (local & 1), chunk = lp % n_chunks;
  const int k0 = chunk * tiles_per_block, k1 = min(n_tiles, k0 + tiles_per_block);
  if (f >= M_dim || k0 >= k1) return;  // whole block
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int row = l & 15, kg = l >> 4;
  const int pa = kg >> 1;                   // the lane's offset within a pair (k = 16 pa + ...)
  const int lane_off = 4 * (kg & 1);        // word offset of its 8 reals within a part
  const int sr = tid & 3, su = tid >> 2;    // staging: reals 4sr..4sr+3 of entries su + EPP * i

  // transfer-block parts of frequency f (scaled by tau_f): bq[j][p] holds, for lane
  // (n = l%16, kg), part p of B_{2j + kg/2}[8 (kg%2) + i][n], i < 8
  halfx8 bq[NPAIR][2];
#pragma unroll
  for (int j = 0; j < NPAIR; j++)
#pragma unroll
    for (int p = 0; p < 2; p++)
      bq[j][p] = *reinterpret_cast<const halfx8*>(Vh + ((((int64_t)j * 2 + p) * M_dim + f) * 64 + l) * 8);
  // the run's scale, and the factor that undoes both scales
  const float xsc = run_scale[chunk], ysc = tinv[f] / xsc;
  for (int e = tid; e < EST; e += 256) Xl[POOL * EST + e] = 0u;
  const int nk = k1 - k0;
  for (int i = tid; i < 4 * nk; i += 256) m_g[i] = gmask[4 * k0 + i];
  for (int i = tid; i <= nk; i += 256) {
    m_d[i] = tile_d0[k0 + i];
    m_t[i] = tile_t0[k0 + i];
  }

  int src[NPASS];  // staged entries (source slot << 9 | pool slot)
  int dst[NPASS];  // their pool slots, kept from the global loads to the LDS stores
  floatx4 val[NPASS];
  uint2 uw[UWP];
  // start and length of tile k's delta list (length 0 past the run's end), from the run's
  // copy in LDS, the length uniform over the block
  auto d0_l = [&](int k) __attribute__((always_inline)) { return m_d[min(k, k1 - 1) - k0]; };
  auto nd_l = [&](int k) __attribute__((always_inline)) {
    const int kk = min(k, k1 - 1) - k0;
    const int n = m_d[kk + 1] - m_d[kk];
    return __builtin_amdgcn_readfirstlane(k < k1 ? n : 0);
  };
  auto load_src = [&](int i, const int32_t* list, int at, int n) __attribute__((always_inline)) {
    src[i] = list[at + min(su + EPP * i, max(n - 1, 0))];
  };
  // X row of source slot s at frequency f: xf + s * xstride bytes (M_dim * 64 < 2^32: the
  // host bounds M_dim)
  const char* xf = reinterpret_cast<const char*>(X + (int64_t)f * RW + 4 * sr);
  const uint32_t xstride = (uint32_t)M_dim * RW * sizeof(float);
  auto load_val = [&](int i, int n) __attribute__((always_inline)) {
    const int e = su + EPP * i < n ? src[i] : POOL + 1;  // a select on the input keeps it branch-free
    dst[i] = e & 511;
    val[i] = *reinterpret_cast<const floatx4*>(xf + (uint64_t)((uint32_t)e >> 9) * xstride);
  };
  auto load_uw = [&](int k) __attribute__((always_inline)) {
    const uint2* ug = reinterpret_cast<const uint2*>(uidx + (int64_t)min(k, k1 - 1) * UT);
#pragma unroll
    for (int i = 0; i < UWP; i++) uw[i] = ug[tid + 256 * i];
  };
  // entry layout (words): part p at 8p, reals 2q, 2q+1 in word 8p + q
  auto store_lds = [&](int buf, int n) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NPASS; i++) {
      if (EPP * i >= n) break;  // uniform
      halfx4 h0, h1;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float x = val[i][q] * xsc;  // exact (power of two)
        h0[q] = (_Float16)x;
        h1[q] = (_Float16)(x - (float)h0[q]);
      }
      uint32_t* d = &Xl[dst[i] * EST + 2 * sr];
      *reinterpret_cast<halfx4*>(d) = h0;
      *reinterpret_cast<halfx4*>(d + 8) = h1;
    }
#pragma unroll
    for (int i = 0; i < UWP; i++) reinterpret_cast<uint2*>(ul[buf])[tid + 256 * i] = uw[i];
  };
  {
    // the run's first tile: its whole union
    const int u00 = tile_u0[k0], nu0 = __builtin_amdgcn_readfirstlane(tile_u0[k0 + 1] - u00);
#pragma unroll
    for (int i = 0; i < NPASS; i++) load_src(i, usrc, u00, nu0);
#pragma unroll
    for (int i = 0; i < NPASS; i++) load_val(i, nu0);
    load_uw(k0);
    __syncthreads();  // m_d
    const int nd1 = nd_l(k0 + 1), d01 = d0_l(k0 + 1);
#pragma unroll
    for (int i = 0; i < NPASS; i++) load_src(i, dent, d01, nd1);
    store_lds(0, nu0);
    // the loop's first use of src[] then waits for nothing: without this wait the compiler
    // carries the prologue's outstanding loads into the loop's wait counts, where they would
    // make every tile wait for the previous tile's Y stores
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  }
#pragma unroll
  for (int j = 0; j < NPAIR; j++)
#pragma unroll
    for (int p = 0; p < 2; p++) asm volatile("" : "+v"(bq[j][p]));
  __syncthreads();

  for (int k = k0; k < k1; k++) {
    const int buf = (k - k0) & 1;
    const int nd1 = nd_l(k + 1), nd2 = nd_l(k + 2), d02 = d0_l(k + 2);
    // offsets at which a target of the wave's group has a partner
    const uint32_t wm = __builtin_amdgcn_readfirstlane(m_g[4 * (k - k0) + w]);
    const uint32_t* xb = Xl;
    // staged positions of the partners of target (w, row) at offsets 2j + pa, j < 13, read at
    // once so that the operand reads of a pair wait for no other LDS read; the 16 rows of a
    // read are consecutive 16-byte words (one per LDS bank group)
    u32x4 pw[2];
#pragma unroll
    for (int h = 0; h < 2; h++) pw[h] = *reinterpret_cast<const u32x4*>(&ul[buf][(((w * 2 + h) * 2 + pa) * 16 + row) * 8]);
    auto pidx = [&](int j) __attribute__((always_inline)) { return (int)((pw[j >> 3][(j >> 1) & 3] >> (16 * (j & 1))) & 0xFFFFu); };
    auto aop = [&](int pos, int p) __attribute__((always_inline)) {
      return *reinterpret_cast<const halfx8*>(&xb[pos * EST + 8 * p + lane_off]);
    };
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    halfx8 an[2];
    {
      const int p0 = pidx(0);
#pragma unroll
      for (int p = 0; p < 2; p++) an[p] = aop(p0, p);
    }
#pragma unroll
    for (int j = 0; j < NPAIR; j++) {
      halfx8 ac[2];
#pragma unroll
      for (int p = 0; p < 2; p++) ac[p] = an[p];
      if (j + 1 < NPAIR) {
        const int pn = pidx(j + 1);
#pragma unroll
        for (int p = 0; p < 2; p++) an[p] = aop(pn, p);
      }
      if (j < NPASS && EPP * j < nd1) {
        load_val(j, nd1);  // tile k+1, from the entries read during tile k-1
      }
      if (j < NPASS && EPP * j < nd2) {
        load_src(j, dent, d02, nd2);  // tile k+2
      }
      if (j == NPASS) load_uw(k + 1);
      __builtin_amdgcn_sched_barrier(0);
      if ((wm >> (2 * j)) & 3u) {
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac[1], bq[j][0], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac[0], bq[j][0], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac[0], bq[j][1], acc1, 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    store_lds(buf ^ 1, nd1);  // tile k+1 (slots that tile k does not read; the other table buffer is free since the last barrier)
    {
      const int t0 = m_t[k - k0], nt = m_t[k - k0 + 1] - t0;
      const floatx4 sum = (acc0 + acc1) * ysc;
      float* yr = Y + ((int64_t)(t0 + 16 * w + 4 * kg) * M_dim + f) * RW + row;
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (16 * w + 4 * kg + r < nt) yr[(int64_t)r * M_dim * RW] = sum[r];
    }
    __syncthreads();
  }
}

// FP32 frequency-domain M2L for Stokes (3x3 blocks, 48 reals per spectrum row) on the fp16
// matrix cores, with the operand split of k_hadamard_split: the tile's partner spectra are
// scaled by the power of two s with s times their largest bound below 2^14 (the bounds are per
// source slot and component) and the transfer blocks of frequency f by tau_f, each split into
// two fp16 parts, and the three largest products are summed in FP32; the result is divided by
// s tau_f.  Offsets go in pairs (2j, 2j+1) of the offset list: the pair's K = 96 (48 reals of
// each partner) is three K = 32 chunks, and lane (row, kg) of chunk c takes reals 8b..8b+7 of
// the partner at offset 2j + a, with q = 4c + kg, a = q / 6, b = q % 6.  A pair of a 16-target
// group then costs 27 MFMAs of K = 32 (3 chunks x 3 output tiles x 3 products) instead of 72
// FP32 MFMAs of K = 4, at half the cycles each.
// A block owns one tile (at most 64 targets in 4 groups of 16, at most UMAX partners) and one
// frequency; the two frequencies of a pair take consecutive blocks of each tile on one XCD, so
// the 384 bytes of a partner's two spectrum rows are read into that XCD's L2 once.  Wave w
// computes the K chunks w, w + 4, ... of the 39 (13 pairs x 3) for all 4 groups, reading the
// transfer-block parts of a chunk from global memory two chunks ahead; chunks at which none of
// a group's targets has a partner are skipped.  The waves' partial sums are added through LDS
// in wave order, wave w adding and writing group w.
template <int UMAX>
__global__ __launch_bounds__(256, 1) void k_hadamard_split3(int M_dim, int n_tiles, const int32_t* __restrict__ tile_t0,
                                                           const int32_t* __restrict__ tile_u0,
                                                           const int32_t* __restrict__ usrc, const uint16_t* __restrict__ uidx,
                                                           const uint32_t* __restrict__ gmask, const uint16_t* __restrict__ Vh,
                                                           const float* __restrict__ tinv, const float* __restrict__ bound,
                                                           const float* __restrict__ X, float* __restrict__ Y) {
  constexpr int NOFF = 26, NPAIR = NOFF / 2, RW = 48;
  constexpr int NB = 3, NCH = 3;          // output tiles of 16 reals, K chunks of a pair
  constexpr int EST = 52;                 // words per staged entry: 2 parts x 24 words, padded
  constexpr int V4E = RW / 4;             // 4-vectors of a spectrum row
  constexpr int NPASS = (UMAX * V4E + 255) / 256;
  constexpr int XW = (UMAX + 1) * EST;    // staging (entry UMAX stays zero)
  constexpr int RED = 4 * 4 * NB * 64 * 4;  // partial sums [wave][group][tile][lane][4]
  constexpr int LW = XW > RED ? XW : RED;
  constexpr int UW = 64 * NOFF / 4;       // partner table as 8-byte words
  __shared__ __align__(16) uint32_t Xl[LW];
  __shared__ __align__(8) uint16_t ul[64 * NOFF];
  __shared__ float redm[4];
  const int b = blockIdx.x, xcd = b & 7, local = b >> 3;
  const int fp = (local / (2 * n_tiles)) * 8 + xcd, rem = local % (2 * n_tiles);
  const int tile = rem >> 1, f = 2 * fp + (rem & 1);
  if (f >= M_dim) return;  // whole block
  const int t0 = tile_t0[tile], nt = tile_t0[tile + 1] - t0;
  const int u0 = tile_u0[tile], nu = tile_u0[tile + 1] - u0;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, row = l & 15, kg = l >> 4;
  // transfer-block parts of frequency f: [pair][chunk][tile][part] blocks of 64 lanes x 8, so
  // global chunk ci = 3 j + c starts at block 6 ci.  Wave w takes chunks w, w + 4, ... (10, 10,
  // 10 and 9 of the 39), their parts read PF chunks ahead into a ring.
  constexpr int NCK = NPAIR * NCH, CPW = (NCK + 3) / 4, BPC = NB * 2, PF = 2;
  const halfx8* vb = reinterpret_cast<const halfx8*>(Vh) + (int64_t)f * NCK * BPC * 64 + l;
  halfx8 bq[PF + 1][BPC];
#pragma unroll
  for (int it = 0; it < PF; it++)
#pragma unroll
    for (int i = 0; i < BPC; i++) bq[it][i] = vb[((int64_t)(w + 4 * it) * BPC + i) * 64];  // w + 4 (PF - 1) < NCK
  // partner spectra of the tile, and the largest bound of their components
  floatx4 val[NPASS];
  float bmx = 0.f;
#pragma unroll
  for (int i = 0; i < NPASS; i++) {
    const int e = tid + 256 * i, u = e / V4E, r = e - u * V4E;
    val[i] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (u < nu) {
      const int s = usrc[u0 + u];
      val[i] = *reinterpret_cast<const floatx4*>(X + ((int64_t)s * M_dim + f) * RW + 4 * r);
      if (r < 3) bmx = fmaxf(bmx, bound[(int64_t)s * 3 + r]);
    }
  }
  {
    const uint2* ug = reinterpret_cast<const uint2*>(uidx + (int64_t)tile * 64 * NOFF);
    for (int e = tid; e < UW; e += 256) reinterpret_cast<uint2*>(ul)[e] = ug[e];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bmx = fmaxf(bmx, __shfl_xor(bmx, o));
  if (l == 0) redm[w] = bmx;
  for (int e = tid; e < EST; e += 256) Xl[UMAX * EST + e] = 0u;
  __syncthreads();
  const float bm = fmaxf(fmaxf(redm[0], redm[1]), fmaxf(redm[2], redm[3]));
  float xsc = 1.f;
  if (bm > 0.f && bm <= 3.4e38f) {
    int ex = 0;
    (void)frexpf(bm, &ex);  // bm < 2^ex
    xsc = ldexpf(1.f, min(100, max(-100, 14 - ex)));
  }
  // entry layout (words): part p at 24p, reals 2q, 2q+1 in word 24p + q
#pragma unroll
  for (int i = 0; i < NPASS; i++) {
    const int e = tid + 256 * i, u = e / V4E, r = e - u * V4E;
    if (u < nu) {
      halfx4 h0, h1;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float x = val[i][q] * xsc;  // exact (power of two)
        h0[q] = (_Float16)x;
        h1[q] = (_Float16)(x - (float)h0[q]);
      }
      uint32_t* d = &Xl[u * EST + 2 * r];
      *reinterpret_cast<halfx4*>(d) = h0;
      *reinterpret_cast<halfx4*>(d + 24) = h1;
    }
  }
  uint32_t gm[4];
#pragma unroll
  for (int g = 0; g < 4; g++) gm[g] = gmask[4 * tile + g];
  __syncthreads();
  floatx4 acc[4][NB];
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[g][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int it = 0; it < CPW; it++) {
    const int ci = w + 4 * it;
    if (ci >= NCK) break;  // wave-uniform
    if (ci + 4 * PF < NCK) {
#pragma unroll
      for (int i = 0; i < BPC; i++) bq[(it + PF) % (PF + 1)][i] = vb[((int64_t)(ci + 4 * PF) * BPC + i) * 64];
    }
    const halfx8* bc = bq[it % (PF + 1)];
    const int j = ci / NCH, c = ci - NCH * j;
    const int q = 4 * c + kg, a = q / 6, bo = q - 6 * a;
    const int oi = 2 * j + a;
    halfx8 A[4][2];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int pos = ul[(oi * 16 + row) * 4 + g];  // a missing partner reads the zero entry
#pragma unroll
      for (int p = 0; p < 2; p++) A[g][p] = *reinterpret_cast<const halfx8*>(&Xl[pos * EST + 24 * p + 4 * bo]);
    }
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const uint32_t pm = (gm[g] >> (2 * j)) & 3u;  // offsets 2j, 2j+1 present in group g
      if (!(c == 0 ? (pm & 1u) : c == 2 ? (pm & 2u) : pm)) continue;  // block-uniform
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        acc[g][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[g][0], bc[2 * nb], acc[g][nb], 0, 0, 0);
        acc[g][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[g][1], bc[2 * nb], acc[g][nb], 0, 0, 0);
        acc[g][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[g][0], bc[2 * nb + 1], acc[g][nb], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // every wave is done reading the staged spectra
  float* red = reinterpret_cast<float*>(Xl);
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) *reinterpret_cast<floatx4*>(&red[(((w * 4 + g) * NB + nb) * 64 + l) * 4]) = acc[g][nb];
  __syncthreads();
  const float ysc = tinv[f] / xsc;
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    floatx4 s = *reinterpret_cast<const floatx4*>(&red[((w * NB + nb) * 64 + l) * 4]);
#pragma unroll
    for (int v = 1; v < 4; v++) s += *reinterpret_cast<const floatx4*>(&red[(((v * 4 + w) * NB + nb) * 64 + l) * 4]);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int tg = 16 * w + 4 * kg + r;
      if (tg < nt) Y[((int64_t)(t0 + tg) * M_dim + f) * RW + 16 * nb + row] = s[r] * ysc;
    }
  }
}

// The same products for 16 density vectors of one tree at once: the MFMA rows are the 16
// vectors of one target (instead of 16 targets of one vector), so the 16 rows of an operand
// read one partner's spectra, 16 consecutive 64-byte rows of X in the vector-group layout
// [slot][f][vector 16][16 reals]; no partner table, no padding of absent partners and no LDS
// staging are needed, the loads are whole 1 KiB blocks straight from L2.  A block owns one
// frequency (its transfer-block parts in registers) and 4 waves; wave w computes targets
// [t0 + w tpw, t0 + (w + 1) tpw) in turn, over the 13 offset pairs each, absent partners
// read as zero.
template <int NPAIR>
__global__ __launch_bounds__(256, 2) void k_hadamard_vec16(int M_dim, int n_trg, int tpw, int n_chunks,
                                                          const int32_t* __restrict__ partner, const uint16_t* __restrict__ Vb,
                                                          const float* __restrict__ X, float* __restrict__ Y) {
  constexpr int NOFF = 2 * NPAIR;
  // all target chunks of one frequency run on one XCD, back to back (block b runs on XCD b mod
  // 8), so the partners' spectra at that frequency are read into that XCD's L2 once and served
  // from there to every target that lists them
  const int b = blockIdx.x, xcd = b & 7, local = b >> 3;
  const int chunk = local % n_chunks, f = (local / n_chunks) * 8 + xcd;
  if (f >= M_dim) return;  // whole block
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int row = l & 15, kg = l >> 4, pa = kg >> 1, half = kg & 1;
  bf16x8 bq[NPAIR][3];
#pragma unroll
  for (int j = 0; j < NPAIR; j++)
#pragma unroll
    for (int p = 0; p < 3; p++)
      bq[j][p] = *reinterpret_cast<const bf16x8*>(Vb + ((((int64_t)j * 3 + p) * M_dim + f) * 64 + l) * 8);
  const int t_begin = (chunk * 4 + w) * tpw, t_end = min(n_trg, t_begin + tpw);
  for (int t = t_begin; t < t_end; t++) {
    const int32_t* pt = partner + (int64_t)t * NOFF;
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < NPAIR; j++) {
      const int sl = pt[2 * j + pa];
      floatx4 v0 = {0.f, 0.f, 0.f, 0.f}, v1 = {0.f, 0.f, 0.f, 0.f};
      if (sl >= 0) {
        const float* src = X + (((int64_t)sl * M_dim + f) * 16 + row) * 16 + 8 * half;
        v0 = *reinterpret_cast<const floatx4*>(src);
        v1 = *reinterpret_cast<const floatx4*>(src + 4);
      }
      uint32_t h[3][8];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        split3(v0[q], h[0][q], h[1][q], h[2][q]);
        split3(v1[q], h[0][q + 4], h[1][q + 4], h[2][q + 4]);
      }
      bf16x8 a[3];
#pragma unroll
      for (int p = 0; p < 3; p++) {
        const u32x4 wd = {pack_hi(h[p][0], h[p][1]), pack_hi(h[p][2], h[p][3]), pack_hi(h[p][4], h[p][5]),
                          pack_hi(h[p][6], h[p][7])};
        a[p] = __builtin_bit_cast(bf16x8, wd);
      }
      acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bq[j][0], acc1, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bq[j][0], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bq[j][1], acc1, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bq[j][1], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bq[j][2], acc1, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bq[j][0], acc0, 0, 0, 0);
    }
    const floatx4 sum = acc0 + acc1;
#pragma unroll
    for (int r = 0; r < 4; r++) Y[(((int64_t)t * M_dim + f) * 16 + 4 * kg + r) * 16 + row] = sum[r];
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
// zeroes n values of T: 16-byte stores over the aligned whole words, scalar stores for the rest
template <typename T>
static void fill_any(hipStream_t s, T* p, int64_t n) {
  if (n == 0) return;
  int64_t head = 0;
  if ((reinterpret_cast<uintptr_t>(p) % 16) == 0) {
    const int64_t n16 = n * (int64_t)sizeof(T) / 16;
    if (n16 > 0) {
      hipLaunchKernelGGL(k_fill16, dim3(grid_for(n16, 256)), dim3(256), 0, s, reinterpret_cast<uint4*>(p), n16);
      HIP_LAUNCH_CHECK();
    }
    head = n16 * 16 / (int64_t)sizeof(T);
  }
  if (head < n) {
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n - head, 256)), dim3(256), 0, s, p + head, n - head);
    HIP_LAUNCH_CHECK();
  }
}
template <typename Real>
void DeviceKernels<Real>::fill(hipStream_t s, Real* p, int64_t n) {
  fill_any<Real>(s, p, n);
}
template <typename Real>
void DeviceKernels<Real>::vol_correction(hipStream_t s, int kind, int64_t n, const double* x, const Real* eq_root, int K0,
                                         int e0, double* val, int nv, int64_t eq_vs, int64_t val_vs) {
  if (n == 0 || nv <= 0) return;
  if ((kind != 0 && kind != 2) || e0 < 1 || e0 > 3) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  hipLaunchKernelGGL(k_vol_correction<Real>, dim3(std::min(grid_for(n, 256), 1024), nv), dim3(256), 0, s, kind, n, x, eq_root,
                     K0, e0, val, eq_vs, val_vs);
  HIP_LAUNCH_CHECK();
}
template <typename Real>
int DeviceKernels<Real>::vol_adjoint_blocks() {
  return kVolAdjBlocks;
}
template <typename Real>
void DeviceKernels<Real>::vol_correction_adjoint(hipStream_t s, int kind, int64_t n, const double* x, const Real* g, int K0,
                                                 int e0, Real* eq_root, double* part, int nv, int64_t g_vs, int64_t eq_vs) {
  if (n == 0 || nv <= 0) return;
  if ((kind != 0 && kind != 2) || e0 < 1 || e0 > 3 || !part) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  hipLaunchKernelGGL(k_vol_adjoint_partial<Real>, dim3(kVolAdjBlocks, nv), dim3(256), 0, s, kind, n, x, g, g_vs, part);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_vol_adjoint_apply<Real>, dim3(nv), dim3(256), 0, s, kVolAdjBlocks, (const double*)part, K0, e0, eq_root,
                     eq_vs);
  HIP_LAUNCH_CHECK();
}
template <typename Real>
void DeviceKernels<Real>::fill64(hipStream_t s, double* p, int64_t n) {
  fill_any<double>(s, p, n);
}

// Density vectors in chunks of 8, 4, 2 and 1 per launch (the chunk shares the geometry).
template <typename F>
static void for_density_chunks(int nv, F&& launch) {
  int v0 = 0;
  while (v0 < nv) {
    const int left = nv - v0, vb = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    launch(v0, vb);
    v0 += vb;
  }
}

template <typename Real>
void DeviceKernels<Real>::p2p_surface(hipStream_t s, int kind, int n_rows, const int32_t* tnode, const int32_t* item_ptr,
                                      const int32_t* items, const int32_t* src_begin, const int32_t* src_end,
                                      const double* center, const float* node_size, double rad, const Real* unit_surf, int ns,
                                      const Real* sx, const Real* sden, int64_t n_src, Real* out, const int32_t* orow,
                                      int64_t ld, bool accumulate, int nv, int64_t den_vs, int64_t out_vs) {
  if (n_rows == 0 || nv <= 0) return;
  if (kind != 0 && kind != 2) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  for_density_chunks(nv, [&](int v0, int vb) {
    const Real* d = sden + v0 * den_vs;
    Real* o = out + v0 * out_vs;
    // two surface points per thread for one density vector, one for several
#define SURF_LAUNCH(KD, VB)                                                                                                 \
  {                                                                                                                         \
    constexpr int PPT = VB == 1 ? 2 : 1;                                                                                    \
    const dim3 g(n_rows, (ns + 256 * PPT - 1) / (256 * PPT));                                                               \
    hipLaunchKernelGGL((k_p2p_surface<Real, KD, VB, PPT>), g, dim3(256), 0, s, tnode, item_ptr, items, src_begin, src_end, \
                       center, node_size, (Real)rad, unit_surf, ns, sx, d, n_src, den_vs, o, out_vs, orow, ld, accumulate); \
  }
#define SURF_VB(KD)                   \
  switch (vb) {                       \
    case 8: SURF_LAUNCH(KD, 8); break; \
    case 4: SURF_LAUNCH(KD, 4); break; \
    case 2: SURF_LAUNCH(KD, 2); break; \
    default: SURF_LAUNCH(KD, 1);       \
  }
    if (kind == 0) { SURF_VB(0) } else { SURF_VB(2) }
#undef SURF_VB
#undef SURF_LAUNCH
    HIP_LAUNCH_CHECK();
  });
}

template <typename Real>
void DeviceKernels<Real>::p2p_points(hipStream_t s, int kind, int n_rows, const int32_t* tleaf, const int32_t* item_ptr,
                                     const int32_t* items, const int32_t* trg_begin, const int32_t* trg_end,
                                     const int32_t* src_begin, const int32_t* src_end, const double* center, const double* tx,
                                     int64_t n_trg, const double* sx, const Real* sden, int64_t n_src, double* out,
                                     bool accumulate, int nv, int64_t den_vs, int64_t out_vs) {
  if (n_rows == 0 || nv <= 0) return;
  for_density_chunks(nv, [&](int v0, int vb) {
    const Real* d = sden + v0 * den_vs;
    double* o = out + v0 * out_vs;
#define LEAF_LAUNCH(KD, VB)                                                                                                 \
  hipLaunchKernelGGL((k_p2p_leaf<Real, KD, VB>), dim3(n_rows), dim3(256), 0, s, tleaf, item_ptr, items, trg_begin, trg_end, \
                     src_begin, src_end, center, tx, n_trg, sx, d, n_src, den_vs, o, out_vs, accumulate)
#define LEAF_VB(KD)                   \
  switch (vb) {                       \
    case 8: LEAF_LAUNCH(KD, 8); break; \
    case 4: LEAF_LAUNCH(KD, 4); break; \
    case 2: LEAF_LAUNCH(KD, 2); break; \
    default: LEAF_LAUNCH(KD, 1);       \
  }
    if (kind == 0) { LEAF_VB(0) } else if (kind == 1) { LEAF_VB(1) } else { LEAF_VB(2) }
#undef LEAF_VB
#undef LEAF_LAUNCH
    HIP_LAUNCH_CHECK();
  });
}

template <typename Real>
void DeviceKernels<Real>::far_points(hipStream_t s, int kind, int n_rows, const int32_t* tleaf, const int32_t* item_ptr,
                                     const int32_t* item_node, const int64_t* item_den, const double* item_rad,
                                     const int32_t* trg_begin, const int32_t* trg_end, const double* center,
                                     const float* node_size, const Real* unit_surf, int ns, const Real* tx, int64_t n_trg,
                                     const Real* dens, double* out, bool accumulate, int nv, int64_t den_vs, int64_t out_vs) {
  if (n_rows == 0 || nv <= 0) return;
  for_density_chunks(nv, [&](int v0, int vb) {
    const Real* d = dens + v0 * den_vs;
    double* o = out + v0 * out_vs;
#define FAR_LAUNCH(KD, VB)                                                                                                   \
  hipLaunchKernelGGL((k_far_leaf<Real, KD, VB>), dim3(n_rows), dim3(256), 0, s, tleaf, item_ptr, item_node, item_den, item_rad, \
                     trg_begin, trg_end, center, node_size, unit_surf, ns, tx, n_trg, d, den_vs, o, out_vs, accumulate)
#define FAR_VB(KD)                   \
  switch (vb) {                      \
    case 8: FAR_LAUNCH(KD, 8); break; \
    case 4: FAR_LAUNCH(KD, 4); break; \
    case 2: FAR_LAUNCH(KD, 2); break; \
    default: FAR_LAUNCH(KD, 1);       \
  }
    if (kind == 0) { FAR_VB(0) } else if (kind == 1) { FAR_VB(1) } else { FAR_VB(2) }
#undef FAR_VB
#undef FAR_LAUNCH
    HIP_LAUNCH_CHECK();
  });
}

template <typename Real>
void DeviceKernels<Real>::gemm_rows(hipStream_t s, int M, int N, int K, const Real* A, const int32_t* ra, int64_t lda,
                                    const Real* B, Real* C, const int32_t* rc, int64_t ldc, const Real* alpha, bool accumulate,
                                    Real* work, int64_t work_elems, int split_blocks, int nv, int64_t avs, int64_t cvs) {
  if (M <= 0 || N <= 0 || K <= 0 || nv <= 0) return;
  const int Mv = M;
  if ((int64_t)M * nv >= (1ll << 31)) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  M *= nv;
  const int gx = (N + 63) / 64, gy = (M + 63) / 64;
  // split K when the tile grid cannot fill the 256 CUs, in chunks of whole 16-wide k steps
  int ksplit = 1;
  const int blocks = gx * gy;
  if ((int64_t)blocks + 8 >= (1ll << 31)) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  if (work && blocks < split_blocks) {
    ksplit = std::min((2 * split_blocks + blocks - 1) / blocks, (K + 63) / 64);
    while (ksplit > 1 && (int64_t)ksplit * M * N > work_elems) ksplit--;
  }
  const int kchunk = ((K + ksplit - 1) / ksplit + 31) / 32 * 32;
  ksplit = (K + kchunk - 1) / kchunk;
  dim3 g((unsigned)((blocks + 7) / 8 * 8), 1, ksplit);
  // 4-vector loads when every row start and the B rows are 16-byte aligned
  const bool vec = lda % 4 == 0 && avs % 4 == 0 && N % 4 == 0 && (reinterpret_cast<uintptr_t>(A) % 32) == 0 &&
                   (reinterpret_cast<uintptr_t>(B) % 32) == 0;
#define GEMM_LAUNCH(ACCV, SPLITV)                                                                                     \
  do {                                                                                                                \
    if (vec)                                                                                                          \
      hipLaunchKernelGGL((k_gemm_rows<Real, ACCV, SPLITV, true>), g, dim3(256), 0, s, M, N, K, kchunk, A, ra, lda, B, C, rc, \
                         ldc, alpha, work, Mv, avs, cvs);                                                             \
    else                                                                                                              \
      hipLaunchKernelGGL((k_gemm_rows<Real, ACCV, SPLITV, false>), g, dim3(256), 0, s, M, N, K, kchunk, A, ra, lda, B, C, rc, \
                         ldc, alpha, work, Mv, avs, cvs);                                                             \
  } while (0)
  if (ksplit > 1) {
    GEMM_LAUNCH(false, true);
    HIP_LAUNCH_CHECK();
    const int64_t tot = (int64_t)M * N;
    if (accumulate)
      hipLaunchKernelGGL((k_gemm_reduce<Real, true>), dim3(grid_for(tot, 256)), dim3(256), 0, s, M, N, ksplit, work, C, rc, ldc, alpha,
                         Mv, cvs);
    else
      hipLaunchKernelGGL((k_gemm_reduce<Real, false>), dim3(grid_for(tot, 256)), dim3(256), 0, s, M, N, ksplit, work, C, rc, ldc, alpha,
                         Mv, cvs);
  } else if (accumulate) {
    GEMM_LAUNCH(true, false);
  } else {
    GEMM_LAUNCH(false, false);
  }
#undef GEMM_LAUNCH
  HIP_LAUNCH_CHECK();
}

// children per block: all 8 when the LDS working set fits in 64 KiB, else fewer
template <typename Real>
static int fft_children_per_block(size_t per_child, size_t fixed) {
  for (int cb = 8; cb > 1; cb >>= 1)
    if (fixed + cb * per_child <= 64 * 1024) return cb;
  return 1;
}

template <typename Real>
void DeviceKernels<Real>::fft_up(hipStream_t s, int n_slots, const int32_t* first_child, const Real* up, int K0, int e0, int m,
                                 const int32_t* lattice, const int32_t* lat_inv, const Real* scal, Real* X, int vg, int vi,
                                 float* bound) {
  if (n_slots == 0) return;
  if (bound && vg != 1) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  const dim3 g((unsigned)((n_slots + 7) / 8 * 8 * e0));  // whole XCD rounds of slots
#define PVFMM_FFT_UP_M(MM)                                                                                        \
  case MM:                                                                                                        \
    if (e0 == 1 && vg == 1)                                                                                       \
      hipLaunchKernelGGL((k_fft_up_m<Real, MM, true>), g, dim3(DftShape<Real, MM>::NT), 0, s, first_child, up, K0, e0, \
                         lat_inv, scal, X, n_slots, vg, vi, bound);                                               \
    else                                                                                                          \
      hipLaunchKernelGGL((k_fft_up_m<Real, MM, false>), g, dim3(DftShape<Real, MM>::NT), 0, s, first_child, up, K0, e0, \
                         lat_inv, scal, X, n_slots, vg, vi, bound);                                               \
    HIP_LAUNCH_CHECK();                                                                                           \
    return;
  switch (m) {
    PVFMM_FFT_UP_M(4)
    PVFMM_FFT_UP_M(6)
    PVFMM_FFT_UP_M(8)
    PVFMM_FFT_UP_M(10)
    PVFMM_FFT_UP_M(12)
    default: break;
  }
#undef PVFMM_FFT_UP_M
  const int n1 = 2 * m;
  const size_t per_child = sizeof(Real) * (m * m * m + 2 * m * m + 2 * m * n1), fixed = sizeof(Real) * 2 * n1;
  const int cb = fft_children_per_block<Real>(per_child, fixed);
  const size_t lds = fixed + cb * per_child;
  hipLaunchKernelGGL(k_fft_up<Real>, dim3(n_slots * e0 * (8 / cb)), dim3(512), lds, s, first_child, up, K0, e0, m, cb, lattice,
                     scal, X, vg, vi);
  HIP_LAUNCH_CHECK();
  if (bound) {
    hipLaunchKernelGGL(k_spectrum_bound<Real>, dim3(n_slots), dim3(256), 0, s, n_slots, first_child, up, K0, e0,
                       6 * (m - 1) * (m - 1) + 2, scal, bound);
    HIP_LAUNCH_CHECK();
  }
}

template <typename Real>
int DeviceKernels<Real>::hadamard_union_max(int e0) {
  // scalar kernels: FP32 250, FP64 256 partners per tile; Stokes: 160, which fits three FP32
  // blocks (one frequency each, fp16 parts) or one FP64 block (2 frequencies) in a CU's 160 KiB
  // of LDS (fuller tiles pay for the per-tile transfer-block loads over more targets)
  if (e0 != 1) return 160;
  return sizeof(Real) == 4 ? 250 : 256;
}

template <typename Real>
int DeviceKernels<Real>::hadamard_pool_max() {
  return kSplitPool;
}

template <typename Real>
void DeviceKernels<Real>::hadamard_tiled(hipStream_t s, int n_tiles, int M_dim, int e0, int e1, int n_off, const int32_t* tile_t0,
                                         const int32_t* tile_u0, const int32_t* usrc, const int32_t* tile_d0,
                                         const int32_t* dent, const uint16_t* uidx, const uint32_t* gmask,
                                         const int32_t* offsets, const Real* Vp, const Real* X, Real* Y, int tiles_per_block,
                                         const HadamardParts* hp) {
  if (n_tiles == 0) return;
  if (n_off != 26) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);  // the 26 V1 parent offsets
  constexpr int FB = sizeof(Real) == 4 ? 4 : 2;
  const int n_fg = (M_dim + FB - 1) / FB;
  const int64_t nb = (int64_t)n_tiles * ((n_fg + 7) / 8) * 8;
  if (nb >= (1ll << 31)) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  if (e0 == 1 && e1 == 1) {
    if constexpr (std::is_same<Real, float>::value) {
      // runs of tiles_per_block (at most 64) tiles per block, one frequency per block
      if (!hp || !hp->Vh || !hp->tinv || !hp->bound || !hp->run_scale || (int64_t)M_dim * 64 >= (1ll << 32))
        throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
      const int TPB = std::min(64, std::max(1, tiles_per_block));
      const int n_chunks = (n_tiles + TPB - 1) / TPB;
      const int64_t nbs = (int64_t)n_chunks * ((M_dim + 15) / 16) * 16;
      if (nbs >= (1ll << 31) || n_chunks > hp->run_scale_cap) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
      hipLaunchKernelGGL(k_run_scale, dim3((unsigned)n_chunks), dim3(256), 0, s, n_tiles, TPB, tile_u0, usrc, tile_d0, dent,
                         hp->bound, hp->run_scale);
      HIP_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_hadamard_split<250, kSplitPool>), dim3((unsigned)nbs), dim3(256), 0, s, M_dim, n_tiles, TPB,
                         n_chunks, tile_t0, tile_u0, usrc, tile_d0, dent, uidx, gmask, hp->Vh, hp->tinv, hp->run_scale, X, Y);
    } else {
      const int n_fgt = (M_dim + 1) / 2;
      const int64_t nbt = (int64_t)n_tiles * ((n_fgt + 7) / 8) * 8;
      if (nbt >= (1ll << 31)) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
      hipLaunchKernelGGL((k_hadamard_tiled<Real, 1, 1, 2, 256, 1>), dim3((unsigned)nbt), dim3(256), 0, s, M_dim, n_tiles,
                         tile_t0, tile_u0, usrc, uidx, gmask, offsets, Vp, X, Y);
    }
  } else if (e0 == 3 && e1 == 3) {
    // tiles of up to 160 partners; FP32: one frequency per block on the fp16 matrix cores,
    // FP64: 2 frequencies per block
    constexpr int UMS = 160;
    const int n_fg2 = (M_dim + 1) / 2;
    const int64_t nb2 = (int64_t)n_tiles * ((n_fg2 + 7) / 8) * 8;
    if (2 * nb2 >= (1ll << 31)) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
    if constexpr (std::is_same<Real, float>::value) {
      if (!hp || !hp->Vh || !hp->tinv || !hp->bound) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
      hipLaunchKernelGGL((k_hadamard_split3<UMS>), dim3((unsigned)(2 * nb2)), dim3(256), 0, s, M_dim, n_tiles, tile_t0,
                         tile_u0, usrc, uidx, gmask, hp->Vh, hp->tinv, hp->bound, X, Y);
    } else {
      hipLaunchKernelGGL((k_hadamard_tiled<Real, 3, 3, 2, UMS, 1>), dim3((unsigned)nb2), dim3(256), 0, s, M_dim, n_tiles,
                         tile_t0, tile_u0, usrc, uidx, gmask, offsets, Vp, X, Y);
    }
  } else {
    throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
  }
  HIP_LAUNCH_CHECK();
}

template <typename Real>
void DeviceKernels<Real>::hadamard_vec16(hipStream_t s, int n_trg, int M_dim, const int32_t* partner, const uint16_t* Vb,
                                         const Real* X, Real* Y, int targets_per_wave) {
  if (n_trg == 0) return;
  if constexpr (std::is_same<Real, float>::value) {
    if (!Vb) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
    const int tpw = std::max(1, targets_per_wave);
    const int n_chunks = (n_trg + 4 * tpw - 1) / (4 * tpw);
    const int64_t nb = (int64_t)n_chunks * ((M_dim + 7) / 8) * 8;
    if (nb >= (1ll << 31)) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);
    hipLaunchKernelGGL(k_hadamard_vec16<13>, dim3((unsigned)nb), dim3(256), 0, s, M_dim, n_trg, tpw, n_chunks, partner, Vb,
                       X, Y);
    HIP_LAUNCH_CHECK();
  } else {
    throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);  // FP32 scalar kernels only
  }
}

template <typename Real>
void DeviceKernels<Real>::ifft_down(hipStream_t s, int n_slots, const int32_t* first_child, const Real* Y, Real* dn, int K1, int e1,
                                    int m, const int32_t* lattice, const int32_t* lat_inv, int ns, const Real* scal, int vg,
                                    int vi) {
  if (n_slots == 0) return;
  const dim3 g((unsigned)((n_slots + 7) / 8 * 8 * e1));  // whole XCD rounds of slots
#define PVFMM_IFFT_DN_M(MM)                                                                                       \
  case MM:                                                                                                        \
    hipLaunchKernelGGL((k_ifft_down_m<Real, MM>), g, dim3(DftShape<Real, MM>::NT), 0, s, first_child, Y, dn, K1, e1,     \
                       lat_inv, scal, n_slots, vg, vi);                                                           \
    HIP_LAUNCH_CHECK();                                                                                           \
    return;
  switch (m) {
    PVFMM_IFFT_DN_M(4)
    PVFMM_IFFT_DN_M(6)
    PVFMM_IFFT_DN_M(8)
    PVFMM_IFFT_DN_M(10)
    PVFMM_IFFT_DN_M(12)
    default: break;
  }
#undef PVFMM_IFFT_DN_M
  const int n1 = 2 * m;
  const size_t per_child = sizeof(Real) * (2 * n1 * n1 + 2 * n1 * m + 2 * m * m), fixed = sizeof(Real) * 2 * n1;
  const int cb = fft_children_per_block<Real>(per_child, fixed);
  if (cb * ns > 24 * 512) throw_device_error(hipErrorInvalidValue, __FILE__, __LINE__);  // register accumulators
  const size_t lds = fixed + cb * per_child;
  hipLaunchKernelGGL(k_ifft_down<Real>, dim3(n_slots * e1 * (8 / cb)), dim3(512), lds, s, first_child, Y, dn, K1, e1, m, cb, lattice,
                     ns, scal, vg, vi);
  HIP_LAUNCH_CHECK();
}

template struct DeviceKernels<float>;
template struct DeviceKernels<double>;

}  // namespace pvfmm_amd
