// FP64 operator precompute on the host.  Pseudo-inverses come from a one-sided (Hestenes)
// Jacobi SVD written here (high relative accuracy for the ill-conditioned equivalent-surface
// systems); M2L transfer spectra come from a separable real DFT written here.  Both are
// independent of the oracle's numpy/LAPACK path that checks them.
#include "operators.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <limits>
#include <memory>
#include <string>

#include <omp.h>

#include <hip/hip_runtime_api.h>

#include "device_fmm.hpp"
#include "octree.hpp"
#include "kernels.hpp"
#include "precompute.hpp"

namespace pvfmm_amd {

// ----------------------------------------------------------------------------- kernels
static inline void eval_pair(int kind, const double r[3], double* out /* e0 x e1 */) {
  const double pi = 3.14159265358979323846;
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (!(r2 > 0)) {
    int n = kind == 2 ? 9 : (kind == 1 ? 3 : 1);
    for (int i = 0; i < n; i++) out[i] = 0;
    return;
  }
  double rinv = 1.0 / std::sqrt(r2);
  if (kind == 0) { out[0] = rinv / (4 * pi); return; }
  if (kind == 1) {
    double f = rinv * rinv * rinv / (4 * pi);
    out[0] = r[0] * f; out[1] = r[1] * f; out[2] = r[2] * f;
    return;
  }
  // stokes velocity: u_j = (f_j + r_j (f.r)/r^2)/r/(8 pi); entry [i][j] for force component i
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) out[3 * i + j] = rinv * ((i == j) + r[i] * r[j] * rinv * rinv) / (8 * pi);
}

static inline void kind_dims(int kind, int& k0, int& k1) {
  k0 = kind == 2 ? 3 : 1;
  k1 = kind == 0 ? 1 : 3;
}

void kernel_matrix(int kind, const double* src, int n_src, const double* trg, int n_trg, double* out) {
  int k0, k1;
  kind_dims(kind, k0, k1);
  #pragma omp parallel for schedule(static)
  for (int s = 0; s < n_src; s++) {
    double blk[9];
    for (int t = 0; t < n_trg; t++) {
      double r[3] = {trg[3 * t] - src[3 * s], trg[3 * t + 1] - src[3 * s + 1], trg[3 * t + 2] - src[3 * s + 2]};
      eval_pair(kind, r, blk);
      for (int i = 0; i < k0; i++)
        for (int j = 0; j < k1; j++) out[((size_t)s * k0 + i) * ((size_t)n_trg * k1) + (size_t)t * k1 + j] = blk[i * k1 + j];
    }
  }
}

// ----------------------------------------------------------------------------- Jacobi SVD
void pinv_jacobi(const std::vector<double>& A, int r, int c, double cutoff_rel, std::vector<double>& F0,
                 std::vector<double>& F1, int& rank_out) {
  // columns of A as rows of G; V^T as rows of W
  const int cp = c + (c & 1);  // padded to even for the round-robin pairing
  std::vector<double> G((size_t)cp * r, 0.0), W((size_t)cp * c, 0.0);
  for (int i = 0; i < r; i++)
    for (int j = 0; j < c; j++) G[(size_t)j * r + i] = A[(size_t)i * c + j];
  for (int j = 0; j < c; j++) W[(size_t)j * c + j] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<int> ring(cp);
  for (int i = 0; i < cp; i++) ring[i] = i;
  for (int sweep = 0; sweep < 60; sweep++) {
    int rotations = 0;
    for (int round = 0; round < cp - 1; round++) {
      #pragma omp parallel for schedule(static) reduction(+ : rotations)
      for (int k = 0; k < cp / 2; k++) {
        int p = ring[k], q = ring[cp - 1 - k];
        if (p >= c || q >= c) continue;
        if (p > q) std::swap(p, q);
        double* gp = &G[(size_t)p * r];
        double* gq = &G[(size_t)q * r];
        double a = 0, b = 0, g = 0;
        for (int i = 0; i < r; i++) { a += gp[i] * gp[i]; b += gq[i] * gq[i]; g += gp[i] * gq[i]; }
        if (g == 0 || std::fabs(g) <= eps * std::sqrt(a * b)) continue;
        double zeta = (b - a) / (2 * g);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
        double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
        for (int i = 0; i < r; i++) { double x = gp[i], y = gq[i]; gp[i] = cs * x - sn * y; gq[i] = sn * x + cs * y; }
        double* wp = &W[(size_t)p * c];
        double* wq = &W[(size_t)q * c];
        for (int i = 0; i < c; i++) { double x = wp[i], y = wq[i]; wp[i] = cs * x - sn * y; wq[i] = sn * x + cs * y; }
        rotations++;
      }
      // rotate the ring keeping element 0 fixed
      int last = ring[cp - 1];
      for (int i = cp - 1; i > 1; i--) ring[i] = ring[i - 1];
      ring[1] = last;
    }
    if (rotations == 0) break;
  }
  pinv_factors(G, W, r, c, cutoff_rel, F0, F1, rank_out);
}

// The pseudo-inverse factors from converged Jacobi rotations: G (c rows of length r) holds
// U S as rows, W (c rows of length c) holds V^T.  Singular values below cutoff_rel * max are
// dropped (include/fmm_pts.txx:498-503), the kept ones ordered largest first.
void pinv_factors(const std::vector<double>& G, const std::vector<double>& W, int r, int c, double cutoff_rel,
                  std::vector<double>& F0, std::vector<double>& F1, int& rank_out) {
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<double> s2(c);
  double smax = 0;
  for (int j = 0; j < c; j++) {
    double a = 0;
    for (int i = 0; i < r; i++) a += G[(size_t)j * r + i] * G[(size_t)j * r + i];
    s2[j] = a;
    smax = std::max(smax, std::sqrt(a));
  }
  if (cutoff_rel <= 0) cutoff_rel = 4 * eps;
  // keep singular values above the cutoff, largest first
  std::vector<int> keep;
  for (int j = 0; j < c; j++)
    if (std::sqrt(s2[j]) > cutoff_rel * smax) keep.push_back(j);
  std::sort(keep.begin(), keep.end(), [&](int a, int b) { return s2[a] > s2[b]; });
  const int rank = (int)keep.size();
  // F0 = V diag(1/s) (c x rank), F1 = U^T (rank x r) with U_j = G_j / s_j
  F0.assign((size_t)c * rank, 0.0);
  F1.assign((size_t)rank * r, 0.0);
  for (int k = 0; k < rank; k++) {
    const int j = keep[k];
    const double s = std::sqrt(s2[j]);
    for (int a = 0; a < c; a++) F0[(size_t)a * rank + k] = W[(size_t)j * c + a] / s;
    for (int b = 0; b < r; b++) F1[(size_t)k * r + b] = G[(size_t)j * r + b] / s;
  }
  rank_out = rank;
}

// ----------------------------------------------------------------------------- helpers
static void matmul(const double* A, const double* B, double* C, int n, int k, int m) {
  // C[n x m] = A[n x k] * B[k x m]
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double* Ci = &C[(size_t)i * m];
    std::fill(Ci, Ci + m, 0.0);
    for (int l = 0; l < k; l++) {
      double a = A[(size_t)i * k + l];
      if (a == 0) continue;
      const double* Bl = &B[(size_t)l * m];
      for (int j = 0; j < m; j++) Ci[j] += a * Bl[j];
    }
  }
}

// points of a surface of radius factor `rad` around a box of size `size` with centre `c`
static std::vector<double> surf_points(const Operators& O, const double* c, double size, double rad) {
  std::vector<double> p((size_t)O.ns * 3);
  for (int i = 0; i < O.ns; i++)
    for (int d = 0; d < 3; d++) p[3 * i + d] = c[d] + rad * size * O.unit_surf[3 * i + d];
  return p;
}

// Separable forward real DFT on an n^3 grid (x fastest); x axis halved; output [z][y][kx].
static void dft3_r2c(int n, const std::vector<double>& in, std::vector<std::complex<double>>& out) {
  const int nh = n / 2 + 1;
  const double pi = 3.14159265358979323846;
  std::vector<std::complex<double>> w(n);
  for (int k = 0; k < n; k++) w[k] = std::polar(1.0, -2 * pi * k / n);
  std::vector<std::complex<double>> a((size_t)n * n * nh), b((size_t)n * n * nh);
  for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
      for (int k = 0; k < nh; k++) {
        std::complex<double> s = 0;
        for (int x = 0; x < n; x++) s += in[((size_t)z * n + y) * n + x] * w[(k * x) % n];
        a[((size_t)z * n + y) * nh + k] = s;
      }
  for (int z = 0; z < n; z++)
    for (int ky = 0; ky < n; ky++)
      for (int k = 0; k < nh; k++) {
        std::complex<double> s = 0;
        for (int y = 0; y < n; y++) s += a[((size_t)z * n + y) * nh + k] * w[(ky * y) % n];
        b[((size_t)z * n + ky) * nh + k] = s;
      }
  out.assign((size_t)n * n * nh, 0);
  for (int kz = 0; kz < n; kz++)
    for (int ky = 0; ky < n; ky++)
      for (int k = 0; k < nh; k++) {
        std::complex<double> s = 0;
        for (int z = 0; z < n; z++) s += b[((size_t)z * n + ky) * nh + k] * w[(kz * z) % n];
        out[((size_t)kz * n + ky) * nh + k] = s;
      }
}

// ----------------------------------------------------------------------------- device precompute
namespace {

struct DevArray {
  void* p = nullptr;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() { if (p) (void)hipFree(p); }
  template <typename T>
  T* alloc(size_t n) {
    HIP_CHECK(hipMalloc(&p, sizeof(T) * std::max<size_t>(1, n)));
    return (T*)p;
  }
  template <typename T>
  T* upload(const std::vector<T>& v) {
    T* d = alloc<T>(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return d;
  }
};

struct DevStream {
  hipStream_t s = nullptr;
  DevStream() { HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~DevStream() { if (s) (void)hipStreamDestroy(s); }
};

// pinv of the r x c device matrix A by Jacobi sweeps on the device; factors as pinv_jacobi
void pinv_device(hipStream_t s, const double* A, int r, int c, double cutoff_rel, std::vector<double>& F0,
                 std::vector<double>& F1, int& rank) {
  const int cp = c + (c & 1);
  DevArray g, w, f;
  double* G = g.alloc<double>((size_t)cp * r);
  double* W = w.alloc<double>((size_t)cp * c);
  int* flag = f.alloc<int>(1);
  jacobi_init_device(s, A, r, c, cp, G, W);
  jacobi_sweeps_device(s, G, W, r, c, cp, flag, 60);
  std::vector<double> hG((size_t)cp * r), hW((size_t)cp * c);
  HIP_CHECK(hipMemcpyAsync(hG.data(), G, sizeof(double) * hG.size(), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(hW.data(), W, sizeof(double) * hW.size(), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  pinv_factors(hG, hW, r, c, cutoff_rel, F0, F1, rank);
}

// C = A (M x K) * B (K x N) on the device (FP64 matrix cores), C rows with stride ldc
void gemm_device(hipStream_t s, int M, int N, int K, const double* A, const double* B, double* C, int64_t ldc,
                 const int32_t* iota) {
  DeviceKernels<double>::gemm_rows(s, M, N, K, A, iota, K, B, C, iota, ldc, nullptr, false);
}

}  // namespace

bool device_precompute_enabled() {
  const char* e = std::getenv("PVFMM_PRECOMP");
  if (e && std::string(e) == "host") return false;
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

// ----------------------------------------------------------------------------- stages
// Parent surfaces of a unit box centred at 0.5 and the child surfaces of its 8 children.
static std::vector<double> unit_box_surface(const Operators& O, int which) {
  const double c0[3] = {0.5, 0.5, 0.5};
  switch (which) {
    case 0: return surf_points(O, c0, 1.0, kRad0);  // upward equivalent
    case 1: return surf_points(O, c0, 1.0, kRad1);  // upward check
    case 2: return surf_points(O, c0, 1.0, kRad1);  // downward equivalent
    case 3: return surf_points(O, c0, 1.0, kRad0);  // downward check
    default: break;
  }
  const int c = (which - 4) % 8;  // 4 + c: child upward equivalent, 12 + c: child downward check
  const double cc[3] = {0.25 + 0.5 * (c & 1), 0.25 + 0.5 * ((c >> 1) & 1), 0.25 + 0.5 * ((c >> 2) & 1)};
  return surf_points(O, cc, 0.5, kRad0);
}

// Surface lattice, unit surface offsets and spectrum sizes (no arithmetic of the operators).
static void operators_geometry(Operators& O, int kernel_id, int m, double cutoff_rel, int precision) {
  if (m < 2 || (m & 1)) throw Error(kErrInvalidArgument, "multipole order must be even and >= 2");
  O = Operators();
  O.m = m;
  O.ns = n_surf(m);
  O.ker = kernel_info(kernel_id);
  // 4*eps*max_S with eps of the evaluation precision (include/fmm_pts.txx:498-503; the
  // reference's float build applies the same rule with the float epsilon)
  if (cutoff_rel <= 0) cutoff_rel = 4.0 * (precision == 0 ? std::numeric_limits<double>::epsilon()
                                                          : (double)std::numeric_limits<float>::epsilon());
  O.pinv_rel_cutoff = cutoff_rel;
  O.n1 = 2 * m;
  O.nh = m + 1;
  O.M_dim = O.n1 * O.n1 * O.nh;
  // Surface lattice in the reference's point order (surface(), include/fmm_pts.txx:54-97):
  // corner point, three face blocks, then the point reflection of the first half.
  O.lattice.clear();
  auto push = [&](int a, int b, int c) { O.lattice.push_back(a); O.lattice.push_back(b); O.lattice.push_back(c); };
  push(0, 0, 0);
  for (int i = 0; i < m - 1; i++) for (int j = 0; j < m - 1; j++) push(0, i + 1, j);
  for (int i = 0; i < m - 1; i++) for (int j = 0; j < m - 1; j++) push(i, 0, j + 1);
  for (int i = 0; i < m - 1; i++) for (int j = 0; j < m - 1; j++) push(i + 1, j, 0);
  const int half = (int)O.lattice.size() / 3;
  for (int i = 0; i < half; i++) push(m - 1 - O.lattice[3 * i], m - 1 - O.lattice[3 * i + 1], m - 1 - O.lattice[3 * i + 2]);
  O.unit_surf.resize((size_t)O.ns * 3);
  for (int i = 0; i < 3 * O.ns; i++) O.unit_surf[i] = 0.5 * (2.0 * O.lattice[i] - (m - 1)) / (m - 1);
  O.v1_offsets.clear();
  for (int R = 0; R < 27; R++)
    if (R != 13) O.v1_offsets.push_back(R);
}

// Check -> equivalent pseudo-inverse factors at box size 1 (include/fmm_pts.txx:478-581).
static void compute_factors(Operators& O) {
  const int ns = O.ns, e0 = O.ker.e0, e1 = O.ker.e1, kind = O.ker.exp_kind, K0 = ns * e0, K1 = ns * e1;
  const std::vector<double> ue = unit_box_surface(O, 0), uc = unit_box_surface(O, 1);
  const std::vector<double> de = unit_box_surface(O, 2), dc = unit_box_surface(O, 3);
  if (O.on_device) {
    DevStream st;
    DevArray a, b, dm;
    double* M = dm.alloc<double>((size_t)K0 * K1);
    build_matrix_device(st.s, kind, a.upload(ue), ns, b.upload(uc), ns, M);
    pinv_device(st.s, M, K0, K1, O.pinv_rel_cutoff, O.uc2ue0, O.uc2ue1, O.rank_u);
    DevArray c, d;
    build_matrix_device(st.s, kind, c.upload(de), ns, d.upload(dc), ns, M);
    pinv_device(st.s, M, K0, K1, O.pinv_rel_cutoff, O.dc2de0, O.dc2de1, O.rank_d);
    return;
  }
  std::vector<double> M((size_t)K0 * K1);
  kernel_matrix(kind, ue.data(), ns, uc.data(), ns, M.data());
  pinv_jacobi(M, K0, K1, O.pinv_rel_cutoff, O.uc2ue0, O.uc2ue1, O.rank_u);
  kernel_matrix(kind, de.data(), ns, dc.data(), ns, M.data());
  pinv_jacobi(M, K0, K1, O.pinv_rel_cutoff, O.dc2de0, O.dc2de1, O.rank_d);
}

// U2U = (K_ce->pc V S^+) U^T and D2D = V S^+ (U^T K_pe->cc), the association order of
// include/fmm_pts.txx:582-643; parent of size 1, child of size 1/2.
static void compute_translations(Operators& O) {
  const int ns = O.ns, e0 = O.ker.e0, e1 = O.ker.e1, kind = O.ker.exp_kind, K0 = ns * e0, K1 = ns * e1;
  const int ru = O.rank_u, rd = O.rank_d;
  O.u2u.assign((size_t)8 * K0 * K0, 0.0);
  O.d2d.assign((size_t)K1 * 8 * K1, 0.0);
  const std::vector<double> uc = unit_box_surface(O, 1), de = unit_box_surface(O, 2);
  if (O.on_device) {
    DevStream st;
    const hipStream_t s = st.s;
    std::vector<int32_t> iota(std::max(K0, K1));
    for (size_t i = 0; i < iota.size(); i++) iota[i] = (int32_t)i;
    DevArray di, dm, dt, du0, du1, dd0, dd1, du2u, dd2d, duc, dde;
    const int32_t* I = di.upload(iota);
    const double* UC = duc.upload(uc);
    const double* DE = dde.upload(de);
    double* M = dm.alloc<double>((size_t)K0 * K1);
    const double* U0 = du0.upload(O.uc2ue0);  // K1 x ru
    const double* U1 = du1.upload(O.uc2ue1);  // ru x K0
    const double* D0 = dd0.upload(O.dc2de0);  // K1 x rd
    const double* D1 = dd1.upload(O.dc2de1);  // rd x K0
    double* T = dt.alloc<double>((size_t)std::max(K0, K1) * std::max(std::max(ru, rd), K1));
    double* u2u = du2u.alloc<double>((size_t)8 * K0 * K0);
    double* d2d = dd2d.alloc<double>((size_t)K1 * 8 * K1);
    for (int c = 0; c < 8; c++) {
      DevArray dce, dcd;
      build_matrix_device(s, kind, dce.upload(unit_box_surface(O, 4 + c)), ns, UC, ns, M);
      gemm_device(s, K0, ru, K1, M, U0, T, ru, I);
      gemm_device(s, K0, K0, ru, T, U1, u2u + (size_t)c * K0 * K0, K0, I);
      build_matrix_device(s, kind, DE, ns, dcd.upload(unit_box_surface(O, 12 + c)), ns, M);
      gemm_device(s, rd, K1, K0, D1, M, T, K1, I);
      gemm_device(s, K1, K1, rd, D0, T, d2d + (size_t)c * K1, 8 * K1, I);
      HIP_CHECK(hipStreamSynchronize(s));  // the child surfaces are freed at the end of the iteration
    }
    HIP_CHECK(hipMemcpyAsync(O.u2u.data(), u2u, sizeof(double) * O.u2u.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(O.d2d.data(), d2d, sizeof(double) * O.d2d.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  std::vector<double> M((size_t)K0 * K1), T((size_t)K0 * K1), Tu((size_t)K0 * std::max(ru, rd));
  for (int c = 0; c < 8; c++) {
    const std::vector<double> ce = unit_box_surface(O, 4 + c);  // child upward equivalent
    kernel_matrix(kind, ce.data(), ns, uc.data(), ns, M.data());
    matmul(M.data(), O.uc2ue0.data(), Tu.data(), K0, K1, ru);
    matmul(Tu.data(), O.uc2ue1.data(), &O.u2u[(size_t)c * K0 * K0], K0, ru, K0);
    const std::vector<double> cd = unit_box_surface(O, 12 + c);  // child downward check
    kernel_matrix(kind, de.data(), ns, cd.data(), ns, M.data());
    matmul(O.dc2de1.data(), M.data(), Tu.data(), rd, K0, K1);
    matmul(O.dc2de0.data(), Tu.data(), T.data(), K1, rd, K1);
    for (int i = 0; i < K1; i++)
      std::memcpy(&O.d2d[(size_t)i * 8 * K1 + (size_t)c * K1], &T[(size_t)i * K1], sizeof(double) * K1);
  }
}

// M2L transfer spectra (include/fmm_pts.txx:671-764).  Child boxes of size 1: equivalent and
// check surfaces share the lattice spacing h = RAD0/(m-1).  For source-minus-target child offset
// r, the kernel of lattice displacement w = v - u is K(w h - r), stored at w mod n1 (circular
// convolution).
static void compute_v1(Operators& O) {
  const int m = O.m, e0 = O.ker.e0, e1 = O.ker.e1, kind = O.ker.exp_kind;
  const int n1 = O.n1, M_dim = O.M_dim, ne = e0 * e1;
  const double h = kRad0 / (m - 1);
  std::vector<std::vector<std::complex<double>>> spec(343);  // offsets r in [-3,3]^3
  #pragma omp parallel for schedule(dynamic, 1)
  for (int ri = 0; ri < 343; ri++) {
    int rx = ri % 7 - 3, ry = (ri / 7) % 7 - 3, rz = ri / 49 - 3;
    if (std::max(std::abs(rx), std::max(std::abs(ry), std::abs(rz))) < 2) continue;
    std::vector<std::vector<double>> grid(ne, std::vector<double>((size_t)n1 * n1 * n1, 0.0));
    double blk[9];
    for (int wz = -(m - 1); wz <= m - 1; wz++)
      for (int wy = -(m - 1); wy <= m - 1; wy++)
        for (int wx = -(m - 1); wx <= m - 1; wx++) {
          double r[3] = {wx * h - rx, wy * h - ry, wz * h - rz};
          eval_pair(kind, r, blk);
          size_t g = ((size_t)((wz + n1) % n1) * n1 + (wy + n1) % n1) * n1 + (wx + n1) % n1;
          for (int q = 0; q < ne; q++) grid[q][g] = blk[q];
        }
    std::vector<std::complex<double>> sp, all((size_t)ne * M_dim);
    for (int q = 0; q < ne; q++) {
      dft3_r2c(n1, grid[q], sp);
      std::copy(sp.begin(), sp.end(), all.begin() + (size_t)q * M_dim);
    }
    spec[ri] = std::move(all);
  }
  const double inv_n3 = 1.0 / ((double)n1 * n1 * n1);
  const int bw = 8 * e1, bh = 8 * e0;  // block: rows t*e1+j, cols s*e0+i
  O.v1.assign((size_t)27 * M_dim * bw * bh, 0.0);
  for (int R = 0; R < 27; R++) {
    if (R == 13) continue;
    int Rx = R % 3 - 1, Ry = (R / 3) % 3 - 1, Rz = R / 9 - 1;
    for (int t = 0; t < 8; t++)
      for (int s = 0; s < 8; s++) {
        int rx = 2 * Rx + (s & 1) - (t & 1), ry = 2 * Ry + ((s >> 1) & 1) - ((t >> 1) & 1), rz = 2 * Rz + ((s >> 2) & 1) - ((t >> 2) & 1);
        if (std::max(std::abs(rx), std::max(std::abs(ry), std::abs(rz))) < 2) continue;
        const auto& sp = spec[(rz + 3) * 49 + (ry + 3) * 7 + (rx + 3)];
        for (int i = 0; i < e0; i++)
          for (int j = 0; j < e1; j++)
            for (int f = 0; f < M_dim; f++)
              O.v1[(((size_t)R * M_dim + f) * bw + (t * e1 + j)) * bh + (s * e0 + i)] = sp[(size_t)(i * e1 + j) * M_dim + f] * inv_n3;
      }
  }
}

// ----------------------------------------------------------------------------- build
void build_operators(Operators& O, int kernel_id, int m, double cutoff_rel, int precision) {
  operators_geometry(O, kernel_id, m, cutoff_rel, precision);
  O.on_device = device_precompute_enabled();
  compute_factors(O);
  compute_translations(O);
  compute_v1(O);
}

// ----------------------------------------------------------------------------- periodic boundary
namespace {

// Gauss-Legendre rule on [0, 1] (Newton iteration on P_n from the Chebyshev-like first guess)
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  const double pi = 3.14159265358979323846;
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; i++) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1;
    for (int it = 0; it < 100; it++) {
      double p1 = 1, p0 = 0;
      for (int j = 1; j <= n; j++) {
        const double pm = p0;
        p0 = p1;
        p1 = ((2 * j - 1) * z * p0 - (j - 1) * pm) / j;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    x[i] = 0.5 * (1 - z);
    w[i] = 1.0 / ((1 - z * z) * dp * dp);
  }
}

// One Duffy pyramid of the signed box [0, E]^3 (E relative to the point, singular corner at the
// point): for a kernel homogeneous of degree -1 the radial integral is 1/2 and what is left is
// the smooth integral over (s, t) in [0, 1]^2 of K at E[ax] along axis ax and E[p] s, E[q] t
// along the others, graded geometrically toward s = t = 0 when E[ax] is small.
void duffy_pyramid(int kind, int ax, const double* E, int ne, const std::vector<double>& gx, const std::vector<double>& gw,
                   double* acc) {
  const int p = (ax + 1) % 3, q = (ax + 2) % 3;
  const double lead = std::fabs(E[ax]), rho = std::max(std::fabs(E[p]), std::fabs(E[q])) / lead;
  const int L = rho > 1 ? std::min(30, (int)std::ceil(std::log(rho) / std::log(5.0))) : 0;
  std::vector<double> sx, sw;  // graded nodes on [0, 1]
  for (int l = L; l >= 0; l--) {
    const double a = l == L ? 0.0 : std::pow(5.0, -l - 1), b = std::pow(5.0, -l);
    for (size_t i = 0; i < gx.size(); i++) {
      sx.push_back(a + (b - a) * gx[i]);
      sw.push_back((b - a) * gw[i]);
    }
  }
  double blk[9], r[3];
  r[ax] = E[ax];
  for (size_t i = 0; i < sx.size(); i++)
    for (size_t j = 0; j < sx.size(); j++) {
      r[p] = E[p] * sx[i];
      r[q] = E[q] * sx[j];
      eval_pair(kind, r, blk);
      const double wt = sw[i] * sw[j];
      for (int c = 0; c < ne; c++) acc[c] += wt * blk[c];
    }
}

// row-major 8x8 inverse by Gauss-Jordan with partial pivoting
void invert8(const double* A, double* Ainv) {
  double a[8][16];
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 16; j++) a[i][j] = j < 8 ? A[i * 8 + j] : (j - 8 == i);
  for (int c = 0; c < 8; c++) {
    int piv = c;
    for (int i = c + 1; i < 8; i++)
      if (std::fabs(a[i][c]) > std::fabs(a[piv][c])) piv = i;
    for (int j = 0; j < 16; j++) std::swap(a[c][j], a[piv][j]);
    const double d = a[c][c];
    for (int j = 0; j < 16; j++) a[c][j] /= d;
    for (int i = 0; i < 8; i++)
      if (i != c) {
        const double f = a[i][c];
        for (int j = 0; j < 16; j++) a[i][j] -= f * a[c][j];
      }
  }
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) Ainv[i * 8 + j] = a[i][8 + j];
}

// the analytic potential of a uniform unit density (laplace_vol_poten, stokes_vol_poten,
// include/kernel.txx:1289-1295,1456-1467): e0 x (n*e1)
std::vector<double> vol_poten(int kind, const double* x, int n) {
  const int e0 = kind == 2 ? 3 : 1, e1 = e0;
  std::vector<double> out((size_t)e0 * n * e1, 0.0);
  for (int i = 0; i < n; i++) {
    const double r2 = x[3 * i] * x[3 * i] + x[3 * i + 1] * x[3 * i + 1] + x[3 * i + 2] * x[3 * i + 2];
    if (kind == 0) out[i] = -r2 / 6;
    else
      for (int k = 0; k < 3; k++) out[(size_t)k * n * 3 + (size_t)i * 3 + k] = -(r2 - x[3 * i + k] * x[3 * i + k]) / 4;
  }
  return out;
}

// C = A (n x k) B (k x m), on the FP64 matrix cores when dev
std::vector<double> mul(const std::vector<double>& A, const std::vector<double>& B, int n, int k, int m, bool dev) {
  std::vector<double> C((size_t)n * m);
  if (!dev) {
    matmul(A.data(), B.data(), C.data(), n, k, m);
    return C;
  }
  DevStream st;
  DevArray a, b, c, io;
  std::vector<int32_t> iota(n);
  for (int i = 0; i < n; i++) iota[i] = i;
  double* dc = c.alloc<double>((size_t)n * m);
  gemm_device(st.s, n, m, k, a.upload(A), b.upload(B), dc, m, io.upload(iota));
  HIP_CHECK(hipMemcpyAsync(C.data(), dc, sizeof(double) * C.size(), hipMemcpyDeviceToHost, st.s));
  HIP_CHECK(hipStreamSynchronize(st.s));
  return C;
}

}  // namespace

void cube_integral(int kind, const double* pts, int n, double lo, double size, double* out) {
  const int e0 = kind == 2 ? 3 : 1, e1 = kind == 0 ? 1 : 3, ne = e0 * e1;
  if (kind == 1) throw Error(kErrInvalidArgument, "cube_integral: expansion kernels only (laplace, stokes)");
  std::vector<double> gx, gw;
  gauss_legendre01(20, gx, gw);
  #pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n; i++) {
    // inclusion-exclusion over the 8 corners: signed boxes from the point to each corner
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 8; c++) {
      double E[3], sign = 1;
      for (int d = 0; d < 3; d++) {
        const bool hi = (c >> d) & 1;
        E[d] = (hi ? lo + size : lo) - pts[3 * i + d];
        if (!hi) sign = -sign;
      }
      const double V = E[0] * E[1] * E[2];
      if (V == 0) continue;
      double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int ax = 0; ax < 3; ax++) duffy_pyramid(kind, ax, E, ne, gx, gw, J);
      for (int q = 0; q < ne; q++) acc[q] += sign * 0.5 * V * J[q];
    }
    for (int q = 0; q < ne; q++) out[(size_t)i * ne + q] = acc[q];
  }
}

// The boundary operators of every type in `types` (bits 1 << bndry), sharing the order m + 2
// factors and the products common to all types.
static void build_bc_operators(Operators& O, int types) {
  for (int b = 1; b <= 3; b++)
    if (!O.bc[b].empty()) types &= ~(1 << b);
  if (!types) return;
  const int m = O.m, mo = m + 2, kind = O.ker.exp_kind, e0 = O.ker.e0, e1 = O.ker.e1;
  const bool dev = O.on_device;
  Operators G;  // order m + 2: surfaces and check -> equivalent factors of the unit box
  operators_geometry(G, O.ker.id, mo, O.pinv_rel_cutoff, 0);
  G.on_device = dev;
  compute_factors(G);
  const int n = G.ns, K0 = n * e0, K1 = n * e1, ru = G.rank_u, rd = G.rank_d;
  // the factors of a box of size 2 are those of the unit box times 2^(src_scal + trg_scal)
  const double s2 = std::pow(2.0, O.ker.src_scal + O.ker.trg_scal);
  auto surf = [](const Operators& Q, double cx, double cy, double cz, int depth, double rad) {  // corner-anchored
    const double size = std::ldexp(1.0, -depth), c[3] = {cx + 0.5 * size, cy + 0.5 * size, cz + 0.5 * size};
    return surf_points(Q, c, size, rad);
  };
  auto K = [&](const std::vector<double>& a, const std::vector<double>& b) {
    const int na = (int)a.size() / 3, nb = (int)b.size() / 3;
    std::vector<double> M((size_t)na * e0 * nb * e1);
    kernel_matrix(kind, a.data(), na, b.data(), nb, M.data());
    return M;
  };
  auto axpy = [](std::vector<double>& y, const std::vector<double>& x, double a) {
    for (size_t i = 0; i < y.size(); i++) y[i] += a * x[i];
  };

  // zero-average projections (include/fmm_pts.txx:862-913): Z removes the total equivalent
  // charge spread as the equivalent density of a uniformly charged unit cube, C the mean check
  // potential of each component
  std::vector<double> S2E;  // e0 x K0, rows summing to 1
  {
    const std::vector<double> uc = surf(G, 0, 0, 0, 0, kRad1);
    std::vector<double> gi((size_t)n * e0 * e1), s2c((size_t)e0 * K1);
    cube_integral(kind, uc.data(), n, 0.0, 1.0, gi.data());
    for (int j = 0; j < e0; j++)
      for (int i = 0; i < n; i++)
        for (int k = 0; k < e1; k++) s2c[(size_t)j * K1 + (size_t)i * e1 + k] = gi[((size_t)i * e0 + k) * e1 + j];
    S2E = mul(mul(s2c, G.uc2ue0, e0, K1, ru, false), G.uc2ue1, e0, ru, K0, false);
    for (int j = 0; j < e0; j++) {
      double t = 0;
      for (int i = 0; i < K0; i++) t += S2E[(size_t)j * K0 + i];
      for (int i = 0; i < K0; i++) S2E[(size_t)j * K0 + i] /= t;
    }
  }
  auto z_left = [&](std::vector<double>& X, int cols) {  // X <- Z X, X: K0 x cols
    std::vector<double> R = mul(S2E, X, e0, K0, cols, false);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < e0; k++) {
        double* x = &X[((size_t)i * e0 + k) * cols];
        for (int c = 0; c < cols; c++) x[c] -= R[(size_t)k * cols + c];
      }
  };
  auto z_right = [&](std::vector<double>& X, int rows) {  // X <- X Z, X: rows x K0
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
      double* x = &X[(size_t)r * K0];
      double t[3] = {0, 0, 0};
      for (int i = 0; i < n; i++)
        for (int k = 0; k < e0; k++) t[k] += x[i * e0 + k];
      for (int j = 0; j < K0; j++)
        for (int k = 0; k < e0; k++) x[j] -= t[k] * S2E[(size_t)k * K0 + j];
    }
  };
  auto c_left = [&](std::vector<double>& X, int cols) {  // X <- C X, X: K1 x cols
    for (int k = 0; k < e1; k++) {
      std::vector<double> mean(cols, 0.0);
      for (int i = 0; i < n; i++)
        for (int c = 0; c < cols; c++) mean[c] += X[((size_t)i * e1 + k) * cols + c];
      for (int i = 0; i < n; i++)
        for (int c = 0; c < cols; c++) X[((size_t)i * e1 + k) * cols + c] -= mean[c] / n;
    }
  };
  auto c_right = [&](std::vector<double>& X, int rows) {  // X <- X C, X: rows x K1
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
      double* x = &X[(size_t)r * K1];
      double t[3] = {0, 0, 0};
      for (int i = 0; i < n; i++)
        for (int k = 0; k < e1; k++) t[k] += x[i * e1 + k];
      for (int i = 0; i < n; i++)
        for (int k = 0; k < e1; k++) x[i * e1 + k] -= t[k] / n;
    }
  };

  // L2L: [-1, 1]^3 -> its child [-1, 0]^3 (include/fmm_pts.txx:942-951)
  std::vector<double> L2L;
  {
    const std::vector<double> MM = K(surf(G, -1, -1, -1, -1, kRad1), surf(G, -1, -1, -1, 0, kRad0));
    L2L = mul(G.dc2de0, mul(G.dc2de1, MM, rd, K0, K1, dev), K1, rd, K1, dev);
    for (double& v : L2L) v *= s2;
  }
  c_left(L2L, K1); c_right(L2L, K1);
  // order m + 2 -> m (include/fmm_pts.txx:1141-1160)
  const int K0m = O.K0(), K1m = O.K1();
  const std::vector<double> M2Md =
      mul(mul(K(surf(O, -0.5, -0.5, -0.5, 0, kRad0), surf(G, -0.5, -0.5, -0.5, 0, kRad1)), G.uc2ue0, K0m, K1, ru, dev),
          G.uc2ue1, K0m, ru, K0, dev);
  const std::vector<double> L2Ld =
      mul(G.dc2de0, mul(G.dc2de1, K(surf(G, -0.5, -0.5, -0.5, 0, kRad1), surf(O, -0.5, -0.5, -0.5, 0, kRad0)), rd, K0, K1m, dev),
          K1, rd, K1m, dev);

  for (int bndry = 1; bndry <= 3; bndry++) {
    if (!(types & (1 << bndry))) continue;
    auto keep = [&](int a, int b, int c) { return !((bndry == kPX && (b != 0 || c != 0)) || (bndry == kPXY && c != 0)); };
    const bool vol = bndry == kPXYZ && O.ker.id != 1;  // laplace_grad's expansion kernel has no volume potential
    // M2M: the 8 unit boxes of [-1, 1]^3 -> that box (include/fmm_pts.txx:917-941)
    std::vector<double> M2M;
    {
      const std::vector<double> chk = surf(G, -1, -1, -1, -1, kRad1);
      std::vector<double> S((size_t)K0 * K1, 0.0);
      for (int c = 0; c < 8; c++)
        if (keep(c & 1, (c >> 1) & 1, (c >> 2) & 1))
          axpy(S, K(surf(G, -1 + (c & 1), -1 + ((c >> 1) & 1), -1 + ((c >> 2) & 1), 0, kRad0), chk), 1.0);
      M2M = mul(mul(S, G.uc2ue0, K0, K1, ru, dev), G.uc2ue1, K0, ru, K0, dev);
      for (double& v : M2M) v *= s2;
    }
    // M2L: unit boxes at offsets in [-2, 3]^3 not adjacent to [0, 1]^3 (include/fmm_pts.txx:952-975)
    std::vector<double> M2L((size_t)K0 * K1, 0.0);
    {
      const std::vector<double> dnc = surf(G, 0, 0, 0, 0, kRad0);
      for (int c = -2; c < 4; c++)
        for (int b = -2; b < 4; b++)
          for (int a = -2; a < 4; a++)
            if (keep(a, b, c) && std::max(std::abs(a), std::max(std::abs(b), std::abs(c))) > 1)
              axpy(M2L, K(surf(G, a, b, c, 0, kRad0), dnc), 1.0);
    }
    z_left(M2M, K0); z_right(M2M, K0);
    z_left(M2L, K1); c_right(M2L, K0);

    // the levels: M <- M2L + M2M (2^-(src_scal + trg_scal) M) L2L, PVFMM_BC_LEVELS = 45 in all
    const double down = 1.0 / s2;
    for (double& v : M2M) v *= down;
    std::vector<double> M = M2L;
    const int levels = 45;
    if (dev) {
      DevStream st;
      DevArray a, b, c, dm, dt, io;
      std::vector<int32_t> iota(std::max(K0, K1));
      for (size_t i = 0; i < iota.size(); i++) iota[i] = (int32_t)i;
      const int32_t* I = io.upload(iota);
      const double* d_m2m = a.upload(M2M);
      const double* d_l2l = b.upload(L2L);
      const double* d_m2l = c.upload(M2L);
      double* d_M = dm.upload(M);
      double* d_T = dt.alloc<double>((size_t)K0 * K1);
      for (int l = 1; l < levels; l++) {
        gemm_device(st.s, K0, K1, K0, d_m2m, d_M, d_T, K1, I);
        HIP_CHECK(hipMemcpyAsync(d_M, d_m2l, sizeof(double) * M.size(), hipMemcpyDeviceToDevice, st.s));
        DeviceKernels<double>::gemm_rows(st.s, K0, K1, K1, d_T, I, K1, d_l2l, d_M, I, K1, nullptr, true);
      }
      HIP_CHECK(hipMemcpyAsync(M.data(), d_M, sizeof(double) * M.size(), hipMemcpyDeviceToHost, st.s));
      HIP_CHECK(hipStreamSynchronize(st.s));
    } else {
      for (int l = 1; l < levels; l++) {
        std::vector<double> T = mul(mul(M2M, M, K0, K0, K1, false), L2L, K0, K1, K1, false);
        for (size_t i = 0; i < M.size(); i++) M[i] = M2L[i] + T[i];
      }
    }

    // far field of the analytic volume potential (include/fmm_pts.txx:1001-1033): the potential of
    // a uniform density everywhere minus that of the 27 boxes around [1, 2]^3
    if (vol) {
      const std::vector<double> dc = surf(G, 1, 1, 1, 0, kRad0);
      std::vector<double> gi((size_t)n * e0 * e1);
      cube_integral(kind, dc.data(), n, 0.0, 3.0, gi.data());
      std::vector<double> far = vol_poten(kind, dc.data(), n);  // e0 x K1
      for (int j = 0; j < e0; j++)
        for (int i = 0; i < n; i++)
          for (int k = 0; k < e1; k++) far[(size_t)j * K1 + (size_t)i * e1 + k] -= gi[((size_t)i * e0 + k) * e1 + j];
      for (int i = 0; i < n; i++)
        for (int k = 0; k < e0; k++)
          for (int c = 0; c < K1; c++) M[((size_t)i * e0 + k) * K1 + c] += far[(size_t)k * K1 + c];
    }

    // corner correction (include/fmm_pts.txx:1036-1138): fit a + bx + cy + dz + exy + fxz + gyz +
    // hxyz to the potential at the 8 corners of [0, 1]^3 (local expansion plus the root's images
    // next to each corner), and subtract the enabled terms on the check surface
    {
      const double corners[24] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1};
      const std::vector<double> cpts(corners, corners + 24);
      const std::vector<double> ue = surf(G, 0, 0, 0, 0, kRad0), de = surf(G, 0, 0, 0, 0, kRad1);
      const std::vector<double> dnc = surf(G, 0, 0, 0, 0, kRad0);
      std::vector<double> cv = mul(mul(M, G.dc2de0, K0, K1, rd, dev), mul(G.dc2de1, K(de, cpts), rd, K0, 8 * e1, dev), K0, rd,
                                   8 * e1, dev);  // K0 x 8 e1
      for (int kk = 0; kk < 8; kk++)
        for (int j0 = -1; j0 <= 1; j0++)
          for (int j1 = -1; j1 <= 1; j1++)
            for (int j2 = -1; j2 <= 1; j2++) {
              if ((bndry == kPXY && j2 != 0) || (bndry == kPX && (j1 != 0 || j2 != 0))) continue;
              const double pt[3] = {corners[3 * kk] - j0, corners[3 * kk + 1] - j1, corners[3 * kk + 2] - j2};
              if (std::fabs(pt[0] - 0.5) < 1 && std::fabs(pt[1] - 0.5) < 1 && std::fabs(pt[2] - 0.5) < 1) continue;
              const std::vector<double> E = K(ue, std::vector<double>(pt, pt + 3));  // K0 x e1
              for (int r = 0; r < K0; r++)
                for (int j = 0; j < e1; j++) cv[(size_t)r * 8 * e1 + kk * e1 + j] += E[(size_t)r * e1 + j];
            }
      if (vol) {
        const std::vector<double> va = vol_poten(kind, corners, 8);  // e0 x 8 e1
        for (int i = 0; i < n; i++)
          for (int k = 0; k < e0; k++)
            for (int c = 0; c < 8 * e1; c++) cv[((size_t)i * e0 + k) * 8 * e1 + c] -= va[(size_t)k * 8 * e1 + c];
      }
      auto basis = [](const double* p, double* b) {
        b[0] = 1; b[1] = p[0]; b[2] = p[1]; b[3] = p[2];
        b[4] = p[0] * p[1]; b[5] = p[0] * p[2]; b[6] = p[1] * p[2]; b[7] = p[0] * p[1] * p[2];
      };
      double V[64], Vinv[64];  // V[coeff][corner]
      for (int kk = 0; kk < 8; kk++) {
        double b[8];
        basis(&corners[3 * kk], b);
        for (int c = 0; c < 8; c++) V[c * 8 + kk] = b[c];
      }
      invert8(V, Vinv);  // Vinv[corner][coeff]
      const int enable = bndry == kPXYZ ? 4 : bndry == kPXY ? 3 : 2;  // terms 1, x, y, z in use
      std::vector<double> cvec((size_t)8 * n, 0.0);  // cvec[coeff][s]
      for (int s = 0; s < n; s++) {
        double b[8];
        basis(&dnc[3 * s], b);
        for (int c = 0; c < enable; c++) cvec[(size_t)c * n + s] = b[c];
      }
      #pragma omp parallel for schedule(static)
      for (int i = 0; i < K0; i++) {
        double coeff[8][3] = {};
        for (int c = 0; c < 8; c++)
          for (int kk = 0; kk < 8; kk++)
            for (int j = 0; j < e1; j++) coeff[c][j] += cv[(size_t)i * 8 * e1 + kk * e1 + j] * Vinv[kk * 8 + c];
        for (int s = 0; s < n; s++)
          for (int j = 0; j < e1; j++) {
            double t = 0;
            for (int c = 0; c < enable; c++) t += coeff[c][j] * cvec[(size_t)c * n + s];
            M[(size_t)i * K1 + (size_t)s * e1 + j] -= t;
          }
      }
    }

    O.bc[bndry] = mul(mul(M2Md, M, K0m, K0, K1, dev), L2Ld, K0m, K1, K1m, dev);
  }
}

void build_bc_operator(Operators& O, int bndry) {
  if (bndry < 1 || bndry > 3) throw Error(kErrInvalidArgument, "periodic boundary type must be 1 (PX), 2 (PXY) or 3 (PXYZ)");
  build_bc_operators(O, 1 << bndry);
}

// ----------------------------------------------------------------------------- cache file
// The container of PrecompMat::Save2File / LoadFile (include/precomp_mat.txx:212-347): int
// sizeof(T), int scale_invar, then for each of the 16 Mat_Type values (precomp_mat.hpp:21-39)
// int n and n matrices, each int rows, int cols and rows*cols values of T.
namespace {
enum : int { kUC2UE0 = 0, kUC2UE1 = 1, kDC2DE0 = 2, kDC2DE1 = 3, kU2UT = 5, kD2DT = 6, kV1T = 14, kBCT = 15, kMatTypes = 16 };
struct Mat {
  int rows = 0, cols = 0;
  std::vector<double> v;
};
}  // namespace

static bool read_container(const std::string& path, std::vector<std::vector<Mat>>& mats) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
  int32_t sz = 0, scale_invar = 0;
  if (std::fread(&sz, 4, 1, f) != 1 || std::fread(&scale_invar, 4, 1, f) != 1) return false;
  if (sz != (int32_t)sizeof(double) || scale_invar != 1) return false;  // FP64 data of a scale-invariant kernel
  mats.assign(kMatTypes, {});
  for (int t = 0; t < kMatTypes; t++) {
    int32_t n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n < 0 || n > 4096) return false;
    mats[t].resize(n);
    for (int j = 0; j < n; j++) {
      int32_t r = 0, c = 0;
      if (std::fread(&r, 4, 1, f) != 1 || std::fread(&c, 4, 1, f) != 1 || r < 0 || c < 0) return false;
      Mat& M = mats[t][j];
      M.rows = r;
      M.cols = c;
      M.v.resize((size_t)r * c);
      if ((size_t)r * c && std::fread(M.v.data(), sizeof(double), M.v.size(), f) != M.v.size()) return false;
    }
  }
  return true;
}

static void write_container(const std::string& path, const std::vector<std::vector<Mat>>& mats) {
  const std::string tmp = path + ".tmp" + std::to_string((long long)std::time(nullptr));
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;  // an unwritable directory only costs the next run its precompute
  bool ok = true;
  auto put = [&](const void* p, size_t sz, size_t n) { ok = ok && std::fwrite(p, sz, n, f) == n; };
  const int32_t sz = (int32_t)sizeof(double), scale_invar = 1;
  put(&sz, 4, 1);
  put(&scale_invar, 4, 1);
  for (int t = 0; t < kMatTypes; t++) {
    const int32_t n = t < (int)mats.size() ? (int32_t)mats[t].size() : 0;
    put(&n, 4, 1);
    for (int j = 0; j < n; j++) {
      const Mat& M = mats[t][j];
      const int32_t r = M.rows, c = M.cols;
      put(&r, 4, 1);
      put(&c, 4, 1);
      if (!M.v.empty()) put(M.v.data(), sizeof(double), M.v.size());
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  if (ok) ok = std::rename(tmp.c_str(), path.c_str()) == 0;
  if (!ok) std::remove(tmp.c_str());
}

std::string precomp_file_name(const Operators& O, int precision) {
  // Precomp_<ker_name>_m<m>.data as the reference names its FP64 files (include/fmm_pts.txx:
  // 244-262); the mixed evaluator's operators (FP64 values, FP32 cutoff) get their own suffix
  return std::string("Precomp_") + O.ker.name + "_m" + std::to_string(O.m) + (precision ? "_mixed" : "") + ".data";
}

void build_operators_cached(Operators& O, int kernel_id, int m, double cutoff_rel, int precision, const char* dir) {
  if (!dir || !*dir || cutoff_rel > 0) {  // custom cutoffs are not cached
    build_operators(O, kernel_id, m, cutoff_rel, precision);
    return;
  }
  operators_geometry(O, kernel_id, m, cutoff_rel, precision);
  O.on_device = device_precompute_enabled();
  const std::string path = std::string(dir) + "/" + precomp_file_name(O, precision);
  const int K0 = O.K0(), K1 = O.K1(), M_dim = O.M_dim, bw = 8 * O.ker.e1, bh = 8 * O.ker.e0;
  std::vector<std::vector<Mat>> mats;
  const bool exists = read_container(path, mats);
  // factors: F0 (K1 x r), F1 (r x K0) with the same r; translations: 8 + 8 full blocks;
  // spectra: 27 entries, 26 of M_dim x (2 bw bh)
  bool have_f = exists && mats[kUC2UE0].size() == 1 && mats[kUC2UE1].size() == 1 && mats[kDC2DE0].size() == 1 &&
                mats[kDC2DE1].size() == 1;
  if (have_f) {
    const Mat &a = mats[kUC2UE0][0], &b = mats[kUC2UE1][0], &c = mats[kDC2DE0][0], &d = mats[kDC2DE1][0];
    have_f = a.rows == K1 && b.cols == K0 && a.cols == b.rows && a.cols > 0 && c.rows == K1 && d.cols == K0 &&
             c.cols == d.rows && c.cols > 0;
  }
  bool have_t = have_f && mats[kU2UT].size() == 8 && mats[kD2DT].size() == 8;
  for (int c = 0; have_t && c < 8; c++)
    have_t = mats[kU2UT][c].rows == K0 && mats[kU2UT][c].cols == K0 && mats[kD2DT][c].rows == K1 && mats[kD2DT][c].cols == K1;
  bool have_v = exists && mats[kV1T].size() == 27;
  for (int R = 0; have_v && R < 27; R++)
    have_v = R == 13 ? mats[kV1T][R].v.empty() : (mats[kV1T][R].rows == M_dim && mats[kV1T][R].cols == 2 * bw * bh);
  if (have_f) {
    O.uc2ue0 = mats[kUC2UE0][0].v; O.uc2ue1 = mats[kUC2UE1][0].v; O.rank_u = mats[kUC2UE0][0].cols;
    O.dc2de0 = mats[kDC2DE0][0].v; O.dc2de1 = mats[kDC2DE1][0].v; O.rank_d = mats[kDC2DE0][0].cols;
  } else {
    compute_factors(O);
  }
  if (have_t) {
    O.u2u.resize((size_t)8 * K0 * K0);
    O.d2d.resize((size_t)K1 * 8 * K1);
    for (int c = 0; c < 8; c++) {
      std::copy(mats[kU2UT][c].v.begin(), mats[kU2UT][c].v.end(), O.u2u.begin() + (size_t)c * K0 * K0);
      for (int i = 0; i < K1; i++)
        std::copy(mats[kD2DT][c].v.begin() + (size_t)i * K1, mats[kD2DT][c].v.begin() + (size_t)(i + 1) * K1,
                  O.d2d.begin() + (size_t)i * 8 * K1 + (size_t)c * K1);
    }
  } else {
    compute_translations(O);
  }
  if (have_v) {
    O.v1.assign((size_t)27 * M_dim * bw * bh, 0.0);
    for (int R = 0; R < 27; R++)
      if (R != 13) std::memcpy(&O.v1[(size_t)R * M_dim * bw * bh], mats[kV1T][R].v.data(), sizeof(double) * mats[kV1T][R].v.size());
  } else {
    compute_v1(O);
  }
  // boundary operators, one per BoundaryType (FreeSpace stored as zeros, include/fmm_pts.txx:
  // 288-300): read where present; a new file gets all of them, as the reference computes them
  // in Initialize
  bool have_bc = exists && have_f && mats[kBCT].size() == 4;
  for (int b = 0; have_bc && b < 4; b++) have_bc = mats[kBCT][b].rows == K0 && mats[kBCT][b].cols == K1;
  if (have_bc)
    for (int b = 1; b <= 3; b++) O.bc[b] = mats[kBCT][b].v;
  O.from_cache = have_f && have_t && have_v;
  if (exists) return;  // an existing file is never replaced (PrecompMat::Save2File with replace = false)
  build_bc_operators(O, (1 << kPX) | (1 << kPXY) | (1 << kPXYZ));
  std::vector<Mat> bcm{Mat{K0, K1, std::vector<double>((size_t)K0 * K1, 0.0)}};
  for (int b = 1; b <= 3; b++) bcm.push_back(Mat{K0, K1, O.bc[b]});
  std::vector<std::vector<Mat>> out(kMatTypes);
  out[kUC2UE0] = {Mat{K1, O.rank_u, O.uc2ue0}};
  out[kUC2UE1] = {Mat{O.rank_u, K0, O.uc2ue1}};
  out[kDC2DE0] = {Mat{K1, O.rank_d, O.dc2de0}};
  out[kDC2DE1] = {Mat{O.rank_d, K0, O.dc2de1}};
  for (int c = 0; c < 8; c++) {
    out[kU2UT].push_back(Mat{K0, K0, std::vector<double>(O.u2u.begin() + (size_t)c * K0 * K0, O.u2u.begin() + (size_t)(c + 1) * K0 * K0)});
    Mat D{K1, K1, std::vector<double>((size_t)K1 * K1)};
    for (int i = 0; i < K1; i++)
      std::copy(O.d2d.begin() + (size_t)i * 8 * K1 + (size_t)c * K1, O.d2d.begin() + (size_t)i * 8 * K1 + (size_t)(c + 1) * K1,
                D.v.begin() + (size_t)i * K1);
    out[kD2DT].push_back(std::move(D));
  }
  for (int R = 0; R < 27; R++) {
    if (R == 13) { out[kV1T].push_back(Mat{}); continue; }
    const double* b = reinterpret_cast<const double*>(&O.v1[(size_t)R * M_dim * bw * bh]);
    out[kV1T].push_back(Mat{M_dim, 2 * bw * bh, std::vector<double>(b, b + (size_t)2 * M_dim * bw * bh)});
  }
  out[kBCT] = std::move(bcm);
  write_container(path, out);
}

}  // namespace pvfmm_amd
