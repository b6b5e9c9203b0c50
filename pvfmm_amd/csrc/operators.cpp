// FP64 operator precompute on the host.  Pseudo-inverses come from a one-sided (Hestenes)
// Jacobi SVD written here (high relative accuracy for the ill-conditioned equivalent-surface
// systems); M2L transfer spectra come from a separable real DFT written here.  Both are
// independent of the oracle's numpy/LAPACK path that checks them.
#include "operators.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include <omp.h>

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

// ----------------------------------------------------------------------------- build
void build_operators(Operators& O, int kernel_id, int m, double cutoff_rel, int precision) {
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
  const int ns = O.ns, e0 = O.ker.e0, e1 = O.ker.e1, kind = O.ker.exp_kind;

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
  O.unit_surf.resize((size_t)ns * 3);
  for (int i = 0; i < 3 * ns; i++) O.unit_surf[i] = 0.5 * (2.0 * O.lattice[i] - (m - 1)) / (m - 1);

  const double c0[3] = {0.5, 0.5, 0.5};
  // check -> equivalent pseudo-inverses at box size 1 (include/fmm_pts.txx:478-581)
  {
    std::vector<double> ue = surf_points(O, c0, 1.0, kRad0), uc = surf_points(O, c0, 1.0, kRad1);
    std::vector<double> M((size_t)ns * e0 * ns * e1);
    kernel_matrix(kind, ue.data(), ns, uc.data(), ns, M.data());
    pinv_jacobi(M, ns * e0, ns * e1, cutoff_rel, O.uc2ue0, O.uc2ue1, O.rank_u);
    std::vector<double> de = surf_points(O, c0, 1.0, kRad1), dc = surf_points(O, c0, 1.0, kRad0);
    kernel_matrix(kind, de.data(), ns, dc.data(), ns, M.data());
    pinv_jacobi(M, ns * e0, ns * e1, cutoff_rel, O.dc2de0, O.dc2de1, O.rank_d);
  }
  // U2U = (K_ce->pc V S^+) U^T and D2D = V S^+ (U^T K_pe->cc), the association order of
  // include/fmm_pts.txx:582-643; parent of size 1, child of size 1/2
  {
    const int K0 = ns * e0, K1 = ns * e1, ru = O.rank_u, rd = O.rank_d;
    O.u2u.assign((size_t)8 * K0 * K0, 0.0);
    O.d2d.assign((size_t)K1 * 8 * K1, 0.0);
    std::vector<double> uc = surf_points(O, c0, 1.0, kRad1);
    std::vector<double> de = surf_points(O, c0, 1.0, kRad1);
    std::vector<double> M((size_t)K0 * K1), T((size_t)K0 * K1), Tu((size_t)K0 * std::max(ru, rd));
    for (int c = 0; c < 8; c++) {
      double cc[3] = {0.25 + 0.5 * (c & 1), 0.25 + 0.5 * ((c >> 1) & 1), 0.25 + 0.5 * ((c >> 2) & 1)};
      std::vector<double> ce = surf_points(O, cc, 0.5, kRad0);  // child upward equivalent
      kernel_matrix(kind, ce.data(), ns, uc.data(), ns, M.data());
      matmul(M.data(), O.uc2ue0.data(), Tu.data(), K0, K1, ru);
      matmul(Tu.data(), O.uc2ue1.data(), &O.u2u[(size_t)c * K0 * K0], K0, ru, K0);
      std::vector<double> cd = surf_points(O, cc, 0.5, kRad0);  // child downward check
      kernel_matrix(kind, de.data(), ns, cd.data(), ns, M.data());
      matmul(O.dc2de1.data(), M.data(), Tu.data(), rd, K0, K1);
      matmul(O.dc2de0.data(), Tu.data(), T.data(), K1, rd, K1);
      for (int i = 0; i < K1; i++)
        std::memcpy(&O.d2d[(size_t)i * 8 * K1 + (size_t)c * K1], &T[(size_t)i * K1], sizeof(double) * K1);
    }
  }
  // M2L transfer spectra.  Child boxes of size 1: equivalent and check surfaces share the
  // lattice spacing h = RAD0/(m-1).  For source-minus-target child offset r, the kernel of
  // lattice displacement w = v - u is K(w h - r), stored at w mod n1 (circular convolution).
  {
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
      std::vector<std::complex<double>> s, all((size_t)ne * M_dim);
      for (int q = 0; q < ne; q++) {
        dft3_r2c(n1, grid[q], s);
        std::copy(s.begin(), s.end(), all.begin() + (size_t)q * M_dim);
      }
      spec[ri] = std::move(all);
    }
    const double inv_n3 = 1.0 / ((double)n1 * n1 * n1);
    const int bw = 8 * e1, bh = 8 * e0;  // block: rows t*e1+j, cols s*e0+i
    O.v1.assign((size_t)27 * M_dim * bw * bh, 0.0);
    O.v1_offsets.clear();
    for (int R = 0; R < 27; R++) {
      if (R == 13) continue;
      O.v1_offsets.push_back(R);
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
}

}  // namespace pvfmm_amd
