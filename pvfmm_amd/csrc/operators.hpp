// Host FP64 precompute of the translation operators (the work of FMM_Pts::Initialize,
// include/fmm_pts.txx:223-331 and 453-764), written for the GPU data layout.
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace pvfmm_amd {

struct Operators {
  int m = 0;          // multipole order
  int ns = 0;         // surface points, 6(m-1)^2+2
  KernelInfo ker{};
  int n1 = 0, nh = 0, M_dim = 0;  // FFT grid 2m, half axis, spectrum length n1*n1*nh
  double pinv_rel_cutoff = 0;

  // surface lattice (reference point order): integer coordinates in {0..m-1}^3
  std::vector<int32_t> lattice;        // ns x 3
  // surface offsets from the box centre for a unit box, (lattice - (m-1)/2) * 2/(m-1) * 0.5
  std::vector<double> unit_surf;       // ns x 3; scale by rad * box size

  // expansion-space operators, row-vector convention y = x * M, all FP64.
  // Check -> equivalent maps are kept as the two SVD factors V*S^+ and U^T (as the reference's
  // UC2UE0/UC2UE1, include/fmm_pts.txx:478-581): with condition numbers near 1e16 at m=10 the
  // explicit product loses the accuracy that the factored application keeps.
  std::vector<double> uc2ue0, uc2ue1;  // (ns*e1) x rank_u, rank_u x (ns*e0); box size 1
  std::vector<double> dc2de0, dc2de1;  // (ns*e1) x rank_d, rank_d x (ns*e0); box size 1
  int rank_u = 0, rank_d = 0;
  std::vector<double> u2u;     // 8 stacked: (8*ns*e0) x (ns*e0), child c rows [c*ns*e0, (c+1)*ns*e0)
  std::vector<double> d2d;     // (ns*e1) x (8*ns*e1), child c columns [c*ns*e1, (c+1)*ns*e1)
  // M2L transfer blocks: [R 27][f M_dim][t*e1+j][s*e0+i] complex, R = (dx+1)+3(dy+1)+9(dz+1)
  // (R = 13 unused), 1/n1^3 included; zero blocks for adjacent child pairs.
  std::vector<std::complex<double>> v1;
  std::vector<int32_t> v1_offsets;  // the 26 offsets R in use

  // Periodic boundary operators by boundary type (1 PX, 2 PXY, 3 PXYZ; empty until
  // build_bc_operator): root upward equivalent density -> root downward check potential of all
  // periodic images beyond the root's colleagues, (ns*e0) x (ns*e1).
  std::vector<double> bc[4];

  bool on_device = false;   // the factors and U2U/D2D came from the device precompute
  bool from_cache = false;  // every operator was read from a precompute file

  int K0() const { return ns * ker.e0; }
  int K1() const { return ns * ker.e1; }
};

// cutoff_rel <= 0 selects the reference's rule 4*eps*max_S with eps of `precision`
// (0: FP64, 1: FP32 as the reference's float build).
void build_operators(Operators& O, int kernel_id, int m, double cutoff_rel = 0.0, int precision = 0);

// The same through a precompute file in directory `dir` (no file I/O when dir is null or empty,
// or for a custom cutoff): Precomp_<ker_name>_m<m>[_mixed].data in the container format of the
// reference's PrecompMat::Save2File / LoadFile (include/precomp_mat.txx:212-347).  Operators
// found in the file with the expected shapes are read, the others computed; a missing file is
// written (atomically, through a temporary name), an existing one is never replaced.
void build_operators_cached(Operators& O, int kernel_id, int m, double cutoff_rel, int precision, const char* dir);
std::string precomp_file_name(const Operators& O, int precision);

// The periodic boundary operator of boundary type bndry (1-3) into O.bc[bndry], unless it is
// there already: FMM_Pts::Precomp(BC_Type) of the reference's default build (include/fmm_pts.txx:
// 790-1163) -- the lattice sum over PVFMM_BC_LEVELS = 45 levels of images, each level the same
// operator for a box twice as large, with the zero-charge and zero-mean projections, the far
// field of the kernel's analytic volume potential (PXYZ; laplace potential and stokes velocity),
// and the subtraction of the constant and linear terms fitted at the cube corners; computed at
// order m + 2 and mapped to order m.  FP64; the products run on the GPU when the precompute does.
void build_bc_operator(Operators& O, int bndry);

// Integral over the cube [lo, lo + size]^3 of the expansion kernel (kind 0 laplace, 2 stokes)
// from a uniform unit density to each of n points (the reference's cheb_integ of degree 0,
// include/cheb_utils.txx:1037-1081): out[(p*e0 + a)*e1 + b] for source component a, target
// component b.  Points may lie inside the cube.
void cube_integral(int kind, const double* pts, int n, double lo, double size, double* out);

// Expansion-kernel matrix K[src*e0+i][trg*e1+j] (Kernel::BuildMatrix layout,
// include/kernel.txx:1059-1135), FP64.
void kernel_matrix(int kind, const double* src, int n_src, const double* trg, int n_trg, double* out);

// pinv of an (r x c) row-major matrix through one-sided Jacobi SVD, singular values below
// cutoff_rel*max are dropped.  Returns the factors F0 = V S^+ (c x rank) and F1 = U^T
// (rank x r) with pinv = F0 F1.
void pinv_jacobi(const std::vector<double>& A, int r, int c, double cutoff_rel, std::vector<double>& F0,
                 std::vector<double>& F1, int& rank);
// The factors from converged rotations (G: c rows U_j s_j of length r, W: c rows of V^T).
void pinv_factors(const std::vector<double>& G, const std::vector<double>& W, int r, int c, double cutoff_rel,
                  std::vector<double>& F0, std::vector<double>& F1, int& rank);

// The SVD, BuildMatrix and U2U/D2D products run on the GPU when one is visible (unless the
// environment sets PVFMM_PRECOMP=host); the host Jacobi path serves machines without one.
bool device_precompute_enabled();

}  // namespace pvfmm_amd
