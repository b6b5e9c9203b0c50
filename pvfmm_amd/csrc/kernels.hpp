// Launch wrappers of the gfx950 kernels (implemented in kernels.hip).
#pragma once
#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace pvfmm_amd {

// Operands of the FP32 Hadamard kernels on the fp16 matrix cores (see hadamard_tiled).
struct HadamardParts {
  const uint16_t* Vh = nullptr;  // transfer blocks as two fp16 parts, scaled by tau_f
  const float* tinv = nullptr;   // 1 / tau_f per frequency (powers of two)
  const float* bound = nullptr;  // spectrum bound per source slot and component (fft_up's `bound`)
  float* run_scale = nullptr;    // scratch: one scale per run of tiles (scalar kernels)
  int run_scale_cap = 0;         // its length
};

// Point-kernel kinds of the device code: 0 laplace potential (1x1), 1 laplace gradient
// (1x3, r = x_trg - x_src, i.e. minus the gradient of the potential), 2 stokes velocity (3x3).
template <typename Real>
struct DeviceKernels {
  // sorted SoA densities from user AoS densities: den[c*n + i] = in[k0*perm[i] + c] * scal
  static void gather_density(hipStream_t s, int64_t n, int k0, const double* in, const int64_t* perm,
                             double scal, Real* den, int n_vec, int64_t in_stride);
  // out[k1*perm[i] + j] = val[j*n + i] * scal (user AoS output, in input order)
  static void scatter_output(hipStream_t s, int64_t n, int k1, const double* val, const int64_t* perm,
                             double scal, double* out, int n_vec, int64_t out_stride);
  static void fill(hipStream_t s, Real* p, int64_t n);
  // Periodic (PXYZ) post-processing of FMM_Pts::PostProcessing (include/fmm_pts.txx:5740-5774):
  // val[j*n + i] -= sum_c avg_c V_cj(x_i), avg_c the sum of component c over the root's K0
  // equivalent values eq_root[p*e0 + c], V the kernel's analytic potential of a uniform unit
  // density (kind 0: -|x|^2/6; kind 2: -(|x|^2 - x_j^2)/4 for c = j); x SoA absolute coordinates.
  static void vol_correction(hipStream_t s, int kind, int64_t n, const double* x, const Real* eq_root, int K0, int e0,
                             double* val, int nv, int64_t eq_vs, int64_t val_vs);
  // Its transpose: eq_root[p*e0 + c] += sum_i sum_j g[j*n + i] W_cj(x_i) for every p, W the
  // correction's factor (-V); per-block partial sums in a fixed order (part: vol_adjoint_blocks()
  // * 3 * nv doubles), then added in block order (deterministic).
  static int vol_adjoint_blocks();
  static void vol_correction_adjoint(hipStream_t s, int kind, int64_t n, const double* x, const Real* g, int K0, int e0,
                                     Real* eq_root, double* part, int nv, int64_t g_vs, int64_t eq_vs);
  static void fill64(hipStream_t s, double* p, int64_t n);

  // P2P into surface points (check surfaces): per row, targets are ns surface points of
  // node tnode[row] (offsets rad*size*unit_surf from its centre); sources are the point ranges
  // of the leaves items[item_ptr[row] .. item_ptr[row+1]).  out + orow[row]*ld (+)= result.
  static void p2p_surface(hipStream_t s, int kind, int n_rows, const int32_t* tnode, const int32_t* item_ptr,
                          const int32_t* items, const int32_t* src_begin, const int32_t* src_end,
                          const double* center, const float* node_size, double rad, const Real* unit_surf, int ns,
                          const Real* sx, const Real* sden, int64_t n_src, Real* out, const int32_t* orow,
                          int64_t ld, bool accumulate, int nv = 1, int64_t den_vs = 0, int64_t out_vs = 0);

  // All P2P launchers take nv density vectors: vector v reads densities at + v * den_vs and
  // writes values at + v * out_vs; up to 8 vectors share one pass over the pair geometry.
  // P2P into target points of leaves: point-range sources (U list).  tx/sx are FP64
  // leaf-local coordinates; the displacement is formed in FP64, the kernel runs in Real.
  static void p2p_points(hipStream_t s, int kind, int n_rows, const int32_t* tleaf, const int32_t* item_ptr,
                         const int32_t* items, const int32_t* trg_begin, const int32_t* trg_end,
                         const int32_t* src_begin, const int32_t* src_end, const double* center,
                         const double* tx, int64_t n_trg, const double* sx, const Real* sden, int64_t n_src,
                         double* out, bool accumulate, int nv = 1, int64_t den_vs = 0, int64_t out_vs = 0);

  // P2P into target points from surface sources (L2T own surface and W-list equivalent
  // surfaces).  items are pairs (node, density row pointer index): den row = dens + drow*ld.
  static void far_points(hipStream_t s, int kind, int n_rows, const int32_t* tleaf, const int32_t* item_ptr,
                         const int32_t* item_node, const int64_t* item_den, const double* item_rad,
                         const int32_t* trg_begin, const int32_t* trg_end, const double* center,
                         const float* node_size, const Real* unit_surf, int ns, const Real* tx, int64_t n_trg,
                         const Real* dens, double* out, bool accumulate, int nv = 1, int64_t den_vs = 0,
                         int64_t out_vs = 0);

  // C[rc[i]*ldc : +N] (+)= alpha[i] * A[ra[i]*lda : +K] * B[K x N] on MFMA; `work` (work_elems
  // Reals, may be null) enables a deterministic split-K for grids too small to fill the GPU.
  // split_blocks: K is split (through work[], work_elems reals) when the tile grid has fewer
  // blocks than this, so that about twice as many run.  nv density vectors run as one GEMM of
  // nv * M rows: vector v reads A + v * avs and writes C + v * cvs (same ra, rc, alpha).
  static void gemm_rows(hipStream_t s, int M, int N, int K, const Real* A, const int32_t* ra, int64_t lda,
                        const Real* B, Real* C, const int32_t* rc, int64_t ldc, const Real* alpha,
                        bool accumulate, Real* work = nullptr, int64_t work_elems = 0, int split_blocks = 1024,
                        int nv = 1, int64_t avs = 0, int64_t cvs = 0);

  // V-list: forward DFT of the 8 children's equivalent densities of each source parent.
  // X layout [slot][f][8*e0] complex, component-major within a row ([slot][f][i][child]).  lattice[3p..3p+2] is the m^3 lattice cell of surface
  // point p; lat_inv[(z*m + y)*m + x] is its inverse (-1 inside the cube).
  // vg, vi: X holds vector groups of vg density vectors, [slot][f][vector vg][8*e0 complex];
  // this call writes vector vi of the group (vg = 1, vi = 0: one vector).  bound (vg = 1 only;
  // may be null): bound[slot * e0 + i] = |scal[slot]| * sum of |up| of component i over the
  // slot's children, which no frequency of that component's spectrum exceeds.
  static void fft_up(hipStream_t s, int n_slots, const int32_t* first_child, const Real* up, int K0, int e0,
                     int m, const int32_t* lattice, const int32_t* lat_inv, const Real* scal, Real* X, int vg = 1,
                     int vi = 0, float* bound = nullptr);
  // Y[slot][f][8*e1] = sum_R V[R][f] X[partner[slot][R]][f] on MFMA (rows component-major,
  // as fft_up writes X and ifft_down reads Y); Vp is the transfer block
  // in real form, pre-permuted per lane (layout [R][f][kb][nb][lane 64][4]).  The target slots
  // form tiles of at most 64 consecutive slots [tile_t0[k], tile_t0[k+1])
  // whose distinct partner source slots usrc[tile_u0[k] .. tile_u0[k+1]) (at most
  // hadamard_union_max(e0)) are staged in LDS; uidx holds, for offset o and tile target
  // 16 * g + r, the partner's position in that list, or hadamard_union_max(e0) (the zero
  // entry) for none, at [k][o][r][g] ([k][n_off][16][4]), or for the FP32 scalar kernel at
  // [k][g][o / 16][o % 2][r][(o / 2) % 8] ([k][4][2][2][16][8]); bit o of gmask[k][g] is set when a
  // target of group g of tile k has a partner at offset o.  tiles_per_block: length of the
  // run of tiles one block of the FP32 scalar kernel handles.  hp (FP32 only, else null): the
  // FP32 kernels run on the fp16 matrix cores and need it.  hp->Vh the transfer blocks as two
  // fp16 parts of tau_f B for the offsets in the order of uidx, taken in pairs (2j, 2j+1):
  // scalar kernels [pair j][part p][f][lane 64][8], lane l holding part p of
  // B_{2j + l/32}[8 ((l/16) % 2) + i][l % 16], i < 8; Stokes [f][pair j][chunk c][tile nb]
  // [part p][lane 64][8], lane l holding part p of B_{2j + a}[8 b + i][16 nb + l % 16] with
  // q = 4c + l/16, a = q / 6, b = q % 6; hp->tinv 1 / tau_f; hp->bound the spectrum bounds of
  // X's slots and components (fft_up); hp->run_scale (scalar kernels) room for
  // ceil(n_tiles / tiles_per_block) scales.
  // They stage into an LDS pool of hadamard_pool_max() entries kept across a block's run of
  // tiles: usrc entries are (source slot << 9 | pool slot), uidx holds pool slots (the pool
  // size for none), and tile k's partners new since tile k-1 are dent[tile_d0[k] ..
  // tile_d0[k+1]) in the same packing; each list has one readable entry past its end.  The
  // other kernels ignore tile_d0 and dent.
  static int hadamard_union_max(int e0);
  static int hadamard_pool_max();
  static void hadamard_tiled(hipStream_t s, int n_tiles, int M_dim, int e0, int e1, int n_off, const int32_t* tile_t0,
                             const int32_t* tile_u0, const int32_t* usrc, const int32_t* tile_d0, const int32_t* dent,
                             const uint16_t* uidx, const uint32_t* gmask, const int32_t* offsets,
                             const Real* Vp, const Real* X, Real* Y, int tiles_per_block = 16,
                             const HadamardParts* hp = nullptr);
  // inverse DFT, pick the check-surface points, dn[child] += scal * value
  static void ifft_down(hipStream_t s, int n_slots, const int32_t* first_child, const Real* Y, Real* dn,
                        int K1, int e1, int m, const int32_t* lattice, const int32_t* lat_inv, int ns,
                        const Real* scal, int vg = 1, int vi = 0);
  // The FP32 scalar-kernel Hadamard for a group of 16 density vectors: the MFMA rows are the
  // 16 vectors of one target, X and Y in the vector-group layout of fft_up / ifft_down with
  // vg = 16, partner[t * 26 + o] the source slot of target slot t at offset v1_offsets[o] (or
  // -1), Vb the transfer blocks as three bf16 parts ([pair j][part p][f][lane 64][8], the lane
  // layout of hadamard_tiled's fp16 parts, unscaled).  targets_per_wave: targets one wave
  // computes in turn.
  static void hadamard_vec16(hipStream_t s, int n_trg, int M_dim, const int32_t* partner, const uint16_t* Vb,
                             const Real* X, Real* Y, int targets_per_wave = 16);
};

// FP64 all-pairs sum on the device (AoS src[3n], den[k0 n], trg[3n], out[k1 n] device arrays);
// kind as in DeviceKernels (0 laplace potential, 1 laplace gradient, 2 stokes velocity).
// The sources are split into n_chunks chunks of `chunk` points when the targets alone would
// not fill the GPU; the chunks' partial sums go to a scratch buffer of scratch_bytes (0 for one
// chunk) and are added in chunk order, so a result depends only on (kind, n_src, n_trg) and
// the inputs, whichever buffer holds the partials.
struct DirectSplit {
  int64_t chunk = 0;
  int n_chunks = 1;
  size_t scratch_bytes = 0;
};
DirectSplit direct_split(int kind, int64_t n_src, int64_t n_trg);
// part: device buffer of direct_split(...).scratch_bytes (may be null when that is 0)
void direct_sum(hipStream_t s, int kind, const double* src, int64_t n_src, const double* den, const double* trg, int64_t n_trg,
                double* out, double* part);
// The same sum with the partials in the library's static device array (direct_static.hip; no
// allocation): kDirectScratchBytes per device, shared by every call on that device; a larger
// need is rejected (kErrInvalidArgument).
constexpr size_t kDirectScratchBytes = (size_t)32 << 20;
void direct_sum_static(hipStream_t s, int kind, const double* src, int64_t n_src, const double* den, const double* trg,
                       int64_t n_trg, double* out);

}  // namespace pvfmm_amd
