/*
 * C ABI of the MI355X particle FMM (libpvfmm.so built from pvfmm_amd/csrc).
 *
 * The PVFMM* entry points are the particle subset of the reference's C interface
 * (include/pvfmm.h:234-303 of dmalhotra/pvfmm, implemented there in
 * src/pvfmm-wrapper.cpp:365-933) with the same names, argument meaning and ownership:
 * the caller owns every array, the context owns the tree, operators and device buffers.
 * The reference functions return void and abort on errors; these keep the void signatures
 * and record the failure instead: PVFMMGetLastError() / PVFMMErrorString() report it and
 * no entry point calls exit().
 *
 * The run is one GPU taking the MPI_COMM_SELF share of the work, so the communicator
 * argument is accepted for signature compatibility and ignored.  It is declared as a
 * pointer-sized integer so that both MPICH (int) and Open MPI (pointer) handles pass
 * through the same register.
 *
 * The pvfmm_amd_* entry points are the library's own tree-level interface (used by the C++
 * header include/pvfmm.hpp and by the Python package); each cites the reference call it
 * stands for.
 */
#ifndef PVFMM_AMD_PVFMM_H
#define PVFMM_AMD_PVFMM_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef intptr_t pvfmm_comm_t;

/* include/pvfmm.h:21-28.  LaplacePotential, LaplaceGradient and StokesVelocity are
 * evaluated; the others are rejected with PVFMM_ERR_UNSUPPORTED. */
enum PVFMMKernel {
  PVFMMLaplacePotential = 0,
  PVFMMLaplaceGradient = 1,
  PVFMMStokesPressure = 2,
  PVFMMStokesVelocity = 3,
  PVFMMStokesVelocityGrad = 4,
  PVFMMBiotSavartPotential = 5
};

/* Error codes of PVFMMGetLastError() (no counterpart in the reference, which aborts). */
enum PVFMMError {
  PVFMM_OK = 0,
  PVFMM_ERR_INVALID_ARGUMENT = 1,
  PVFMM_ERR_UNSUPPORTED = 2,
  PVFMM_ERR_DEVICE = 3,
  PVFMM_ERR_OUT_OF_MEMORY = 4,
  PVFMM_ERR_STATE = 5,
  PVFMM_ERR_INTERNAL = 6
};

/* ---------------------------------------------------------------- reference particle C ABI */

/* include/pvfmm.h:254 -- double-precision context; box_size <= 0 selects free space with an
 * automatic bounding box (src/pvfmm-wrapper.cpp:457-514), box_size > 0 the periodic box
 * [0, box_size)^3 (BoundaryType Periodic, src/pvfmm-wrapper.cpp:397-403,514-519). */
void* PVFMMCreateContextD(double box_size, int n, int m, enum PVFMMKernel kernel, pvfmm_comm_t comm);
/* include/pvfmm.h:255 -- single-precision context (FP32 inputs/outputs, FP32 evaluation). */
void* PVFMMCreateContextF(float box_size, int n, int m, enum PVFMMKernel kernel, pvfmm_comm_t comm);
/* include/pvfmm.h:256-257 -- the same without a communicator (MPI_COMM_WORLD there). */
void* PVFMMCreateContextDWorld(double box_size, int n, int m, enum PVFMMKernel kernel);
void* PVFMMCreateContextFWorld(float box_size, int n, int m, enum PVFMMKernel kernel);

/* include/pvfmm.h:289-290 -- evaluate.  src_pos/trg_pos: AoS [3*n]; sl_den: [k0*n_src];
 * dl_den must be NULL (double-layer sources are not supported); trg_val: caller-allocated
 * [k1*n_trg].  setup != 0 rebuilds the tree for new positions, setup == 0 reuses it. */
void PVFMMEvalD(const double* src_pos, const double* sl_den, const double* dl_den, long n_src, const double* trg_pos,
                double* trg_val, long n_trg, const void* ctx, int setup);
void PVFMMEvalF(const float* src_pos, const float* sl_den, const float* dl_den, long n_src, const float* trg_pos,
                float* trg_val, long n_trg, const void* ctx, int setup);

/* include/pvfmm.h:298-299 -- destroy and set *ctx = NULL (src/pvfmm-wrapper.cpp:813-825). */
void PVFMMDestroyContextD(void** ctx);
void PVFMMDestroyContextF(void** ctx);

/* Fortran-callable particle context (src/pvfmm-wrapper.cpp:831-876): every argument by
 * reference, the communicator as a Fortran integer handle (ignored: one GPU).  Errors are
 * recorded as for the C calls; a failed create stores NULL in *ctx. */
void pvfmmcreatecontextd_(void** ctx, double* box_size, int32_t* n, int32_t* m, int32_t* kernel, int32_t* fcomm);
void pvfmmcreatecontextf_(void** ctx, float* box_size, int32_t* n, int32_t* m, int32_t* kernel, int32_t* fcomm);
void pvfmmevald_(const double* src_pos, const double* sl_den, int64_t* n_src, const double* trg_pos, double* trg_val,
                 int64_t* n_trg, void** ctx, int32_t* setup);
void pvfmmevalf_(const float* src_pos, const float* sl_den, int64_t* n_src, const float* trg_pos, float* trg_val,
                 int64_t* n_trg, void** ctx, int32_t* setup);
void pvfmmdestroycontextd_(void** ctx);
void pvfmmdestroycontextf_(void** ctx);

/* Error reporting; no counterpart in the reference (it aborts).  The state is per thread. */
int PVFMMGetLastError(void);
const char* PVFMMErrorString(int code);
const char* PVFMMGetLastErrorMessage(void);

/* Mixed precision for a double-precision context: FP32 P2P, M2L and up/down GEMMs with
 * FP64 operator precompute and an FP64 per-target merge.  No counterpart in the reference;
 * takes effect at the next setup.  Returns an error code. */
int PVFMMSetMixedPrecision(void* ctx, int enable);

/* Volume (Chebyshev) FMM, include/pvfmm.h:43-228 of the reference: not provided.  These
 * entry points exist so that bindings which resolve every symbol when they load the library
 * (python/src/pvfmm/ffi.py of the reference) can load this one; each records
 * PVFMM_ERR_UNSUPPORTED and returns NULL, 0 or nothing (the Destroy calls set *ctx = NULL). */
void* PVFMMCreateVolumeFMMD(int m, int q, enum PVFMMKernel kernel, pvfmm_comm_t comm);
void* PVFMMCreateVolumeFMMF(int m, int q, enum PVFMMKernel kernel, pvfmm_comm_t comm);
void* PVFMMCreateVolumeTreeD(int cheb_deg, int data_dim, void (*fn_ptr)(const double* coord, long n, double* out, const void* ctx),
                             const void* fn_ctx, const double* trg_coord, long n_trg, pvfmm_comm_t comm, double tol,
                             int max_pts, bool periodic, int init_depth);
void* PVFMMCreateVolumeTreeF(int cheb_deg, int data_dim, void (*fn_ptr)(const float* coord, long n, float* out, const void* ctx),
                             const void* fn_ctx, const float* trg_coord, long n_trg, pvfmm_comm_t comm, float tol, int max_pts,
                             bool periodic, int init_depth);
void* PVFMMCreateVolumeTreeFromCoeffD(long n_leaf, int cheb_deg, int data_dim, const double* leaf_coord, const double* fn_coeff,
                                      const double* trg_coord, long n_trg, pvfmm_comm_t comm, bool periodic);
void* PVFMMCreateVolumeTreeFromCoeffF(long n_leaf, int cheb_deg, int data_dim, const float* leaf_coord, const float* fn_coeff,
                                      const float* trg_coord, long n_trg, pvfmm_comm_t comm, bool periodic);
void PVFMMEvaluateVolumeFMMD(double* trg_value, void* tree, const void* fmm, long loc_size);
void PVFMMEvaluateVolumeFMMF(float* trg_value, void* tree, const void* fmm, long loc_size);
void PVFMMDestroyVolumeFMMD(void** ctx);
void PVFMMDestroyVolumeFMMF(void** ctx);
void PVFMMDestroyVolumeTreeD(void** ctx);
void PVFMMDestroyVolumeTreeF(void** ctx);
long PVFMMGetLeafCountD(const void* tree);
long PVFMMGetLeafCountF(const void* tree);
void PVFMMGetLeafCoordD(double* leaf_coord, const void* tree);
void PVFMMGetLeafCoordF(float* leaf_coord, const void* tree);
void PVFMMGetPotentialCoeffD(double* coeff, const void* tree);
void PVFMMGetPotentialCoeffF(float* coeff, const void* tree);
void PVFMMCoeff2NodesD(double* node_val, long n_leaf, int cheb_deg, int dof, const double* coeff);
void PVFMMCoeff2NodesF(float* node_val, long n_leaf, int cheb_deg, int dof, const float* coeff);
void PVFMMNodes2CoeffD(double* coeff, long n_leaf, int cheb_deg, int dof, const double* node_val);
void PVFMMNodes2CoeffF(float* coeff, long n_leaf, int cheb_deg, int dof, const float* node_val);

/* ---------------------------------------------------------------- tree-level interface */

/* FMM_Pts<...>::Initialize(m, comm, kernel) (include/fmm_pts.txx:223-331): operator
 * precompute for a kernel and multipole order.  precision: 0 FP64, 1 mixed FP32/FP64.
 * pinv_cutoff <= 0 keeps the reference's 4*eps*max_S truncation. */
void* pvfmm_amd_fmm_create(int kernel, int m, int precision, double pinv_cutoff);
void pvfmm_amd_fmm_destroy(void* fmm);

/* PtFMM_CreateTree(src_coord, src_value, trg_coord, comm, max_pts, FreeSpace)
 * (include/pvfmm.txx:215-245): coordinates in [0,1)^3, AoS. */
void* pvfmm_amd_tree_create(const double* src, int64_t n_src, const double* trg, int64_t n_trg, int max_pts);
/* PtFMM_CreateTree(..., max_pts, bndry) with the reference's boundary types (BoundaryType,
 * include/mpi_tree.hpp:22-29; include/pvfmm.txx:215-245): 0 FreeSpace, 1 PX, 2 PXY, 3 PXYZ
 * (= Periodic, unit period along the periodic axes).  pt: tree-defining points, or NULL for
 * the targets (n_pt is then ignored). */
void* pvfmm_amd_tree_create_bndry(const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt,
                                  int64_t n_pt, int max_pts, int bndry);
/* Variant with an explicit tree-defining point set (MPI_Node::pt_coord). */
void* pvfmm_amd_tree_create_pts(const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt,
                                int64_t n_pt, int max_pts);
void pvfmm_amd_tree_destroy(void* tree);

/* FMM_Tree::SetupFMM(fmm) (include/fmm_tree.txx:72-183): builds the device plan. */
int pvfmm_amd_tree_setup(void* tree, void* fmm, int n_vec_max);

/* PtFMM_Evaluate(tree, trg_val, n_trg, &sl_den) (include/pvfmm.txx:256-328) on host arrays:
 * den [n_vec][k0*n_src] in input order, out [n_vec][k1*n_trg] in input order. */
int pvfmm_amd_tree_evaluate(void* tree, const double* den, double* out, int n_vec);
/* Same with device-resident arrays; the work is queued on the tree's stream and this call
 * returns after it completes unless `async` is nonzero. */
int pvfmm_amd_tree_evaluate_device(void* tree, const double* d_den, double* d_out, int n_vec, int async);
/* The transpose of the evaluation's linear map (the adjoint FMM, for gradients with respect to
 * the densities): out = A^T g where pvfmm_amd_tree_evaluate computes out = A den.  g holds
 * [n_vec][k1*n_trg] values in target input order, out receives [n_vec][k0*n_src] in source
 * input order.  It runs the evaluation's phases in reverse order with the transposed
 * translations (S2M^T, U2U^T, the V-list M2L of the opposite offset, D2D^T, L2T^T and the
 * exchanged near-field interactions), so <g, A q> = <A^T g, q> holds to rounding.  The
 * reference has no backward pass: its nearest interface is PtFMM_Evaluate
 * (include/pvfmm.txx:256-328), whose translation operators (include/fmm_pts.txx:582-643,
 * 719-764, 2275-2599, 5427-5737) these transpose.  Symmetric kernels only (laplace_poten,
 * stokes_vel; laplace_grad reports PVFMM_ERR_UNSUPPORTED).  The first call builds the
 * transposed lists on the host and uploads them; it shares the tree's scratch buffers and
 * stream with the forward evaluation. */
int pvfmm_amd_tree_evaluate_adjoint(void* tree, const double* g, double* out, int n_vec);
int pvfmm_amd_tree_evaluate_adjoint_device(void* tree, const double* d_g, double* d_out, int n_vec, int async);
/* The HIP stream (hipStream_t) the tree's kernels run on. */
void* pvfmm_amd_tree_stream(void* tree);

/* Launch-shape tuning of a set-up tree; no counterpart in the reference (its CPU kernels have
 * no launch shapes).  params = {hadamard_tiles_per_block, gemm_split_blocks,
 * hadamard_vec_targets_per_wave} (tiles per block >= 1, used up to 64 by the FP32 kernel;
 * split blocks >= 0; targets per wave >= 0, used by the 16-vector FP32 kernel of plans set up
 * for 8 or more vectors, 0 running such passes one vector at a time through the V list).  set_tuning takes 2 or 3 values.  autotune times the candidates on
 * this tree with n_vec device-resident density vectors d_den (1 <= n_vec <= n_vec_max, as
 * pvfmm_amd_tree_evaluate_device; d_out is overwritten), keeps the fastest and writes the
 * chosen values to params[0..cap).  Returns 0 or an error code. */
int pvfmm_amd_tree_autotune(void* tree, const double* d_den, double* d_out, int n_vec, int* params, int cap);
int pvfmm_amd_tree_get_tuning(void* tree, int* params, int cap);
int pvfmm_amd_tree_set_tuning(void* tree, const int* params, int n);

/* Many set-up trees evaluated together (independent point clouds); no counterpart in the
 * reference, whose callers keep one context per cloud.  The first evaluation records every
 * tree's evaluation (one density vector each, device-resident) into one HIP graph with one
 * branch per tree; later calls with the same device pointers replay it with a single graph
 * launch on the batch's stream.  Other pointers, changed launch shapes of a tree, or a tree
 * set up again (its new plan has new device buffers) record the graph again.  A launch is
 * ordered after the work already queued on every tree's stream and before work queued there
 * later, so eager and batched evaluations of one tree may be mixed, also with async = 1.  The
 * trees must outlive the batch. */
void* pvfmm_amd_batch_create(void* const* trees, int n_trees);
int pvfmm_amd_batch_evaluate_device(void* batch, const double* const* d_den, double* const* d_out, int async);
void* pvfmm_amd_batch_stream(void* batch);
/* number of times the graph was recorded (1 while the pointers stay the same) */
int pvfmm_amd_batch_records(void* batch);
void pvfmm_amd_batch_destroy(void* batch);

/* Introspection for parity checks. */
int64_t pvfmm_amd_tree_num_nodes(void* tree);
/* per node (level-major order): x, y, z (units 2^-30), depth, is_leaf, parent, n_src, n_trg,
 * pt_cnt0, pt_cnt1 */
int pvfmm_amd_tree_nodes(void* tree, int64_t* out10);
/* sorted position -> input index for which = 0 (sources) or 1 (targets) */
int pvfmm_amd_tree_point_order(void* tree, int which, int64_t* out);
int pvfmm_amd_list_count(int type);
/* interaction list of Mat_Type `type` (include/precomp_mat.hpp:21-39): [n_nodes][count], -1 = empty */
int pvfmm_amd_tree_lists(void* tree, int type, int32_t* out);
/* per-phase times (ms) of the last evaluation after enabling timing; returns phase count */
int pvfmm_amd_tree_set_timing(void* tree, int on);
int pvfmm_amd_tree_phase_times(void* tree, double* ms, int cap);
const char* pvfmm_amd_phase_name(int phase);
/* plan statistics: n_nodes, n_leaves, max_depth, s2m_rows, l2t_rows, m2m_rows, l2l_rows, x_rows,
 * x_items, v_src_slots, v_trg_slots, v_pairs, u_rows, u_items, u_pairs, far_items, far_pairs,
 * s2m_pairs, u_src_points_streamed, device_bytes, v_tiles, v_tile_partners, v_tile_max_partners,
 * v_tile_staged (partner spectra a pass over the V-list tiles stages into LDS);
 * copies the first `cap` values and returns how many exist (negative error code on failure) */
int pvfmm_amd_tree_stats(void* tree, int64_t* out, int cap);

/* Operator matrices of an fmm handle (FP64), for parity checks against the oracle:
 * which 0 uc2ue0 (V S^+), 1 uc2ue1 (U^T), 2 dc2de0, 3 dc2de1, 4 u2u (8 stacked),
 * 5 d2d (8 side by side), 6 v1 (complex interleaved), 7/8/9 the periodic boundary operator
 * (BC_Type, include/fmm_pts.txx:790-1163) of PX / PXY / PXYZ, computed on first request.
 * Returns the element count; copies when out != NULL. */
int64_t pvfmm_amd_fmm_operator(void* fmm, int which, double* out);

/* Kernel<T>::ker_poten (include/kernel.hpp:84, used by the reference's examples for their
 * error checks): FP64 all-pairs sum on the GPU, out[k1*n_trg] = K(trg, src) den with the
 * kernel's 1/(4 pi) or 1/(8 pi) scale; r = 0 pairs contribute nothing.  Host arrays, AoS. */
int pvfmm_amd_direct(int kernel, const double* src, int64_t n_src, const double* den, const double* trg, int64_t n_trg,
                     double* out);

/* pvfmm_amd_direct on device arrays with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  d_src [3 n_src], d_den [k0 n_src],
 * d_trg [3 n_trg], d_out [k1 n_trg] are device arrays (AoS); the work is queued on `stream`
 * (a hipStream_t; NULL is the null stream) and the call returns without waiting.  The
 * partial sums of the source chunks live in one static 32 MiB device array per GPU; the chunk
 * rule keeps them within it (at most 19 MB), larger needs are rejected with
 * PVFMM_ERR_INVALID_ARGUMENT.  Every call on a GPU shares that array and overwrites it, so
 * calls that overlap on one GPU (other streams, other threads) race.  The first call on a
 * device looks the array up (hipGetSymbolAddress) and should run outside graph capture.
 * Results equal pvfmm_amd_direct's bit for bit. */
int pvfmm_amd_direct_device_noalloc(int kernel, const double* d_src, int64_t n_src, const double* d_den, const double* d_trg,
                                    int64_t n_trg, double* d_out, void* stream);

/* Number of GPUs visible; 0 when the HIP runtime finds none. */
int pvfmm_amd_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* PVFMM_AMD_PVFMM_H */
