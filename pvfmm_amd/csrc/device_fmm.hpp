// Device-resident evaluation plan of one tree: index arrays per phase, operator copies in the
// working precision, expansion buffers, and the launch sequence of one evaluation.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "octree.hpp"
#include "operators.hpp"

namespace pvfmm_amd {

[[noreturn]] void throw_device_error(hipError_t e, const char* file, int line);
#define HIP_CHECK(x)                                              \
  do {                                                            \
    hipError_t e_ = (x);                                          \
    if (e_ != hipSuccess) throw_device_error(e_, __FILE__, __LINE__); \
  } while (0)

// Phases timed by the plan (HIP events on the plan's stream), names follow the reference's
// Profile tags (include/fmm_tree.txx:283-713, include/fmm_pts.txx:3729-3755).
enum Phase : int {
  kPhGather = 0, kPhS2UP2P, kPhS2UGemm, kPhU2U, kPhXList, kPhFFT, kPhHadamard, kPhIFFT, kPhD2D, kPhD2TGemm,
  kPhFarP2P, kPhUList, kPhScatter, kPhaseCount
};
const char* phase_name(int p);

struct PlanStats {
  int64_t n_nodes = 0, n_leaves = 0, max_depth = 0;
  int64_t s2m_rows = 0, l2t_rows = 0, m2m_rows = 0, l2l_rows = 0, x_rows = 0, x_items = 0;
  int64_t v_src_slots = 0, v_trg_slots = 0, v_pairs = 0;
  int64_t u_rows = 0, u_items = 0, u_pairs = 0, far_items = 0, far_pairs = 0, s2m_pairs = 0;
  int64_t u_src_points_streamed = 0;
  int64_t device_bytes = 0;
  int64_t v_tiles = 0, v_tile_partners = 0, v_tile_max_partners = 0;
  int64_t v_tile_staged = 0;  // partner spectra staged per pass (the FP32 scalar kernel stages deltas)
};

// Launch shapes an evaluation uses; the defaults suit large trees, and autotune() picks them
// per tree by timing the candidates.
struct Tuning {
  int hadamard_tpb = 16;        // tiles per block run of the FP32 scalar Hadamard
  int gemm_split_blocks = 1024;  // GEMMs with fewer tile blocks than this split K
  int hadamard_vec_tpw = 16;    // targets per wave of the 16-vector FP32 scalar Hadamard (0: one vector at a time)
};

class DevicePlanBase {
 public:
  DevicePlanBase();
  virtual ~DevicePlanBase() = default;
  // Process-wide unique id of this plan: recorded graphs that bake in a plan's device
  // addresses are keyed on it, so a tree set up again (a new plan) is never replayed stale.
  const uint64_t generation;
  // Times each candidate launch shape on this tree (3 evaluations each, the fastest counts)
  // and keeps the best; d_den / d_out as for evaluate with n_vec vectors (one pass: at most
  // n_vec_max), d_out is overwritten.
  virtual Tuning autotune(const double* d_den, double* d_out, int n_vec) = 0;
  Tuning tuning;
  // d_den: user layout (k0 per source, input order) for n_vec vectors with stride k0*n_src;
  // d_out: user layout (k1 per target, input order), overwritten.
  virtual void evaluate(const double* d_den, double* d_out, int n_vec, double den_scale, double out_scale) = 0;
  // The transpose of evaluate's linear map (same scales): d_g in the user layout of the
  // targets (k1 per target), d_out in that of the sources (k0 per source), overwritten.  Built
  // on first use; symmetric kernels only (adjoint_supported()).
  virtual void evaluate_adjoint(const double* d_g, double* d_out, int n_vec, double den_scale, double out_scale) = 0;
  virtual bool adjoint_supported() const = 0;
  virtual hipStream_t stream() const = 0;
  virtual bool is_fp32() const = 0;
  // per-phase times of the last evaluation (ms); enabled by set_timing(true)
  void set_timing(bool on) { timing_ = on; }
  std::vector<double> phase_ms;
  PlanStats stats;
  int n_vec_max = 1;

 protected:
  bool timing_ = false;
};

// precision: 0 = FP64 throughout, 1 = FP32 bulk with FP64 operators precompute and FP64 merge
std::unique_ptr<DevicePlanBase> make_plan(const Octree& T, const Operators& O, const double* src_xyz, const double* trg_xyz,
                                          int precision, int n_vec_max, hipStream_t stream);

}  // namespace pvfmm_amd
