// C ABI of libpvfmm: the reference's particle context interface (src/pvfmm-wrapper.cpp:365-933)
// and the library's tree-level interface.  All state lives in opaque handles; errors are
// recorded per thread and never abort the caller.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/pvfmm.h"
#include "common.hpp"
#include "device_fmm.hpp"
#include "gpu_octree.hpp"
#include "kernels.hpp"
#include "octree.hpp"
#include "operators.hpp"

using namespace pvfmm_amd;

namespace {

thread_local int g_err = 0;
thread_local std::string g_msg;

void set_error(int code, const std::string& msg) {
  g_err = code;
  g_msg = msg;
}
void clear_error() {
  g_err = 0;
  g_msg.clear();
}

template <typename F>
int guarded(F&& f) {
  clear_error();
  try {
    f();
    return 0;
  } catch (const Error& e) {
    set_error(e.code, e.what());
  } catch (const std::bad_alloc&) {
    set_error(kErrOutOfMemory, "host allocation failed");
  } catch (const std::exception& e) {
    set_error(kErrInternal, e.what());
  }
  return g_err;
}

struct FmmHandle {
  Operators O;
  int precision = 0;
  std::mutex bc_mutex;  // the boundary operators are built on first use, by any tree set up with this handle
};

struct DeviceArray {
  void* p = nullptr;
  size_t bytes = 0;
  ~DeviceArray() { if (p) (void)hipFree(p); }
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    HIP_CHECK(hipMalloc(&p, b));
    bytes = b;
  }
};

struct TreeHandle {
  Octree T;
  std::vector<double> src, trg;
  FmmHandle* fmm = nullptr;
  hipStream_t stream = nullptr;
  std::unique_ptr<DevicePlanBase> plan;
  DeviceArray d_den, d_out;
  ~TreeHandle() {
    plan.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }
  void ensure_stream() {
    if (!stream) HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  }
};

TreeHandle* make_tree(const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt, int64_t n_pt,
                      int max_pts, int bndry = kFreeSpace) {
  if (n_src < 0 || n_trg < 0 || n_pt < 0) throw Error(kErrInvalidArgument, "negative point count");
  if ((n_src && !src) || (n_trg && !trg) || (n_pt && !pt)) throw Error(kErrInvalidArgument, "null coordinate array");
  auto h = std::make_unique<TreeHandle>();
  h->src.assign(src, src + 3 * n_src);
  h->trg.assign(trg, trg + 3 * n_trg);
  // the tree-defining set of PtFMM_CreateTree is the target set itself (include/pvfmm.txx:239)
  const double* pts = (pt == trg && n_pt == n_trg) ? h->trg.data() : pt;
  if (device_tree_enabled())
    build_octree_device(h->T, h->src.data(), n_src, h->trg.data(), n_trg, pts, n_pt, max_pts, bndry);
  else
    build_octree(h->T, h->src.data(), n_src, h->trg.data(), n_trg, pts, n_pt, max_pts, bndry);
  return h.release();
}

void setup_tree(TreeHandle* t, FmmHandle* f, int n_vec_max) {
  if (!t || !f) throw Error(kErrInvalidArgument, "null handle");
  t->ensure_stream();
  t->plan.reset();
  t->fmm = f;
  if (t->T.bndry != kFreeSpace) {  // once per operator set and boundary type
    std::lock_guard<std::mutex> lock(f->bc_mutex);
    build_bc_operator(f->O, t->T.bndry);
  }
  t->plan = make_plan(t->T, f->O, t->src.data(), t->trg.data(), f->precision, n_vec_max, t->stream);
}

void eval_device(TreeHandle* t, const double* d_den, double* d_out, int n_vec, double den_scale, double out_scale) {
  if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
  t->plan->evaluate(d_den, d_out, n_vec, den_scale, out_scale);
}

// Independent trees evaluated as one HIP graph: a fork event on the batch stream, each
// tree's evaluation on its own stream, a join back.  Recorded once per set of device pointers.
struct BatchHandle {
  std::vector<TreeHandle*> trees;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const double*> den_key;
  std::vector<double*> out_key;
  std::vector<std::pair<int, int>> tune_key;  // launch shapes baked into the graph
  std::vector<uint64_t> plan_key;             // plans (device addresses) baked into the graph
  // ordering of a replay against eager work on the trees' own streams (outside the graph)
  std::vector<hipEvent_t> before;
  hipEvent_t after = nullptr;
  int records = 0;
  std::vector<std::pair<int, int>> tunings() const {
    std::vector<std::pair<int, int>> v;
    for (TreeHandle* t : trees)
      v.emplace_back(t->plan->tuning.hadamard_tpb, t->plan->tuning.gemm_split_blocks * 64 + t->plan->tuning.hadamard_vec_tpw);
    return v;
  }
  std::vector<uint64_t> plans() const {
    std::vector<uint64_t> v;
    for (TreeHandle* t : trees) v.push_back(t->plan->generation);
    return v;
  }
  ~BatchHandle() {
    drop_graph();
    for (hipEvent_t e : joins) (void)hipEventDestroy(e);
    for (hipEvent_t e : before) (void)hipEventDestroy(e);
    if (after) (void)hipEventDestroy(after);
    if (fork) (void)hipEventDestroy(fork);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void record(const double* const* d_den, double* const* d_out) {
    drop_graph();
    std::vector<bool> timing(trees.size());
    HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      HIP_CHECK(hipEventRecord(fork, stream));
      for (size_t i = 0; i < trees.size(); i++) {
        TreeHandle* t = trees[i];
        HIP_CHECK(hipStreamWaitEvent(t->stream, fork, 0));
        t->plan->evaluate(d_den[i], d_out[i], 1, 1.0, 1.0);
        HIP_CHECK(hipEventRecord(joins[i], t->stream));
        HIP_CHECK(hipStreamWaitEvent(stream, joins[i], 0));
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(stream, &graph));
    HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    den_key.assign(d_den, d_den + trees.size());
    out_key.assign(d_out, d_out + trees.size());
    tune_key = tunings();
    plan_key = plans();
    records++;
  }
};

BatchHandle* make_batch(void* const* trees, int n) {
  if (n <= 0 || !trees) throw Error(kErrInvalidArgument, "a batch needs at least one tree");
  auto b = std::make_unique<BatchHandle>();
  for (int i = 0; i < n; i++) {
    TreeHandle* t = (TreeHandle*)trees[i];
    if (!t || !t->plan) throw Error(kErrState, "every tree of a batch must be set up");
    b->trees.push_back(t);
  }
  HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  b->joins.resize(n, nullptr);
  b->before.resize(n, nullptr);
  for (int i = 0; i < n; i++) HIP_CHECK(hipEventCreateWithFlags(&b->joins[i], hipEventDisableTiming));
  for (int i = 0; i < n; i++) HIP_CHECK(hipEventCreateWithFlags(&b->before[i], hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&b->after, hipEventDisableTiming));
  return b.release();
}

void batch_eval(BatchHandle* b, const double* const* d_den, double* const* d_out, bool async) {
  if (!b) throw Error(kErrInvalidArgument, "null batch");
  if (!d_den || !d_out) throw Error(kErrInvalidArgument, "null pointer array");
  const size_t n = b->trees.size();
  for (size_t i = 0; i < n; i++) {
    if (!d_den[i] || !d_out[i]) throw Error(kErrInvalidArgument, "null device array");
    if (!b->trees[i]->plan) throw Error(kErrState, "a tree of the batch lost its plan");
    b->trees[i]->plan->set_timing(false);  // event records stay out of the graph
  }
  // a tree set up again has a new plan (new device buffers): the graph is recorded again
  const bool same = b->exec && std::equal(b->den_key.begin(), b->den_key.end(), d_den) &&
                    std::equal(b->out_key.begin(), b->out_key.end(), d_out) && b->tune_key == b->tunings() &&
                    b->plan_key == b->plans();
  if (!same) b->record(d_den, d_out);
  // the replay uses each plan's scratch buffers, as eager evaluations on the tree's own stream
  // do: it starts after the work queued there and that work resumes after it
  for (size_t i = 0; i < n; i++) {
    HIP_CHECK(hipEventRecord(b->before[i], b->trees[i]->stream));
    HIP_CHECK(hipStreamWaitEvent(b->stream, b->before[i], 0));
  }
  HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
  HIP_CHECK(hipEventRecord(b->after, b->stream));
  for (size_t i = 0; i < n; i++) HIP_CHECK(hipStreamWaitEvent(b->trees[i]->stream, b->after, 0));
  if (!async) HIP_CHECK(hipStreamSynchronize(b->stream));
}

void eval_adjoint_device(TreeHandle* t, const double* d_g, double* d_out, int n_vec) {
  if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
  if (n_vec > 0 && (!d_g || !d_out)) throw Error(kErrInvalidArgument, "null device array");
  t->plan->evaluate_adjoint(d_g, d_out, n_vec, 1.0, 1.0);
}

void eval_host(TreeHandle* t, const double* den, double* out, int n_vec, double den_scale, double out_scale) {
  if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
  const KernelInfo& ki = t->fmm->O.ker;
  const size_t nd = (size_t)ki.k0 * t->T.src.n * n_vec, no = (size_t)ki.k1 * t->T.trg.n * n_vec;
  t->d_den.ensure(std::max<size_t>(8, nd * sizeof(double)));
  t->d_out.ensure(std::max<size_t>(8, no * sizeof(double)));
  if (nd) HIP_CHECK(hipMemcpyAsync(t->d_den.p, den, nd * sizeof(double), hipMemcpyHostToDevice, t->stream));
  eval_device(t, (const double*)t->d_den.p, (double*)t->d_out.p, n_vec, den_scale, out_scale);
  if (no) HIP_CHECK(hipMemcpyAsync(out, t->d_out.p, no * sizeof(double), hipMemcpyDeviceToHost, t->stream));
  HIP_CHECK(hipStreamSynchronize(t->stream));
}

// ----------------------------------------------------------------------------- PVFMM context
struct Context {
  double box_size = 0;
  int max_pts = 0, m = 0, kernel = 0;
  bool single = false;  // created by PVFMMCreateContextF
  int precision = 0;
  std::unique_ptr<FmmHandle> fmm;
  std::vector<double> grid;  // initial uniform points of the context tree (src/pvfmm-wrapper.cpp:416-432)
  std::unique_ptr<TreeHandle> tree;
  int64_t n_src = -1, n_trg = -1;
  double scale_x = 1, shift_x[3] = {0, 0, 0};
};

template <typename Real>
Context* create_context(Real box_size, int n, int m, int kernel, bool single) {
  if (n <= 0) throw Error(kErrInvalidArgument, "max points per box must be positive");
  kernel_info(kernel);  // validates the kernel
  auto c = std::make_unique<Context>();
  c->box_size = (double)box_size;
  c->max_pts = n;
  c->m = m;
  c->kernel = kernel;
  c->single = single;
  c->precision = single ? 1 : 0;
  c->fmm = std::make_unique<FmmHandle>();
  build_operators_cached(c->fmm->O, kernel, m, 0.0, c->precision, std::getenv("PVFMM_DIR"));
  c->fmm->precision = c->precision;
  // src/pvfmm-wrapper.cpp:416-432 at np = 1, in the context's precision
  size_t NN = (size_t)std::ceil(std::pow((Real)n, (Real)(1.0 / 3.0)));
  size_t total = NN * NN * NN;
  c->grid.resize(3 * total);
  for (size_t i = 0; i < total; i++) {
    c->grid[3 * i + 0] = (double)((Real)(((i / 1) % NN) + 0.5) / NN);
    c->grid[3 * i + 1] = (double)((Real)(((i / NN) % NN) + 0.5) / NN);
    c->grid[3 * i + 2] = (double)((Real)(((i / (NN * NN)) % NN) + 0.5) / NN);
  }
  return c.release();
}

// Bounding-box map to [0,1)^3, src/pvfmm-wrapper.cpp:457-520
template <typename Real>
void bounding_box(size_t n, const Real* x, Real* scale_xr, Real* shift_xr) {
  double mn[3], mx[3];
  for (int k = 0; k < 3; k++) mn[k] = mx[k] = x[k];
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      mn[k] = std::min(mn[k], (double)x[3 * i + k]);
      mx[k] = std::max(mx[k], (double)x[3 * i + k]);
    }
  Real eps = std::numeric_limits<Real>::epsilon() * 64;
  Real s = 1 / (Real)(mx[0] - mn[0] + 2 * eps);
  for (int k = 0; k < 3; k++) s = std::min(s, (Real)(1.0 / (mx[k] - mn[k] + 2 * eps)));
  if (s * 0.0 != 0.0) s = 1.0;
  *scale_xr = s;
  for (int k = 0; k < 3; k++) shift_xr[k] = (Real)(-mn[k] * s + eps);
}

template <typename Real>
void context_eval(Context* c, const Real* src_pos, const Real* sl_den, const Real* dl_den, long n_src, const Real* trg_pos,
                  Real* trg_val, long n_trg, int setup) {
  if (!c) throw Error(kErrInvalidArgument, "null context");
  if (dl_den) throw Error(kErrUnsupported, "double-layer densities are not supported (pass dl_den = NULL)");
  if (!sl_den && n_src > 0) throw Error(kErrInvalidArgument, "sl_den is NULL");
  if (n_src <= 0 || n_trg <= 0) throw Error(kErrInvalidArgument, "n_src and n_trg must be positive");
  if (!src_pos || !trg_pos || !trg_val) throw Error(kErrInvalidArgument, "null array");
  const KernelInfo ki = kernel_info(c->kernel);
  if (setup || !c->tree) {
    Real scale_x, shift_x[3];
    if (c->box_size <= 0) {
      Real s0, x0[3], s1, x1[3];
      bounding_box<Real>(n_src, src_pos, &s0, x0);
      bounding_box<Real>(n_trg, trg_pos, &s1, x1);
      Real c0[3] = {(Real)(0.5 - x0[0]) / s0, (Real)(0.5 - x0[1]) / s0, (Real)(0.5 - x0[2]) / s0};
      Real c1[3] = {(Real)(0.5 - x1[0]) / s1, (Real)(0.5 - x1[1]) / s1, (Real)(0.5 - x1[2]) / s1};
      scale_x = 0;
      for (int k = 0; k < 3; k++) scale_x = std::max<Real>(scale_x, std::fabs(c0[k] - c1[k]));
      scale_x = 1 / (scale_x + 1 / s0 + 1 / s1);
      for (int k = 0; k < 3; k++) shift_x[k] = (Real)0.5 - (c0[k] + c1[k]) * scale_x / 2;
    } else {  // periodic box [0, box_size)^3 (src/pvfmm-wrapper.cpp:514-519): coordinates wrap into it
      scale_x = 1 / (Real)c->box_size;
      shift_x[0] = shift_x[1] = shift_x[2] = 0;
    }
    auto map = [&](const Real* p, long n, std::vector<double>& out) {
      out.resize(3 * (size_t)n);
      for (long i = 0; i < n; i++)
        for (int j = 0; j < 3; j++) {
          Real v = p[3 * i + j] * scale_x + shift_x[j];
          while (v < 0.0) v += 1;
          while (v >= 1.0) v -= 1;
          out[3 * i + j] = (double)v;
        }
    };
    std::vector<double> sp, tp;
    map(src_pos, n_src, sp);
    map(trg_pos, n_trg, tp);
    c->tree.reset();
    c->tree.reset(make_tree(sp.data(), n_src, tp.data(), n_trg, c->grid.data(), (int64_t)(c->grid.size() / 3), c->max_pts,
                            c->box_size > 0 ? kPXYZ : kFreeSpace));
    if (c->fmm->precision != c->precision) {
      // the pseudo-inverse truncation depends on the evaluation precision
      build_operators_cached(c->fmm->O, c->kernel, c->m, 0.0, c->precision, std::getenv("PVFMM_DIR"));
      c->fmm->precision = c->precision;
    }
    setup_tree(c->tree.get(), c->fmm.get(), 1);
    c->scale_x = (double)scale_x;
    c->n_src = n_src;
    c->n_trg = n_trg;
  } else if (n_src != c->n_src || n_trg != c->n_trg) {
    throw Error(kErrInvalidArgument, "setup = 0 requires the point counts of the last setup");
  }
  // densities * scale_x^src_scal, outputs * scale_x^trg_scal (src/pvfmm-wrapper.cpp:522-541, 805)
  const double den_scale = std::pow(c->scale_x, 0.0), out_scale = std::pow(c->scale_x, ki.out_scal);
  std::vector<double> den((size_t)ki.k0 * n_src), out((size_t)ki.k1 * n_trg);
  for (size_t i = 0; i < den.size(); i++) den[i] = (double)sl_den[i];
  eval_host(c->tree.get(), den.data(), out.data(), 1, den_scale, out_scale);
  for (size_t i = 0; i < out.size(); i++) trg_val[i] = (Real)out[i];
}

}  // namespace

extern "C" {

void* PVFMMCreateContextD(double box_size, int n, int m, enum PVFMMKernel kernel, pvfmm_comm_t) {
  Context* c = nullptr;
  guarded([&] { c = create_context<double>(box_size, n, m, (int)kernel, false); });
  return c;
}
void* PVFMMCreateContextF(float box_size, int n, int m, enum PVFMMKernel kernel, pvfmm_comm_t) {
  Context* c = nullptr;
  guarded([&] { c = create_context<float>(box_size, n, m, (int)kernel, true); });
  return c;
}
void* PVFMMCreateContextDWorld(double box_size, int n, int m, enum PVFMMKernel kernel) {
  return PVFMMCreateContextD(box_size, n, m, kernel, 0);
}
void* PVFMMCreateContextFWorld(float box_size, int n, int m, enum PVFMMKernel kernel) {
  return PVFMMCreateContextF(box_size, n, m, kernel, 0);
}
void PVFMMEvalD(const double* src_pos, const double* sl_den, const double* dl_den, long n_src, const double* trg_pos,
                double* trg_val, long n_trg, const void* ctx, int setup) {
  guarded([&] { context_eval<double>((Context*)ctx, src_pos, sl_den, dl_den, n_src, trg_pos, trg_val, n_trg, setup); });
}
void PVFMMEvalF(const float* src_pos, const float* sl_den, const float* dl_den, long n_src, const float* trg_pos,
                float* trg_val, long n_trg, const void* ctx, int setup) {
  guarded([&] { context_eval<float>((Context*)ctx, src_pos, sl_den, dl_den, n_src, trg_pos, trg_val, n_trg, setup); });
}
void PVFMMDestroyContextD(void** ctx) {
  if (!ctx || !*ctx) return;
  delete (Context*)*ctx;
  *ctx = nullptr;
}
void PVFMMDestroyContextF(void** ctx) { PVFMMDestroyContextD(ctx); }

// Fortran symbols: arguments by reference (src/pvfmm-wrapper.cpp:831-876)
void pvfmmcreatecontextd_(void** ctx, double* box_size, int32_t* n, int32_t* m, int32_t* kernel, int32_t*) {
  if (!ctx) return;
  *ctx = nullptr;
  guarded([&] {
    if (!box_size || !n || !m || !kernel) throw Error(kErrInvalidArgument, "null argument");
    *ctx = create_context<double>(*box_size, *n, *m, *kernel, false);
  });
}
void pvfmmcreatecontextf_(void** ctx, float* box_size, int32_t* n, int32_t* m, int32_t* kernel, int32_t*) {
  if (!ctx) return;
  *ctx = nullptr;
  guarded([&] {
    if (!box_size || !n || !m || !kernel) throw Error(kErrInvalidArgument, "null argument");
    *ctx = create_context<float>(*box_size, *n, *m, *kernel, true);
  });
}
void pvfmmevald_(const double* src_pos, const double* sl_den, int64_t* n_src, const double* trg_pos, double* trg_val,
                 int64_t* n_trg, void** ctx, int32_t* setup) {
  guarded([&] {
    if (!n_src || !n_trg || !ctx || !setup) throw Error(kErrInvalidArgument, "null argument");
    context_eval<double>((Context*)*ctx, src_pos, sl_den, nullptr, (long)*n_src, trg_pos, trg_val, (long)*n_trg, *setup);
  });
}
void pvfmmevalf_(const float* src_pos, const float* sl_den, int64_t* n_src, const float* trg_pos, float* trg_val,
                 int64_t* n_trg, void** ctx, int32_t* setup) {
  guarded([&] {
    if (!n_src || !n_trg || !ctx || !setup) throw Error(kErrInvalidArgument, "null argument");
    context_eval<float>((Context*)*ctx, src_pos, sl_den, nullptr, (long)*n_src, trg_pos, trg_val, (long)*n_trg, *setup);
  });
}
void pvfmmdestroycontextd_(void** ctx) { PVFMMDestroyContextD(ctx); }
void pvfmmdestroycontextf_(void** ctx) { PVFMMDestroyContextD(ctx); }

int PVFMMGetLastError(void) { return g_err; }
const char* PVFMMGetLastErrorMessage(void) { return g_msg.c_str(); }
const char* PVFMMErrorString(int code) {
  switch (code) {
    case PVFMM_OK: return "no error";
    case PVFMM_ERR_INVALID_ARGUMENT: return "invalid argument";
    case PVFMM_ERR_UNSUPPORTED: return "unsupported feature";
    case PVFMM_ERR_DEVICE: return "device (HIP) error";
    case PVFMM_ERR_OUT_OF_MEMORY: return "out of memory";
    case PVFMM_ERR_STATE: return "invalid state";
    case PVFMM_ERR_INTERNAL: return "internal error";
    default: return "unknown error";
  }
}
int PVFMMSetMixedPrecision(void* ctx, int enable) {
  return guarded([&] {
    Context* c = (Context*)ctx;
    if (!c) throw Error(kErrInvalidArgument, "null context");
    if (c->single && !enable) throw Error(kErrUnsupported, "single-precision contexts always evaluate in FP32");
    c->precision = enable ? 1 : 0;
  });
}

void* pvfmm_amd_fmm_create(int kernel, int m, int precision, double pinv_cutoff) {
  FmmHandle* f = nullptr;
  guarded([&] {
    if (precision != 0 && precision != 1) throw Error(kErrInvalidArgument, "precision must be 0 or 1");
    auto h = std::make_unique<FmmHandle>();
    build_operators_cached(h->O, kernel, m, pinv_cutoff, precision, std::getenv("PVFMM_DIR"));
    h->precision = precision;
    f = h.release();
  });
  return f;
}
void pvfmm_amd_fmm_destroy(void* fmm) { delete (FmmHandle*)fmm; }

void* pvfmm_amd_tree_create(const double* src, int64_t n_src, const double* trg, int64_t n_trg, int max_pts) {
  TreeHandle* t = nullptr;
  guarded([&] { t = make_tree(src, n_src, trg, n_trg, trg, n_trg, max_pts); });
  return t;
}
void* pvfmm_amd_tree_create_bndry(const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt,
                                  int64_t n_pt, int max_pts, int bndry) {
  TreeHandle* t = nullptr;
  guarded([&] { t = make_tree(src, n_src, trg, n_trg, pt ? pt : trg, pt ? n_pt : n_trg, max_pts, bndry); });
  return t;
}
void* pvfmm_amd_tree_create_pts(const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt,
                                int64_t n_pt, int max_pts) {
  TreeHandle* t = nullptr;
  guarded([&] { t = make_tree(src, n_src, trg, n_trg, pt, n_pt, max_pts); });
  return t;
}
void pvfmm_amd_tree_destroy(void* tree) { delete (TreeHandle*)tree; }

int pvfmm_amd_tree_setup(void* tree, void* fmm, int n_vec_max) {
  return guarded([&] { setup_tree((TreeHandle*)tree, (FmmHandle*)fmm, n_vec_max); });
}
int pvfmm_amd_tree_evaluate(void* tree, const double* den, double* out, int n_vec) {
  return guarded([&] { eval_host((TreeHandle*)tree, den, out, n_vec, 1.0, 1.0); });
}
int pvfmm_amd_tree_evaluate_device(void* tree, const double* d_den, double* d_out, int n_vec, int async) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    eval_device(t, d_den, d_out, n_vec, 1.0, 1.0);
    if (!async) HIP_CHECK(hipStreamSynchronize(t->stream));
  });
}
int pvfmm_amd_tree_evaluate_adjoint_device(void* tree, const double* d_g, double* d_out, int n_vec, int async) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    eval_adjoint_device(t, d_g, d_out, n_vec);
    if (!async) HIP_CHECK(hipStreamSynchronize(t->stream));
  });
}
int pvfmm_amd_tree_evaluate_adjoint(void* tree, const double* g, double* out, int n_vec) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
    if (n_vec < 1) throw Error(kErrInvalidArgument, "n_vec must be >= 1");
    const KernelInfo& ki = t->fmm->O.ker;
    const size_t ng = (size_t)ki.k1 * t->T.trg.n * n_vec, no = (size_t)ki.k0 * t->T.src.n * n_vec;
    // the host staging buffers of eval_host, with the roles of input and output exchanged
    t->d_out.ensure(std::max<size_t>(8, ng * sizeof(double)));
    t->d_den.ensure(std::max<size_t>(8, no * sizeof(double)));
    if (ng) HIP_CHECK(hipMemcpyAsync(t->d_out.p, g, ng * sizeof(double), hipMemcpyHostToDevice, t->stream));
    eval_adjoint_device(t, (const double*)t->d_out.p, (double*)t->d_den.p, n_vec);
    if (no) HIP_CHECK(hipMemcpyAsync(out, t->d_den.p, no * sizeof(double), hipMemcpyDeviceToHost, t->stream));
    HIP_CHECK(hipStreamSynchronize(t->stream));
  });
}
void* pvfmm_amd_tree_stream(void* tree) { return tree ? (void*)((TreeHandle*)tree)->stream : nullptr; }

int pvfmm_amd_tree_autotune(void* tree, const double* d_den, double* d_out, int n_vec, int* params, int cap) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
    if (!d_den || !d_out) throw Error(kErrInvalidArgument, "null device array");
    const Tuning tu = t->plan->autotune(d_den, d_out, n_vec);
    HIP_CHECK(hipStreamSynchronize(t->stream));
    const int v[3] = {tu.hadamard_tpb, tu.gemm_split_blocks, tu.hadamard_vec_tpw};
    for (int i = 0; i < std::min(cap, 3); i++) params[i] = v[i];
  });
}
int pvfmm_amd_tree_get_tuning(void* tree, int* params, int cap) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
    const int v[3] = {t->plan->tuning.hadamard_tpb, t->plan->tuning.gemm_split_blocks, t->plan->tuning.hadamard_vec_tpw};
    for (int i = 0; i < std::min(cap, 3); i++) params[i] = v[i];
  });
}
int pvfmm_amd_tree_set_tuning(void* tree, const int* params, int n) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    if (!t || !t->plan) throw Error(kErrState, "tree has no device plan; call setup first");
    if (n < 2 || !params) throw Error(kErrInvalidArgument, "at least two tuning values expected");
    if (params[0] < 1 || params[1] < 0 || (n > 2 && params[2] < 0))
      throw Error(kErrInvalidArgument, "tiles per block >= 1, split blocks >= 0, targets per wave >= 0");
    t->plan->tuning.hadamard_tpb = params[0];
    t->plan->tuning.gemm_split_blocks = params[1];
    if (n > 2) t->plan->tuning.hadamard_vec_tpw = params[2];
  });
}

void* pvfmm_amd_batch_create(void* const* trees, int n_trees) {
  BatchHandle* b = nullptr;
  guarded([&] { b = make_batch(trees, n_trees); });
  return b;
}
int pvfmm_amd_batch_evaluate_device(void* batch, const double* const* d_den, double* const* d_out, int async) {
  return guarded([&] { batch_eval((BatchHandle*)batch, d_den, d_out, async != 0); });
}
void* pvfmm_amd_batch_stream(void* batch) { return batch ? (void*)((BatchHandle*)batch)->stream : nullptr; }
int pvfmm_amd_batch_records(void* batch) { return batch ? ((BatchHandle*)batch)->records : -1; }
void pvfmm_amd_batch_destroy(void* batch) { delete (BatchHandle*)batch; }

int64_t pvfmm_amd_tree_num_nodes(void* tree) { return tree ? ((TreeHandle*)tree)->T.n_nodes() : -1; }
int pvfmm_amd_tree_nodes(void* tree, int64_t* out) {
  return guarded([&] {
    const Octree& T = ((TreeHandle*)tree)->T;
    for (int32_t i = 0; i < T.n_nodes(); i++) {
      int64_t* o = &out[10 * (size_t)i];
      o[0] = T.x[i]; o[1] = T.y[i]; o[2] = T.z[i]; o[3] = T.depth[i]; o[4] = T.is_leaf(i); o[5] = T.parent[i];
      o[6] = T.is_leaf(i) ? T.pt_cnt0(i) : 0; o[7] = T.is_leaf(i) ? T.pt_cnt1(i) : 0; o[8] = T.pt_cnt0(i); o[9] = T.pt_cnt1(i);
    }
  });
}
int pvfmm_amd_tree_point_order(void* tree, int which, int64_t* out) {
  return guarded([&] {
    const Octree& T = ((TreeHandle*)tree)->T;
    const PointSet& P = which == 0 ? T.src : T.trg;
    std::copy(P.perm.begin(), P.perm.end(), out);
  });
}
int pvfmm_amd_list_count(int type) {
  if (type < 0 || type >= kTypeCount) return -1;
  return (int)rel_table(type).rel.size();
}
int pvfmm_amd_tree_lists(void* tree, int type, int32_t* out) {
  return guarded([&] {
    const Octree& T = ((TreeHandle*)tree)->T;
    if (type < 0 || type >= kTypeCount) throw Error(kErrInvalidArgument, "bad list type");
    const size_t n = (size_t)T.n_nodes() * T.list_count(type);
    if (T.lists[type].size() == n) std::copy(T.lists[type].begin(), T.lists[type].end(), out);
    else std::fill(out, out + n, -1);
  });
}
int pvfmm_amd_tree_set_timing(void* tree, int on) {
  return guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    if (!t->plan) throw Error(kErrState, "no plan");
    t->plan->set_timing(on != 0);
  });
}
int pvfmm_amd_tree_phase_times(void* tree, double* ms, int cap) {
  TreeHandle* t = (TreeHandle*)tree;
  if (!t || !t->plan) return 0;
  int n = std::min<int>(cap, (int)t->plan->phase_ms.size());
  for (int i = 0; i < n; i++) ms[i] = t->plan->phase_ms[i];
  return (int)t->plan->phase_ms.size();
}
const char* pvfmm_amd_phase_name(int phase) { return phase_name(phase); }
int pvfmm_amd_tree_stats(void* tree, int64_t* o, int cap) {
  int count = 0;
  const int err = guarded([&] {
    TreeHandle* t = (TreeHandle*)tree;
    if (!t || !t->plan) throw Error(kErrState, "no plan");
    if (cap < 0 || (cap > 0 && !o)) throw Error(kErrInvalidArgument, "bad output array");
    const PlanStats& s = t->plan->stats;
    const int64_t v[] = {s.n_nodes, s.n_leaves, s.max_depth, s.s2m_rows, s.l2t_rows, s.m2m_rows, s.l2l_rows, s.x_rows,
                         s.x_items, s.v_src_slots, s.v_trg_slots, s.v_pairs, s.u_rows, s.u_items, s.u_pairs, s.far_items,
                         s.far_pairs, s.s2m_pairs, s.u_src_points_streamed, s.device_bytes, s.v_tiles, s.v_tile_partners,
                         s.v_tile_max_partners, s.v_tile_staged};
    count = (int)(sizeof(v) / sizeof(v[0]));
    std::copy(v, v + std::min(cap, count), o);
  });
  return err ? -err : count;
}
int64_t pvfmm_amd_fmm_operator(void* fmm, int which, double* out) {
  FmmHandle* f = (FmmHandle*)fmm;
  if (!f) return -1;
  if (which >= 7 && which <= 9) {  // periodic boundary operators, built on first request
    int64_t n = -1;
    if (guarded([&] {
          std::lock_guard<std::mutex> lock(f->bc_mutex);
          build_bc_operator(f->O, which - 6);
          n = (int64_t)f->O.bc[which - 6].size();
        }) != 0)
      return -1;
    if (out) std::copy(f->O.bc[which - 6].begin(), f->O.bc[which - 6].end(), out);
    return n;
  }
  const Operators& O = f->O;
  const std::vector<double>* v = nullptr;
  switch (which) {
    case 0: v = &O.uc2ue0; break;
    case 1: v = &O.uc2ue1; break;
    case 2: v = &O.dc2de0; break;
    case 3: v = &O.dc2de1; break;
    case 4: v = &O.u2u; break;
    case 5: v = &O.d2d; break;
    case 6:
      if (out) std::memcpy(out, O.v1.data(), sizeof(double) * 2 * O.v1.size());
      return (int64_t)(2 * O.v1.size());
    default: return -1;
  }
  if (out) std::copy(v->begin(), v->end(), out);
  return (int64_t)v->size();
}
int pvfmm_amd_direct(int kernel, const double* src, int64_t n_src, const double* den, const double* trg, int64_t n_trg,
                     double* out) {
  return guarded([&] {
    const KernelInfo& ki = kernel_info(kernel);
    if (n_src < 0 || n_trg < 0 || (n_src && (!src || !den)) || (n_trg && (!trg || !out)))
      throw Error(kErrInvalidArgument, "bad direct-sum arguments");
    if (n_trg == 0) return;
    if (n_src == 0) {
      std::fill(out, out + ki.k1 * n_trg, 0.0);
      return;
    }
    struct Buf {
      double* p = nullptr;
      ~Buf() { if (p) (void)hipFree(p); }
    } ds, dd, dt, dout, part;
    const DirectSplit sp = direct_split(ki.pt_kind, n_src, n_trg);
    if (sp.scratch_bytes) HIP_CHECK(hipMalloc(&part.p, sp.scratch_bytes));
    HIP_CHECK(hipMalloc(&ds.p, sizeof(double) * 3 * n_src));
    HIP_CHECK(hipMalloc(&dd.p, sizeof(double) * ki.k0 * n_src));
    HIP_CHECK(hipMalloc(&dt.p, sizeof(double) * 3 * n_trg));
    HIP_CHECK(hipMalloc(&dout.p, sizeof(double) * ki.k1 * n_trg));
    HIP_CHECK(hipMemcpy(ds.p, src, sizeof(double) * 3 * n_src, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dd.p, den, sizeof(double) * ki.k0 * n_src, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dt.p, trg, sizeof(double) * 3 * n_trg, hipMemcpyHostToDevice));
    direct_sum(nullptr, ki.pt_kind, ds.p, n_src, dd.p, dt.p, n_trg, dout.p, part.p);
    HIP_CHECK(hipMemcpy(out, dout.p, sizeof(double) * ki.k1 * n_trg, hipMemcpyDeviceToHost));
  });
}
int pvfmm_amd_direct_device_noalloc(int kernel, const double* d_src, int64_t n_src, const double* d_den, const double* d_trg,
                                    int64_t n_trg, double* d_out, void* stream) {
  return guarded([&] {
    const KernelInfo& ki = kernel_info(kernel);
    if (n_src < 0 || n_trg < 0 || (n_src && (!d_src || !d_den)) || (n_trg && (!d_trg || !d_out)))
      throw Error(kErrInvalidArgument, "bad direct-sum arguments");
    if (n_trg == 0) return;
    direct_sum_static(static_cast<hipStream_t>(stream), ki.pt_kind, d_src, n_src, d_den, d_trg, n_trg, d_out);
  });
}
int pvfmm_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ---- volume (Chebyshev) FMM, include/pvfmm.h:43-228 of the reference: not provided.  The
// entry points exist so that bindings resolving every symbol at load time load this
// library; each records PVFMM_ERR_UNSUPPORTED and returns NULL / 0 / nothing.
static void volume_unsupported() { set_error(kErrUnsupported, "the volume (Chebyshev) FMM is not provided by this library"); }
void* PVFMMCreateVolumeFMMD(int, int, enum PVFMMKernel, pvfmm_comm_t) { volume_unsupported(); return nullptr; }
void* PVFMMCreateVolumeFMMF(int, int, enum PVFMMKernel, pvfmm_comm_t) { volume_unsupported(); return nullptr; }
void* PVFMMCreateVolumeTreeD(int, int, void (*)(const double*, long, double*, const void*), const void*, const double*, long,
                             pvfmm_comm_t, double, int, bool, int) {
  volume_unsupported();
  return nullptr;
}
void* PVFMMCreateVolumeTreeF(int, int, void (*)(const float*, long, float*, const void*), const void*, const float*, long,
                             pvfmm_comm_t, float, int, bool, int) {
  volume_unsupported();
  return nullptr;
}
void* PVFMMCreateVolumeTreeFromCoeffD(long, int, int, const double*, const double*, const double*, long, pvfmm_comm_t, bool) {
  volume_unsupported();
  return nullptr;
}
void* PVFMMCreateVolumeTreeFromCoeffF(long, int, int, const float*, const float*, const float*, long, pvfmm_comm_t, bool) {
  volume_unsupported();
  return nullptr;
}
void PVFMMEvaluateVolumeFMMD(double*, void*, const void*, long) { volume_unsupported(); }
void PVFMMEvaluateVolumeFMMF(float*, void*, const void*, long) { volume_unsupported(); }
void PVFMMDestroyVolumeFMMD(void** ctx) { if (ctx) *ctx = nullptr; }
void PVFMMDestroyVolumeFMMF(void** ctx) { if (ctx) *ctx = nullptr; }
void PVFMMDestroyVolumeTreeD(void** ctx) { if (ctx) *ctx = nullptr; }
void PVFMMDestroyVolumeTreeF(void** ctx) { if (ctx) *ctx = nullptr; }
long PVFMMGetLeafCountD(const void*) { volume_unsupported(); return 0; }
long PVFMMGetLeafCountF(const void*) { volume_unsupported(); return 0; }
void PVFMMGetLeafCoordD(double*, const void*) { volume_unsupported(); }
void PVFMMGetLeafCoordF(float*, const void*) { volume_unsupported(); }
void PVFMMGetPotentialCoeffD(double*, const void*) { volume_unsupported(); }
void PVFMMGetPotentialCoeffF(float*, const void*) { volume_unsupported(); }
void PVFMMCoeff2NodesD(double*, long, int, int, const double*) { volume_unsupported(); }
void PVFMMCoeff2NodesF(float*, long, int, int, const float*) { volume_unsupported(); }
void PVFMMNodes2CoeffD(double*, long, int, int, const double*) { volume_unsupported(); }
void PVFMMNodes2CoeffF(float*, long, int, int, const float*) { volume_unsupported(); }

}  // extern "C"
