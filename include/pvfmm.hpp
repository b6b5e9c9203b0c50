// C++ particle-FMM interface of the MI355X library, with the names and call sequence of the
// reference's include/pvfmm.hpp:199-313 (dmalhotra/pvfmm):
//
//   const pvfmm::Kernel<double>& ker = pvfmm::LaplaceKernel<double>::gradient();
//   auto* tree = pvfmm::PtFMM_CreateTree(sl_coord, sl_den, trg_coord, comm, max_pts, pvfmm::FreeSpace);
//   pvfmm::PtFMM<double> mat;
//   mat.Initialize(m, comm, &ker);
//   tree->SetupFMM(&mat);
//   pvfmm::PtFMM_Evaluate(tree, trg_val, n_trg);
//   tree->ClearFMMData();
//   pvfmm::PtFMM_Evaluate(tree, trg_val, n_trg, &new_sl_den);
//   delete tree;
//
// Header-only over the C ABI of include/pvfmm.h (link with -lpvfmm).  The work runs on one
// GPU as the communicator's single rank, so `comm` may be any value (MPI_COMM_WORLD,
// MPI_COMM_SELF or 0) and is not used.  Differences from the reference, each reported by
// throwing pvfmm::Error rather than aborting: double-layer sources, init_depth > 0 and kernels
// other than laplace potential / gradient and stokes velocity are rejected.  Periodic trees
// (PX, PXY, PXYZ = Periodic, unit period) are supported.  Real = float evaluates in the mixed
// FP32/FP64 mode of the library.
#ifndef PVFMM_AMD_PVFMM_HPP
#define PVFMM_AMD_PVFMM_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pvfmm.h"

namespace pvfmm {

// include/mpi_tree.hpp:22-29: boundary condition of the tree
enum BoundaryType { FreeSpace = 0, PX, PXY, PXYZ, Periodic = PXYZ };

// Failure of a library call: code is a PVFMMError value.
class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check(int rc) {
  if (rc != PVFMM_OK) {
    const char* m = PVFMMGetLastErrorMessage();
    throw Error(rc, std::string(PVFMMErrorString(rc)) + (m && *m ? std::string(": ") + m : std::string()));
  }
}
inline void check_last() { check(PVFMMGetLastError()); }
template <class Real>
std::vector<double> to_double(const std::vector<Real>& v) {
  return std::vector<double>(v.begin(), v.end());
}
}  // namespace detail

namespace mem {
// The reference's optional memory manager (include/mem_mgr.hpp); device memory here is owned
// by the tree plan, so this only keeps call sites source-compatible.
class MemoryManager {
 public:
  explicit MemoryManager(size_t = 0) {}
};
}  // namespace mem

// include/kernel.hpp:27-136: kernel descriptor.  ker_dim = {k0, k1} (source and target
// dimensions per point); ker_poten is the direct sum used by the reference's examples.
template <class T>
struct Kernel {
  int ker_dim[2];
  std::string ker_name;
  int id;  // PVFMMKernel value

  // v_trg[k1*trg_cnt] += K(r_trg, r_src) v_src (include/kernel.hpp:40-41; dof must be 1),
  // evaluated as an FP64 all-pairs sum on the GPU.
  void ker_poten(const T* r_src, int src_cnt, const T* v_src, int dof, const T* r_trg, int trg_cnt, T* v_trg,
                 mem::MemoryManager* = nullptr) const {
    if (dof != 1) throw Error(PVFMM_ERR_UNSUPPORTED, "ker_poten supports dof == 1 only");
    std::vector<double> s(r_src, r_src + 3 * (size_t)src_cnt), d(v_src, v_src + (size_t)ker_dim[0] * src_cnt),
        t(r_trg, r_trg + 3 * (size_t)trg_cnt), o((size_t)ker_dim[1] * trg_cnt);
    detail::check(pvfmm_amd_direct(id, s.data(), src_cnt, d.data(), t.data(), trg_cnt, o.data()));
    for (size_t i = 0; i < o.size(); i++) v_trg[i] += (T)o[i];
  }
};

// include/kernel.txx:1333-1344
template <class T>
struct LaplaceKernel {
  static const Kernel<T>& potential() {
    static const Kernel<T> k{{1, 1}, "laplace", PVFMMLaplacePotential};
    return k;
  }
  static const Kernel<T>& gradient() {
    static const Kernel<T> k{{1, 3}, "laplace_grad", PVFMMLaplaceGradient};
    return k;
  }
};

// include/kernel.txx:1469-1473
template <class T>
struct StokesKernel {
  static const Kernel<T>& velocity() {
    static const Kernel<T> k{{3, 3}, "stokes_vel", PVFMMStokesVelocity};
    return k;
  }
};

// FMM_Pts<...> (include/fmm_pts.hpp:131): the operators of one kernel and multipole order.
template <class Real>
class PtFMM {
 public:
  explicit PtFMM(mem::MemoryManager* = nullptr) {}
  PtFMM(const PtFMM&) = delete;
  PtFMM& operator=(const PtFMM&) = delete;
  ~PtFMM() { pvfmm_amd_fmm_destroy(h_); }

  // FMM_Pts::Initialize(mult_order, comm, kernel) (include/fmm_pts.txx:223-331)
  template <class Comm>
  void Initialize(int mult_order, Comm, const Kernel<Real>* kernel) {
    if (!kernel) throw Error(PVFMM_ERR_INVALID_ARGUMENT, "null kernel");
    pvfmm_amd_fmm_destroy(h_);
    h_ = pvfmm_amd_fmm_create(kernel->id, mult_order, precision(), 0.0);
    if (!h_) detail::check_last();
    kernel_ = kernel;
    m_ = mult_order;
  }
  // Library extension: FP32 bulk evaluation with FP64 operators and merge for Real = double
  // (takes effect at the next Initialize).
  void SetMixedPrecision(bool on) { mixed_ = on; }

  void* handle() const { return h_; }
  const Kernel<Real>* kernel() const { return kernel_; }
  int MultipoleOrder() const { return m_; }

 private:
  int precision() const { return (sizeof(Real) == 4 || mixed_) ? 1 : 0; }
  void* h_ = nullptr;
  const Kernel<Real>* kernel_ = nullptr;
  int m_ = 0;
  bool mixed_ = false;
};

// FMM_Tree<PtFMM<Real>> (include/fmm_tree.hpp): the octree, lists and device plan.
template <class Real>
class PtFMM_Tree {
 public:
  PtFMM_Tree(const std::vector<Real>& sl_coord, const std::vector<Real>& sl_density, const std::vector<Real>& trg_coord,
             int max_pts, BoundaryType bndry = FreeSpace)
      : n_src_((int64_t)(sl_coord.size() / 3)), n_trg_((int64_t)(trg_coord.size() / 3)), sl_density_(sl_density) {
    if (sl_coord.size() % 3 || trg_coord.size() % 3) throw Error(PVFMM_ERR_INVALID_ARGUMENT, "coordinates are not [x y z]*n");
    const std::vector<double> s = detail::to_double(sl_coord), t = detail::to_double(trg_coord);
    h_ = pvfmm_amd_tree_create_bndry(s.data(), n_src_, t.data(), n_trg_, nullptr, 0, max_pts, (int)bndry);
    if (!h_) detail::check_last();
  }
  PtFMM_Tree(const PtFMM_Tree&) = delete;
  PtFMM_Tree& operator=(const PtFMM_Tree&) = delete;
  ~PtFMM_Tree() { pvfmm_amd_tree_destroy(h_); }

  // FMM_Tree::SetupFMM (include/fmm_tree.txx:72-183): builds the device plan for fmm's kernel.
  void SetupFMM(PtFMM<Real>* fmm) {
    if (!fmm || !fmm->handle()) throw Error(PVFMM_ERR_STATE, "SetupFMM needs an initialized PtFMM");
    detail::check(pvfmm_amd_tree_setup(h_, fmm->handle(), 1));
    kernel_ = fmm->kernel();
  }
  // FMM_Tree::ClearFMMData: the device plan zeroes its expansions at every evaluation, so this
  // only exists for source compatibility.
  void ClearFMMData() {}

  void* handle() const { return h_; }
  int64_t NumSources() const { return n_src_; }
  int64_t NumTargets() const { return n_trg_; }
  const Kernel<Real>* kernel() const { return kernel_; }
  const std::vector<Real>& sl_density() const { return sl_density_; }

 private:
  void* h_ = nullptr;
  int64_t n_src_ = 0, n_trg_ = 0;
  std::vector<Real> sl_density_;
  const Kernel<Real>* kernel_ = nullptr;
};

// PtFMM_CreateTree (include/pvfmm.hpp:262-265, include/pvfmm.txx:215-245).  Coordinates in
// [0,1)^3; the caller deletes the tree.  bndry: FreeSpace or a periodic type (PX, PXY, PXYZ).
template <class Real, class Comm>
PtFMM_Tree<Real>* PtFMM_CreateTree(const std::vector<Real>& sl_coord, const std::vector<Real>& sl_density,
                                   const std::vector<Real>& trg_coord, Comm, int max_pts = 100, BoundaryType bndry = FreeSpace,
                                   int init_depth = 0) {
  if (init_depth != 0) throw Error(PVFMM_ERR_UNSUPPORTED, "init_depth > 0 is not supported");
  return new PtFMM_Tree<Real>(sl_coord, sl_density, trg_coord, max_pts, bndry);
}

// Variant with double-layer sources (include/pvfmm.hpp:262-265); they must be empty here.
template <class Real, class Comm>
PtFMM_Tree<Real>* PtFMM_CreateTree(const std::vector<Real>& sl_coord, const std::vector<Real>& sl_density,
                                   const std::vector<Real>& dl_coord, const std::vector<Real>& dl_density,
                                   const std::vector<Real>& trg_coord, Comm comm, int max_pts = 100,
                                   BoundaryType bndry = FreeSpace, int init_depth = 0) {
  if (!dl_coord.empty() || !dl_density.empty()) throw Error(PVFMM_ERR_UNSUPPORTED, "double-layer sources are not supported");
  return PtFMM_CreateTree(sl_coord, sl_density, trg_coord, comm, max_pts, bndry, init_depth);
}

// PtFMM_Evaluate (include/pvfmm.hpp:311, include/pvfmm.txx:256-328): trg_val is resized to
// k1 * n_trg and filled in input order; sl_density (k0 * n_src, input order) replaces the
// tree's densities when given.
template <class Real>
void PtFMM_Evaluate(const PtFMM_Tree<Real>* tree, std::vector<Real>& trg_val, size_t loc_size = 0,
                    const std::vector<Real>* sl_density = nullptr, const std::vector<Real>* dl_density = nullptr) {
  if (!tree || !tree->kernel()) throw Error(PVFMM_ERR_STATE, "PtFMM_Evaluate needs a tree after SetupFMM");
  if (dl_density && !dl_density->empty()) throw Error(PVFMM_ERR_UNSUPPORTED, "double-layer sources are not supported");
  if (loc_size != 0 && (int64_t)loc_size != tree->NumTargets())
    throw Error(PVFMM_ERR_INVALID_ARGUMENT, "loc_size must be 0 or the number of targets");
  const int k0 = tree->kernel()->ker_dim[0], k1 = tree->kernel()->ker_dim[1];
  const std::vector<Real>& den = sl_density ? *sl_density : tree->sl_density();
  if ((int64_t)den.size() != k0 * tree->NumSources())
    throw Error(PVFMM_ERR_INVALID_ARGUMENT, "sl_density must hold k0 values per source");
  const std::vector<double> d = detail::to_double(den);
  std::vector<double> out((size_t)k1 * tree->NumTargets());
  detail::check(pvfmm_amd_tree_evaluate(tree->handle(), d.data(), out.data(), 1));
  trg_val.assign(out.begin(), out.end());
}

}  // namespace pvfmm

#endif  // PVFMM_AMD_PVFMM_HPP
