// The reference's example1 flow (examples/src/example1.cpp:59-100) through include/pvfmm.hpp:
// tree, operators, setup, evaluate, re-evaluate with new densities, error against the direct
// sum of Kernel::ker_poten, and a periodic tree checked by translation.  Without a GPU it checks
// the host-side steps and that SetupFMM reports the missing device as a pvfmm::Error.  Exit
// status 0 = pass.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "pvfmm.hpp"

template <class Real>
static double rel_err(const std::vector<Real>& a, const std::vector<Real>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); i++) {
    num += (double)(a[i] - b[i]) * (a[i] - b[i]);
    den += (double)b[i] * b[i];
  }
  return std::sqrt(num / den);
}

template <class Real>
static int run(const pvfmm::Kernel<Real>& ker, int m, double tol, bool gpu) {
  const int n = 4000, comm = 0;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<Real> src(3 * n), trg(3 * n), den(ker.ker_dim[0] * n), den2(ker.ker_dim[0] * n);
  for (auto& v : src) v = (Real)U(rng);
  for (auto& v : trg) v = (Real)U(rng);
  for (auto& v : den) v = (Real)(2 * U(rng) - 1);
  for (auto& v : den2) v = (Real)(2 * U(rng) - 1);
  auto* tree = pvfmm::PtFMM_CreateTree(src, den, trg, comm, 64, pvfmm::FreeSpace);
  pvfmm::PtFMM<Real> mat;
  mat.Initialize(m, comm, &ker);
  if (!gpu) {
    try {
      tree->SetupFMM(&mat);
      std::printf("SetupFMM without a GPU did not fail\n");
      return 1;
    } catch (const pvfmm::Error& e) {
      if (e.code() != PVFMM_ERR_DEVICE) { std::printf("unexpected error %d: %s\n", e.code(), e.what()); return 1; }
    }
    delete tree;
    return 0;
  }
  tree->SetupFMM(&mat);
  std::vector<Real> val, dir(ker.ker_dim[1] * n, 0);
  pvfmm::PtFMM_Evaluate(tree, val, n);
  ker.ker_poten(src.data(), n, den.data(), 1, trg.data(), n, dir.data());
  const double e1 = rel_err(val, dir);
  tree->ClearFMMData();
  pvfmm::PtFMM_Evaluate(tree, val, n, &den2);
  std::fill(dir.begin(), dir.end(), Real(0));
  ker.ker_poten(src.data(), n, den2.data(), 1, trg.data(), n, dir.data());
  const double e2 = rel_err(val, dir);
  std::printf("%s m=%d: rel-L2 %.3e, after new densities %.3e (tolerance %.0e)\n", ker.ker_name.c_str(), m, e1, e2, tol);
  delete tree;
  return (e1 < tol && e2 < tol) ? 0 : 1;
}

// Periodic tree (PXYZ): the field of a neutral charge set moves with the points, so translating
// every point by the same vector (mod 1) leaves the field at each point unchanged.
static int run_periodic(bool gpu) {
  if (!gpu) return 0;
  const int n = 3000, comm = 0;
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> x(3 * n), xs(3 * n), q(n);
  double mean = 0;
  for (auto& v : x) v = U(rng);
  for (auto& v : q) { v = 2 * U(rng) - 1; mean += v / n; }
  for (auto& v : q) v -= mean;
  const double shift[3] = {0.25, 0.5, 0.125};
  for (int i = 0; i < 3 * n; i++) xs[i] = std::fmod(x[i] + shift[i % 3], 1.0);
  const pvfmm::Kernel<double>& ker = pvfmm::LaplaceKernel<double>::gradient();
  pvfmm::PtFMM<double> mat;
  mat.Initialize(8, comm, &ker);
  std::vector<double> g1, g2;
  for (int pass = 0; pass < 2; pass++) {
    auto* tree = pvfmm::PtFMM_CreateTree(pass ? xs : x, q, pass ? xs : x, comm, 40, pvfmm::Periodic);
    tree->SetupFMM(&mat);
    pvfmm::PtFMM_Evaluate(tree, pass ? g2 : g1, n);
    delete tree;
  }
  const double e = rel_err(g2, g1);
  std::printf("periodic laplace_grad m=8: translated field rel-L2 %.3e (tolerance 1e-05)\n", e);
  return e < 1e-5 ? 0 : 1;
}

int main() {
  const bool gpu = pvfmm_amd_device_count() > 0;
  int bad = 0;
  bad += run(pvfmm::LaplaceKernel<double>::potential(), 8, 1e-5, gpu);
  bad += run(pvfmm::LaplaceKernel<double>::gradient(), 8, 1e-4, gpu);
  bad += run(pvfmm::StokesKernel<double>::velocity(), 8, 1e-4, gpu);
  bad += run(pvfmm::LaplaceKernel<float>::gradient(), 8, 1e-4, gpu);
  bad += run_periodic(gpu);
  // rejected configurations throw instead of aborting
  std::vector<double> x(30, 0.5), q(10, 1.0), none;
  try {
    pvfmm::PtFMM_CreateTree(x, q, x, q, x, 0, 64);
    bad++;
  } catch (const pvfmm::Error& e) {
    bad += e.code() != PVFMM_ERR_UNSUPPORTED;
  }
  std::printf("api_example: %s (gpu=%d)\n", bad ? "FAIL" : "OK", (int)gpu);
  return bad ? 1 : 0;
}
