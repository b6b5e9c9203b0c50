// Builds the per-tree device plan and runs one evaluation (the reference's RunFMM,
// include/fmm_tree.txx:237-738, with the list setups of include/fmm_pts.txx:2275-5737
// expressed as flat index arrays).
#include "device_fmm.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>

#include "kernels.hpp"

namespace pvfmm_amd {

void throw_device_error(hipError_t e, const char* file, int line) {
  throw Error(e == hipErrorOutOfMemory ? kErrOutOfMemory : kErrDevice,
              std::string("HIP error '") + hipGetErrorString(e) + "' at " + file + ":" + std::to_string(line));
}

static std::atomic<uint64_t> g_plan_generation{0};
DevicePlanBase::DevicePlanBase() : generation(++g_plan_generation) {}

const char* phase_name(int p) {
  static const char* names[] = {"Gather", "S2U-P2P", "S2U-GEMM", "U2U", "X-List", "V-List:FFT", "V-List:Hadamard",
                                "V-List:IFFT", "D2D", "D2T-GEMM", "D2T+W-List", "U-List", "Scatter"};
  return (p >= 0 && p < kPhaseCount) ? names[p] : "?";
}

namespace {

template <typename T>
struct DBuf {
  T* p = nullptr;
  int64_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { if (p) (void)hipFree(p); }
  void alloc(int64_t count, int64_t* total) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = count;
    if (count > 0) {
      HIP_CHECK(hipMalloc(&p, sizeof(T) * count));
      if (total) *total += sizeof(T) * count;
    }
  }
  void upload(const std::vector<T>& v, int64_t* total) {
    alloc((int64_t)v.size(), total);
    if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  }
};

// U2U and D2D through the rank of the check-to-equivalent factors.  The composite translations
// are u2u = T uc2ue1 (T: 8 K0 x r, uc2ue1: r x K0 with orthonormal rows) and d2d = dc2de0 B
// (dc2de0: K1 x r with orthogonal columns of norms 1/s_j, B: r x 8 K1), so T = u2u uc2ue1^T and
// B = diag(s^2) dc2de0^T d2d (include/fmm_pts.txx:582-643 forms the same products).  Applied as
// two GEMMs they cost 9 K r instead of 8 K^2 multiply-adds per row; the results are returned in
// the padded layouts of the plan (r -> rp columns of T, rows of B, zero-filled).
static std::vector<double> u2u_rank_factor(const Operators& O, int rp) {
  const int K0 = O.K0(), r = O.rank_u;
  std::vector<double> T((size_t)8 * K0 * rp, 0.0);
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < 8 * K0; i++)
    for (int j = 0; j < r; j++) {
      double acc = 0;
      for (int k = 0; k < K0; k++) acc += O.u2u[(size_t)i * K0 + k] * O.uc2ue1[(size_t)j * K0 + k];
      T[(size_t)i * rp + j] = acc;
    }
  return T;
}
static std::vector<double> d2d_rank_factor(const Operators& O, int rp) {
  const int K1 = O.K1(), r = O.rank_d;
  std::vector<double> B((size_t)rp * 8 * K1, 0.0);
  #pragma omp parallel for schedule(static)
  for (int j = 0; j < r; j++) {
    double n2 = 0;
    for (int i = 0; i < K1; i++) n2 += O.dc2de0[(size_t)i * r + j] * O.dc2de0[(size_t)i * r + j];
    double* b = &B[(size_t)j * 8 * K1];
    for (int i = 0; i < K1; i++) {
      const double c = O.dc2de0[(size_t)i * r + j] / n2;
      const double* d = &O.d2d[(size_t)i * 8 * K1];
      for (int n = 0; n < 8 * K1; n++) b[n] += c * d[n];
    }
  }
  return B;
}

// LDS pool slots of the V-list tiles' staged partners for the bf16 kernel (FP32 scalar
// kernels), which keeps one pool of `pool` staged entries across the run of tiles a block
// handles: a partner of tile k that tile k-1 also lists keeps its slot, and only the others
// (the delta) are read and staged, into slots that no partner of tile k-1 holds, while tile
// k-1 is computed.  A block's first tile stages its whole union into the same slots.
// Among the free slots, a new partner takes one by greedy colouring, most-read partners
// first: the 16 lanes that read one staged entry each (the 16 targets of a group at one
// offset) are served together, and slot s lands in LDS bank group 7s mod 16, so the slots of
// the partners a group reads at one offset should differ mod 16.  Output: kslot[u] the slot of
// union entry u, and per tile k the delta dent[d0[k] .. d0[k+1]) as (source slot << 9 | slot).
static void assign_pool_slots(int n_off, int pool, int n_s, const std::vector<int32_t>& t0, const std::vector<int32_t>& u0,
                              const std::vector<int32_t>& usrc, const std::vector<int32_t>& partner,
                              std::vector<int32_t>& kslot, std::vector<int32_t>& d0, std::vector<int32_t>& dent) {
  constexpr int G = 16;
  const int n_tiles = (int)t0.size() - 1;
  kslot.assign(usrc.size(), -1);
  d0.assign(1, 0);
  dent.clear();
  std::vector<int32_t> slot_of(std::max(1, n_s), -1);  // slots of the partners of the last tile
  std::vector<int32_t> loc(std::max(1, n_s), -1);      // local index in the current tile
  std::vector<char> busy(pool);
  for (int k = 0; k < n_tiles; k++) {
    const int nu = u0[k + 1] - u0[k];
    for (int u = u0[k]; u < u0[k + 1]; u++) loc[usrc[u]] = u - u0[k];
    std::fill(busy.begin(), busy.end(), 0);
    if (k > 0)
      for (int u = u0[k - 1]; u < u0[k]; u++) {
        const int32_t sl = usrc[u];
        busy[slot_of[sl]] = 1;
        if (loc[sl] < 0) slot_of[sl] = -1;  // dropped: a later listing starts a new run
      }
    // groups of local indices read together, as (group, offset) of the tile
    std::vector<std::vector<int>> grp, memb(nu);
    for (int g = 0; g < 4; g++)
      for (int o = 0; o < n_off; o++) {
        std::vector<int> m;
        for (int r = 0; r < 16; r++) {
          const int t = t0[k] + 16 * g + r;
          if (t >= t0[k + 1]) break;
          const int32_t sl = partner[(size_t)t * n_off + o];
          if (sl >= 0) m.push_back(loc[sl]);
        }
        std::sort(m.begin(), m.end());
        m.erase(std::unique(m.begin(), m.end()), m.end());
        if (m.size() < 2) continue;
        for (int li : m) memb[li].push_back((int)grp.size());
        grp.push_back(std::move(m));
      }
    std::vector<int> col(nu, -1), order;
    for (int li = 0; li < nu; li++) {
      const int32_t sl = usrc[u0[k] + li];
      if (slot_of[sl] >= 0) col[li] = slot_of[sl] % G;  // kept
      else order.push_back(li);
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return memb[x].size() > memb[y].size(); });
    for (int li : order) {
      int bad[G] = {0};
      for (int gi : memb[li])
        for (int lj : grp[gi])
          if (col[lj] >= 0) bad[col[lj]]++;
      int best = -1, best_slot = -1;
      for (int c = 0; c < G; c++) {
        int s = c;
        while (s < pool && busy[s]) s += G;
        if (s >= pool) continue;
        if (best < 0 || bad[c] < bad[best] || (bad[c] == bad[best] && s < best_slot)) { best = c; best_slot = s; }
      }
      if (best < 0) throw Error(kErrInternal, "V-list tiles exceed the staging pool");
      const int32_t sl = usrc[u0[k] + li];
      busy[best_slot] = 1;
      slot_of[sl] = best_slot;
      col[li] = best;
      dent.push_back((sl << 9) | best_slot);
    }
    for (int u = u0[k]; u < u0[k + 1]; u++) {
      kslot[u] = slot_of[usrc[u]];
      loc[usrc[u]] = -1;
    }
    d0.push_back((int32_t)dent.size());
  }
  // replay the staging as the kernel does it: each tile's delta goes to slots its predecessor
  // does not read, after which every slot of the tile's table holds the listed source
  std::vector<int32_t> held(pool, -1);
  for (int k = 0; k < n_tiles; k++) {
    if (k > 0) std::fill(busy.begin(), busy.end(), 0);
    if (k > 0)
      for (int u = u0[k - 1]; u < u0[k]; u++) busy[kslot[u]] = 1;
    for (int d = d0[k]; d < d0[k + 1]; d++) {
      const int s = dent[d] & 511;
      if (s >= pool || (k > 0 && busy[s])) throw Error(kErrInternal, "V-list pool slot overwritten while in use");
      held[s] = dent[d] >> 9;
    }
    for (int u = u0[k]; u < u0[k + 1]; u++)
      if (kslot[u] < 0 || kslot[u] >= pool || held[kslot[u]] != usrc[u])
        throw Error(kErrInternal, "V-list pool slot does not hold its partner");
  }
}

template <typename Real>
class DevicePlan final : public DevicePlanBase {
 public:
  DevicePlan(const Octree& T, const Operators& O, const double* src_xyz, const double* trg_xyz, int n_vec, hipStream_t s)
      : O_(O), st_(s) {
    n_vec_max = std::max(1, n_vec);
    build(T, src_xyz, trg_xyz);
  }

  hipStream_t stream() const override { return st_; }
  bool is_fp32() const override { return sizeof(Real) == 4; }

  void evaluate(const double* d_den, double* d_out, int n_vec, double den_scale, double out_scale) override {
    if (n_vec < 1) throw Error(kErrInvalidArgument, "n_vec must be >= 1");
    // passes of up to n_vec_max vectors: they share the pair geometry of every P2P and run as
    // the rows of one GEMM per up/down step
    for (int v = 0; v < n_vec; v += nvb_)
      evaluate_one(d_den + (int64_t)v * k0_ * n_src_, d_out + (int64_t)v * k1_ * n_trg_, den_scale, out_scale,
                   std::min(nvb_, n_vec - v));
  }

  Tuning autotune(const double* d_den, double* d_out, int n_vec) override {
    if (n_vec < 1 || n_vec > nvb_) throw Error(kErrInvalidArgument, "autotune takes 1 to n_vec_max density vectors");
    const bool was = timing_;
    timing_ = true;
    // fastest per-phase times over a few evaluations with the current tuning
    auto measure = [&]() {
      std::vector<double> best(kPhaseCount, 1e300);
      for (int r = 0; r < 3; r++) {
        evaluate_one(d_den, d_out, 1.0, 1.0, n_vec);
        for (int p = 0; p < kPhaseCount; p++) best[p] = std::min(best[p], phase_ms[p]);
      }
      return best;
    };
    auto pick = [&](int& knob, std::initializer_list<int> cands, std::initializer_list<int> phases) {
      int best_v = knob;
      double best_t = 1e300;
      for (int v : cands) {
        knob = v;
        const std::vector<double> ph = measure();
        double t = 0;
        for (int p : phases) t += ph[p];
        if (t < best_t) { best_t = t; best_v = v; }
      }
      knob = best_v;
    };
    try {
      if (sizeof(Real) == 4 && e0_ == 1 && vt_.n > 0) {
        // with 8 or more vectors: the 16-vector kernel (targets per wave) or, for 0, one vector
        // at a time through the tiled kernel, compared over the whole V list (their spectra
        // layouts differ)
        if (vec16_ && n_vec >= 8) pick(tuning.hadamard_vec_tpw, {0, 4, 8, 16, 32}, {kPhFFT, kPhHadamard, kPhIFFT});
        if (!use_vec16(n_vec)) pick(tuning.hadamard_tpb, {4, 8, 16, 32, 64}, {kPhHadamard});
      }
      pick(tuning.gemm_split_blocks, {0, 256, 512, 1024, 2048}, {kPhS2UGemm, kPhU2U, kPhD2D, kPhD2TGemm});
    } catch (...) {
      timing_ = was;
      throw;
    }
    timing_ = was;
    return tuning;
  }

 private:
  using K = DeviceKernels<Real>;
  const Operators& O_;
  hipStream_t st_;
  int64_t n_src_ = 0, n_trg_ = 0;
  int k0_ = 1, k1_ = 1, e0_ = 1, e1_ = 1, K0_ = 0, K1_ = 0, ns_ = 0, m_ = 0, M_dim_ = 0;
  int n_nodes_ = 0, max_depth_ = 0;
  int bndry_ = 0;     // boundary type of the tree (0 free space, 1-3 periodic)
  bool vol_ = false;  // PXYZ with a kernel carrying the analytic volume potential
  DBuf<Real> bc_;     // periodic boundary operator (K0_ x K1_)
  DBuf<double> txa_;  // target coordinates in sorted order (SoA), for the volume-potential correction

  // tree arrays (periodic trees append virtual image nodes: a source node's point range and size
  // with its centre moved by whole periods)
  DBuf<double> center_;
  DBuf<float> node_size_;
  DBuf<int32_t> src_begin_, src_end_, trg_begin_, trg_end_;
  DBuf<int64_t> src_perm_, trg_perm_;
  DBuf<Real> sx_, tx_, sden_;
  DBuf<double> sx64_, tx64_;  // FP64 leaf-local coordinates for the U list
  DBuf<double> tval_;
  // expansions: eq_ = [up (n_nodes x K0) | L2T equivalent rows (n_l2t x K0)]
  DBuf<Real> eq_, dn_, chk_;
  // operators
  DBuf<Real> uc2ue0_, uc2ue1_, dc2de0_, dc2de1_, u2u_, d2d_, V_, unit_surf_, ctmp_;
  // U2U / D2D through the check-to-equivalent rank (u2u_rank_factor, d2d_rank_factor) when
  // 9 r < 8 K, at levels of at least kFactorRows parents (a smaller level costs less as one
  // GEMM than as two launches): u2ur_ 8 K0 x ru_, d2dr_ rd_ x 8 K1; lvl_seq_ = 0, 1, ... (rows
  // of ctmp_)
  static constexpr int kFactorRows = 256;
  bool u2u_fact_ = false, d2d_fact_ = false;
  DBuf<Real> u2ur_, d2dr_;
  DBuf<int32_t> lvl_seq_;
  DBuf<uint16_t> Vb_;  // transfer blocks as bf16 parts (FP32 scalar kernels, 16-vector passes)
  // the fp16 form of the FP32 scalar Hadamard: transfer parts, 1 / tau_f, the spectrum bound of
  // each slot of X (or, in the adjoint, of Y) and the scale of each run of tiles
  DBuf<uint16_t> Vh_;
  DBuf<float> vtinv_, vbound_, vrscale_;
  bool split_ = false;
  DBuf<Real> gemm_work_;  // split-K partial tiles of the small coarse-level GEMMs
  int64_t gemm_work_elems_ = 0;
  int nvb_ = 1;  // density vectors per evaluation pass (buffers hold this many)
  int64_t eq_vs_ = 0, dn_vs_ = 0, chk_vs_ = 0, ctmp_vs_ = 0;  // per-vector strides  // room for 4 K-split partials of the largest GEMM output, at most 2^25
  int ru_ = 0, rd_ = 0;
  DBuf<int32_t> lattice_, lattice_inv_;
  // S2M
  DBuf<int32_t> s2m_node_, s2m_ptr_, s2m_items_, s2m_orow_;
  DBuf<Real> s2m_alpha_;
  int n_s2m_ = 0;
  // M2M / L2L per level
  struct LevelRows { int n = 0; DBuf<int32_t> ra, rc; };
  std::vector<std::unique_ptr<LevelRows>> m2m_, l2l_;
  // X list
  DBuf<int32_t> x_node_, x_ptr_, x_items_, x_orow_;
  int n_x_ = 0;
  // V list
  DBuf<int32_t> v_src_fc_, v_trg_fc_, v_partner_, v_offsets_;
  DBuf<Real> v_src_scal_, v_trg_scal_, X_, Y_;
  // spectra of a group of 16 density vectors for the 16-vector Hadamard (FP32 scalar kernels,
  // plans set up for 8 or more vectors): [slot][f][vector 16][16 reals]
  DBuf<Real> Xv_, Yv_;
  bool vec16_ = false;
  bool use_vec16(int nv) const { return vec16_ && nv >= 8 && tuning.hadamard_vec_tpw > 0; }
  int n_vsrc_ = 0, n_vtrg_ = 0, n_off_ = 0;

  // V-list target tiles for the LDS-staged Hadamard
  struct TileSet {
    DBuf<int32_t> t0, u0, usrc;
    DBuf<uint16_t> uidx;
    DBuf<uint32_t> gmask;
    DBuf<int32_t> d0, dent;  // FP32 scalar kernel: per tile, the partners new since the previous tile
    int n = 0;
    int64_t partners = 0, max_partners = 0, staged_delta = 0;
  };
  TileSet vt_;
  // operands of the fp16 Hadamard with run scales in `rs` (kept until the launch returns)
  HadamardParts hp_;
  const HadamardParts& hparts(DBuf<float>& rs) {
    hp_.Vh = Vh_.p;
    hp_.tinv = vtinv_.p;
    hp_.bound = vbound_.p;
    hp_.run_scale = rs.p;
    hp_.run_scale_cap = (int)rs.n;
    return hp_;
  }

  // Tiles of consecutive target slots: at most 64 targets, at most umax distinct partners
  // (greedy in slot order, which is Morton order within a level).  partner[t * n_off + o] is
  // the source slot of target slot t at offset o, or -1.  For the FP32 scalar kernel (LDS pool
  // of assign_pool_slots) a tile's partners together with the previous tile's must also fit
  // the pool, and usrc holds each tile's union as (source slot << 9 | pool slot).
  void build_tiles(const std::vector<int32_t>& v_partner, int n_t, int n_s, TileSet& ts, int64_t* total) {
    const int umax = K::hadamard_union_max(e0_);
    const bool pair_layout = std::is_same<Real, float>::value && e0_ == 1 && e1_ == 1;
    const int pool = pair_layout ? K::hadamard_pool_max() : (1 << 30);
    std::vector<int32_t> t0{0}, u0{0}, usrc;
    std::vector<uint16_t> uidx;
    std::vector<int32_t> pos(std::max(1, n_s), -1);
    std::vector<char> in_prev(std::max(1, n_s), 0);  // partner of the last closed tile
    std::vector<int32_t> cur, prev;  // distinct partners of the open tile and of the last closed one
    int cur_new = 0;                 // partners of the open tile that the last closed one lacks
    auto close_tile = [&](int t_end) {
      for (int32_t sl : cur) pos[sl] = -1;
      usrc.insert(usrc.end(), cur.begin(), cur.end());
      for (int32_t sl : prev) in_prev[sl] = 0;
      prev.swap(cur);
      for (int32_t sl : prev) in_prev[sl] = 1;
      cur.clear();
      cur_new = 0;
      t0.push_back(t_end);
      u0.push_back((int32_t)usrc.size());
    };
    // first pass: tile boundaries and union lists
    for (int t = 0; t < n_t; t++) {
      int add = 0, add_new = 0;
      for (int k = 0; k < n_off_; k++) {
        const int32_t sl = v_partner[(size_t)t * n_off_ + k];
        if (sl >= 0 && pos[sl] < 0) { pos[sl] = -2; add++; add_new += in_prev[sl] ? 0 : 1; }  // provisional mark
      }
      for (int k = 0; k < n_off_; k++) {
        const int32_t sl = v_partner[(size_t)t * n_off_ + k];
        if (sl >= 0 && pos[sl] == -2) pos[sl] = -1;
      }
      if (t > t0.back() && (t - t0.back() == 64 || (int)cur.size() + add > umax ||
                            (int64_t)prev.size() + cur_new + add_new > pool))
        close_tile(t);
      for (int k = 0; k < n_off_; k++) {
        const int32_t sl = v_partner[(size_t)t * n_off_ + k];
        if (sl >= 0 && pos[sl] < 0) {
          pos[sl] = (int32_t)cur.size();
          cur.push_back(sl);
          cur_new += in_prev[sl] ? 0 : 1;
        }
      }
    }
    if (n_t > t0.back()) close_tile(n_t);
    const int n_tiles = (int)t0.size() - 1;
    // a missing partner points at the staging area's zero entry (index umax, or the pool size
    // for the FP32 scalar kernel).  The FP32 scalar kernel reads its table as [k][group]
    // [o / 16][o % 2][row][(o / 2) % 8] (pairs o / 2 padded to 16), the others as
    // [k][o][row][group].
    const size_t per_tile = pair_layout ? (size_t)4 * 16 * 2 * 16 : (size_t)64 * n_off_;
    if (pair_layout && n_off_ > 32) throw Error(kErrInternal, "V-list offsets exceed the pair table");
    if (pair_layout && (pool + 1 > 511 || (int64_t)n_s >= (int64_t(1) << 22)))
      throw Error(kErrInternal, "V-list staging pool or source slots exceed the packed entry");
    std::vector<int32_t> kslot, d0, dent;
    if (pair_layout) assign_pool_slots(n_off_, pool, n_s, t0, u0, usrc, v_partner, kslot, d0, dent);
    uidx.assign((size_t)n_tiles * per_tile, (uint16_t)(pair_layout ? pool : umax));
    std::vector<uint32_t> gmask((size_t)n_tiles * 4, 0u);  // offsets present per 16-target group
    for (int k = 0; k < n_tiles; k++) {
      for (int u = u0[k]; u < u0[k + 1]; u++) pos[usrc[u]] = pair_layout ? kslot[u] : u - u0[k];
      for (int t = t0[k]; t < t0[k + 1]; t++)
        for (int o = 0; o < n_off_; o++) {
          const int32_t sl = v_partner[(size_t)t * n_off_ + o];
          const int tt = t - t0[k];  // target tt = 16 * group + row of the MFMA tiles
          if (sl >= 0) {
            const size_t at = pair_layout ? (size_t)k * per_tile +
                                                (((((size_t)(tt >> 4) * 2 + (o >> 4)) * 2 + (o & 1)) * 16 + (tt & 15)) * 8 +
                                                 ((o >> 1) & 7))
                                          : (((size_t)k * n_off_ + o) * 16 + (tt & 15)) * 4 + (tt >> 4);
            uidx[at] = (uint16_t)pos[sl];
            gmask[(size_t)k * 4 + (tt >> 4)] |= 1u << o;
          }
        }
      for (int u = u0[k]; u < u0[k + 1]; u++) pos[usrc[u]] = -1;
      if (u0[k + 1] - u0[k] > umax || t0[k + 1] - t0[k] > 64)
        throw Error(kErrInternal, "V-list tile exceeds its staging capacity");
      ts.max_partners = std::max<int64_t>(ts.max_partners, u0[k + 1] - u0[k]);
    }
    ts.n = n_tiles;
    ts.partners = (int64_t)usrc.size();
    if (pair_layout) {
      for (size_t u = 0; u < usrc.size(); u++) usrc[u] = (usrc[u] << 9) | kslot[u];
      ts.staged_delta = (int64_t)dent.size();
      // one readable entry past each list's end (source 0 into the unread trash slot): the
      // kernel reads ahead branch-free
      dent.push_back(pool + 1);
      usrc.push_back(pool + 1);
      ts.d0.upload(d0, total);
      ts.dent.upload(dent, total);
    } else {
      ts.staged_delta = ts.partners;
      usrc.push_back(0);  // one readable slot past the end: the kernels read ahead branch-free
    }
    ts.t0.upload(t0, total); ts.u0.upload(u0, total); ts.usrc.upload(usrc, total); ts.uidx.upload(uidx, total);
    ts.gmask.upload(gmask, total);
  }
  // L2T
  DBuf<int32_t> l2t_ra_, l2t_rc_, l2t_seq_;
  DBuf<Real> l2t_alpha_;
  int n_l2t_ = 0;
  // target leaves: far field and U list
  DBuf<int32_t> t_leaf_, far_ptr_, far_node_, u_ptr_, u_items_;
  DBuf<int64_t> far_den_;
  DBuf<double> far_rad_;
  int n_tleaf_ = 0;
  std::vector<hipEvent_t> ev_;
  std::vector<hipEvent_t> vev_;  // per-vector V-list sub-phase events (timing of multi-vector passes)

 public:
  ~DevicePlan() override {
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
    for (hipEvent_t e : vev_) (void)hipEventDestroy(e);
  }

 private:
  void mark(int phase) {
    if (timing_) HIP_CHECK(hipEventRecord(ev_[phase], st_));
  }
  void ensure_vlist_events(int nv) {
    while ((int)vev_.size() < 4 * nv) {
      hipEvent_t e = nullptr;
      HIP_CHECK(hipEventCreate(&e));
      vev_.push_back(e);
    }
  }

  void build(const Octree& T, const double* src_xyz, const double* trg_xyz) {
    int64_t total = 0;
    const KernelInfo& ki = O_.ker;
    k0_ = ki.k0; k1_ = ki.k1; e0_ = ki.e0; e1_ = ki.e1;
    ns_ = O_.ns; m_ = O_.m; K0_ = O_.K0(); K1_ = O_.K1(); M_dim_ = O_.M_dim;
    n_src_ = T.src.n; n_trg_ = T.trg.n;
    n_nodes_ = T.n_nodes();
    max_depth_ = T.max_depth;
    if (n_src_ >= (1ll << 31) || n_trg_ >= (1ll << 31)) throw Error(kErrUnsupported, "more than 2^31 points");
    const int N = n_nodes_;
    bndry_ = T.bndry;

    // ---- periodic images: every (source node, shift) that a U, W or X list of a target node
    // reaches across the boundary becomes a virtual node N + j
    std::unordered_map<int64_t, int32_t> img_of;
    std::vector<std::pair<int32_t, int>> imgs;  // (source node, shift code (sx+1) + 3(sy+1) + 9(sz+1))
    auto img = [&](int32_t t, int type, int k, int32_t s) -> int32_t {
      if (bndry_ == kFreeSpace) return s;
      int sh[3];
      image_shift(T, t, s, type, k, sh);
      const int code = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
      if (code == 13) return s;
      const int64_t key = (int64_t)s * 27 + code;
      auto it = img_of.find(key);
      if (it != img_of.end()) return it->second;
      const int32_t v = N + (int32_t)imgs.size();
      img_of.emplace(key, v);
      imgs.emplace_back(s, code);
      return v;
    };
    if (bndry_ != kFreeSpace)
      for (int t = 0; t < N; t++) {
        if (T.pt_cnt1(t) == 0) continue;
        for (int type : {kU0, kU1, kU2, kW, kX}) {
          const int nl = T.list_count(type);
          for (int k = 0; k < nl; k++) {
            const int32_t s = T.lists[type][(size_t)t * nl + k];
            if (s >= 0) (void)img(t, type, k, s);
          }
        }
      }
    const int NV = N + (int)imgs.size();

    // ---- node arrays
    std::vector<double> center(3 * (size_t)NV);
    std::vector<float> nsz(NV);
    std::vector<int32_t> sb(NV), se(NV), tb(NV, 0), te(NV, 0);
    for (int i = 0; i < N; i++) {
      double c[3];
      T.corner(i, c);
      double s = T.size(i);
      for (int d = 0; d < 3; d++) center[3 * i + d] = c[d] + 0.5 * s;
      nsz[i] = (float)s;
      sb[i] = (int32_t)T.src_begin[i]; se[i] = (int32_t)T.src_end[i];
      tb[i] = (int32_t)T.trg_begin[i]; te[i] = (int32_t)T.trg_end[i];
    }
    for (size_t j = 0; j < imgs.size(); j++) {
      const int32_t s = imgs[j].first, v = N + (int32_t)j, code = imgs[j].second;
      const int sh[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
      for (int d = 0; d < 3; d++) center[3 * (size_t)v + d] = center[3 * (size_t)s + d] + sh[d];
      nsz[v] = nsz[s];
      sb[v] = sb[s]; se[v] = se[s];
    }
    center_.upload(center, &total);
    node_size_.upload(nsz, &total);
    src_begin_.upload(sb, &total); src_end_.upload(se, &total);
    trg_begin_.upload(tb, &total); trg_end_.upload(te, &total);
    h_.center = center; h_.nsz = nsz; h_.sb = sb; h_.se = se; h_.tb = tb; h_.te = te; h_.imgs = imgs;
    src_perm_.upload(T.src.perm, &total);
    trg_perm_.upload(T.trg.perm, &total);

    // ---- leaf-local coordinates (x - leaf centre), SoA
    std::vector<double> sx(3 * (size_t)n_src_), tx(3 * (size_t)n_trg_);
    for (int i = 0; i < N; i++) {
      if (!T.is_leaf(i)) continue;
      for (int64_t k = T.src_begin[i]; k < T.src_end[i]; k++)
        for (int d = 0; d < 3; d++) sx[d * n_src_ + k] = src_xyz[3 * T.src.perm[k] + d] - center[3 * i + d];
      for (int64_t k = T.trg_begin[i]; k < T.trg_end[i]; k++)
        for (int d = 0; d < 3; d++) tx[d * n_trg_ + k] = trg_xyz[3 * T.trg.perm[k] + d] - center[3 * i + d];
    }
    sx64_.upload(sx, &total);
    tx64_.upload(tx, &total);
    sx_.upload(std::vector<Real>(sx.begin(), sx.end()), &total);
    tx_.upload(std::vector<Real>(tx.begin(), tx.end()), &total);
    // per-vector buffers hold n_vec_max density vectors, one after another
    nvb_ = n_vec_max;
    sden_.alloc((int64_t)k0_ * n_src_ * nvb_, &total);
    tval_.alloc((int64_t)k1_ * n_trg_ * nvb_, &total);

    // ---- operators in the working precision
    auto conv = [](const std::vector<double>& v) { return std::vector<Real>(v.begin(), v.end()); };
    // check -> equivalent factors with the rank padded to a multiple of 16 (zero columns of
    // the first factor, zero rows of the second), so the GEMM operands stay 4-vector aligned
    ru_ = (O_.rank_u + 15) / 16 * 16;
    rd_ = (O_.rank_d + 15) / 16 * 16;
    auto pad_cols = [&](const std::vector<double>& m, int rows, int r, int rp) {
      std::vector<Real> out((size_t)rows * rp, Real(0));
      for (int i = 0; i < rows; i++)
        for (int j = 0; j < r; j++) out[(size_t)i * rp + j] = (Real)m[(size_t)i * r + j];
      return out;
    };
    auto pad_rows = [&](const std::vector<double>& m, int r, int cols, int rp) {
      std::vector<Real> out((size_t)rp * cols, Real(0));
      for (int i = 0; i < r; i++)
        for (int j = 0; j < cols; j++) out[(size_t)i * cols + j] = (Real)m[(size_t)i * cols + j];
      return out;
    };
    uc2ue0_.upload(pad_cols(O_.uc2ue0, K1_, O_.rank_u, ru_), &total);
    uc2ue1_.upload(pad_rows(O_.uc2ue1, O_.rank_u, K0_, ru_), &total);
    dc2de0_.upload(pad_cols(O_.dc2de0, K1_, O_.rank_d, rd_), &total);
    dc2de1_.upload(pad_rows(O_.dc2de1, O_.rank_d, K0_, rd_), &total);
    u2u_fact_ = 9 * ru_ < 8 * K0_;
    d2d_fact_ = 9 * rd_ < 8 * K1_;
    if (u2u_fact_) u2ur_.upload(conv(u2u_rank_factor(O_, ru_)), &total);
    if (d2d_fact_) d2dr_.upload(conv(d2d_rank_factor(O_, rd_)), &total);
    u2u_.upload(conv(O_.u2u), &total);
    d2d_.upload(conv(O_.d2d), &total);
    {
      // Transfer blocks in real form for y = x B with x, y interleaved (re, im):
      //   B[2s][2t] = Re M[t][s], B[2s+1][2t] = -Im M[t][s], B[2s][2t+1] = Im M[t][s],
      //   B[2s+1][2t+1] = Re M[t][s],
      // stored per MFMA lane: [R][f][kb][nb][lane][q] = B[16kb + 4(lane/16) + q][16nb + lane%16].
      // Spectrum rows are component-major (the DFTs write and read them so): row position k of
      // an input spectrum holds component k / 16 of child (k % 16) / 2, which is complex source
      // index s_of(k) of M (child-major: s = child * E0 + component), and output position n is
      // complex target index t_of(n).
      const int E0 = e0_, E1 = e1_, NS = 8 * E0, NR = 8 * E1;
      auto s_of = [E0](int k) { return ((k & 15) >> 1) * E0 + (k >> 4); };
      auto t_of = [E1](int n) { return ((n & 15) >> 1) * E1 + (n >> 4); };
      // the FP32 kernels use the fp16 (and, scalar, bf16) parts below instead
      const bool bf16_parts = std::is_same<Real, float>::value && E0 == 1 && E1 == 1;
      const bool stokes_parts = std::is_same<Real, float>::value && E0 == 3 && E1 == 3;
      const bool real_form = !bf16_parts && !stokes_parts;
      std::vector<Real> Vp(real_form ? (size_t)27 * M_dim_ * E0 * E1 * 256 : 0, (Real)0);
      #pragma omp parallel for schedule(static)
      for (int64_t rf = 0; rf < (real_form ? (int64_t)27 * M_dim_ : 0); rf++) {
        const std::complex<double>* Mb = &O_.v1[(size_t)rf * NR * NS];  // [t][s]
        for (int kb = 0; kb < E0; kb++)
          for (int nb = 0; nb < E1; nb++)
            for (int l = 0; l < 64; l++)
              for (int q = 0; q < 4; q++) {
                const int k = 16 * kb + 4 * (l >> 4) + q, n = 16 * nb + (l & 15);
                const std::complex<double> v = Mb[(size_t)t_of(n) * NS + s_of(k)];
                double val = ((k & 1) == 0) ? ((n & 1) == 0 ? v.real() : v.imag()) : ((n & 1) == 0 ? -v.imag() : v.real());
                Vp[((((size_t)rf * E0 + kb) * E1 + nb) * 64 + l) * 4 + q] = (Real)val;
              }
      }
      V_.upload(Vp, &total);
      if (bf16_parts) {
        // the same blocks as three bf16 parts for the matrix-core kernel, offsets taken in
        // pairs in the order of v1_offsets: [pair j][part p][f][lane][i] = part p of
        // B_{2j + lane/32}[8 ((lane/16) % 2) + i][lane % 16]; each part rounds the remainder of
        // the FP64 value to nearest
        auto bf16_rn = [](float x) {
          uint32_t u;
          std::memcpy(&u, &x, 4);
          u += 0x7FFFu + ((u >> 16) & 1u);
          return (uint16_t)(u >> 16);
        };
        auto bf16_val = [](uint16_t h) {
          const uint32_t u = (uint32_t)h << 16;
          float x;
          std::memcpy(&x, &u, 4);
          return (double)x;
        };
        const int n_pair = (int)O_.v1_offsets.size() / 2;
        std::vector<uint16_t> Vb((size_t)n_pair * 3 * M_dim_ * 64 * 8);
        #pragma omp parallel for schedule(static)
        for (int64_t jf = 0; jf < (int64_t)n_pair * M_dim_; jf++) {
          const int j = (int)(jf / M_dim_), f = (int)(jf % M_dim_);
          for (int l = 0; l < 64; l++) {
            const int R = O_.v1_offsets[2 * j + (l >> 5)];
            const std::complex<double>* Mb = &O_.v1[((size_t)R * M_dim_ + f) * NR * NS];
            for (int i = 0; i < 8; i++) {
              const int k = 8 * ((l >> 4) & 1) + i, n = l & 15;
              const std::complex<double> v = Mb[(size_t)(n >> 1) * NS + (k >> 1)];
              double r = ((k & 1) == 0) ? ((n & 1) == 0 ? v.real() : v.imag()) : ((n & 1) == 0 ? -v.imag() : v.real());
              for (int p = 0; p < 3; p++) {
                const uint16_t h = bf16_rn((float)r);
                Vb[((((size_t)j * 3 + p) * M_dim_ + f) * 64 + l) * 8 + i] = h;
                r -= bf16_val(h);
              }
            }
          }
        }
        Vb_.upload(Vb, &total);
        // two fp16 parts of tau_f B for k_hadamard_split: tau_f = 2^(14 - e) with every entry
        // of frequency f's blocks below 2^e (so below 2^14 after scaling); part 0 rounds tau_f t
        // to nearest, part 1 the remainder
        std::vector<uint16_t> Vh((size_t)n_pair * 2 * M_dim_ * 64 * 8);
        std::vector<float> tinv(M_dim_);
        auto entry = [&](int j, int f, int l, int i) {
          const int R = O_.v1_offsets[2 * j + (l >> 5)];
          const std::complex<double> v = O_.v1[((size_t)R * M_dim_ + f) * NR * NS + (size_t)((l & 15) >> 1) * NS +
                                               ((8 * ((l >> 4) & 1) + i) >> 1)];
          const int k = 8 * ((l >> 4) & 1) + i, n = l & 15;
          return ((k & 1) == 0) ? ((n & 1) == 0 ? v.real() : v.imag()) : ((n & 1) == 0 ? -v.imag() : v.real());
        };
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < M_dim_; f++) {
          double mx = 0;
          for (int j = 0; j < n_pair; j++)
            for (int l = 0; l < 64; l++)
              for (int i = 0; i < 8; i++) mx = std::max(mx, std::fabs(entry(j, f, l, i)));
          int e = 0;
          if (mx > 0) (void)std::frexp(mx, &e);
          const double tau = mx > 0 ? std::ldexp(1.0, std::min(100, std::max(-100, 14 - e))) : 1.0;
          tinv[f] = (float)(1.0 / tau);
          for (int j = 0; j < n_pair; j++)
            for (int l = 0; l < 64; l++)
              for (int i = 0; i < 8; i++) {
                const double x = entry(j, f, l, i) * tau;
                const _Float16 h0 = (_Float16)x, h1 = (_Float16)(x - (double)h0);
                std::memcpy(&Vh[((((size_t)j * 2 + 0) * M_dim_ + f) * 64 + l) * 8 + i], &h0, 2);
                std::memcpy(&Vh[((((size_t)j * 2 + 1) * M_dim_ + f) * 64 + l) * 8 + i], &h1, 2);
              }
        }
        Vh_.upload(Vh, &total);
        vtinv_.upload(tinv, &total);
        split_ = true;
      }
      if (stokes_parts) {
        // two fp16 parts of tau_f B for k_hadamard_split3, offsets in pairs (2j, 2j+1) of
        // v1_offsets: [f][pair j][chunk c][tile nb][part p][lane][i] = part p of
        // B_{2j + a}[8 b + i][16 nb + lane % 16], q = 4c + lane / 16, a = q / 6, b = q % 6;
        // tau_f = 2^(14 - e) with every entry of frequency f's blocks below 2^e
        const int n_pair = (int)O_.v1_offsets.size() / 2;
        const size_t per_f = (size_t)n_pair * 3 * 3 * 2 * 64 * 8;
        std::vector<uint16_t> Vh((size_t)M_dim_ * per_f);
        std::vector<float> tinv(M_dim_);
        auto entry = [&](int R, int f, int k, int n) {
          const std::complex<double> v = O_.v1[((size_t)R * M_dim_ + f) * NR * NS + (size_t)t_of(n) * NS + s_of(k)];
          return ((k & 1) == 0) ? ((n & 1) == 0 ? v.real() : v.imag()) : ((n & 1) == 0 ? -v.imag() : v.real());
        };
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < M_dim_; f++) {
          double mx = 0;
          for (int o = 0; o < 2 * n_pair; o++)
            for (int k = 0; k < 2 * NS; k++)
              for (int n = 0; n < 2 * NR; n++) mx = std::max(mx, std::fabs(entry(O_.v1_offsets[o], f, k, n)));
          int e = 0;
          if (mx > 0) (void)std::frexp(mx, &e);
          const double tau = mx > 0 ? std::ldexp(1.0, std::min(100, std::max(-100, 14 - e))) : 1.0;
          tinv[f] = (float)(1.0 / tau);
          uint16_t* out = &Vh[(size_t)f * per_f];
          for (int j = 0; j < n_pair; j++)
            for (int c = 0; c < 3; c++)
              for (int nb = 0; nb < 3; nb++)
                for (int l = 0; l < 64; l++) {
                  const int q = 4 * c + (l >> 4), a = q / 6, b = q % 6, R = O_.v1_offsets[2 * j + a];
                  for (int i = 0; i < 8; i++) {
                    const double x = entry(R, f, 8 * b + i, 16 * nb + (l & 15)) * tau;
                    const _Float16 h0 = (_Float16)x, h1 = (_Float16)(x - (double)h0);
                    const size_t at = (((((size_t)j * 3 + c) * 3 + nb) * 2) * 64 + l) * 8 + i;
                    std::memcpy(&out[at], &h0, 2);
                    std::memcpy(&out[at + 64 * 8], &h1, 2);
                  }
                }
        }
        Vh_.upload(Vh, &total);
        vtinv_.upload(tinv, &total);
        split_ = true;
      }
    }
    unit_surf_.upload(conv(O_.unit_surf), &total);
    lattice_.upload(O_.lattice, &total);
    {
      std::vector<int32_t> inv((size_t)m_ * m_ * m_, -1);
      for (int p = 0; p < ns_; p++)
        inv[((size_t)O_.lattice[3 * p + 2] * m_ + O_.lattice[3 * p + 1]) * m_ + O_.lattice[3 * p]] = p;
      lattice_inv_.upload(inv, &total);
    }

    const double ts = ki.trg_scal + ki.src_scal;
    // ---- S2M rows: leaves holding sources (Source2UpSetup, include/fmm_pts.txx:2287-2290)
    std::vector<int32_t> s2m_node, s2m_ptr{0}, s2m_items, s2m_orow;
    std::vector<Real> s2m_alpha;
    std::vector<int32_t> l2t_ra, l2t_rc;
    std::vector<Real> l2t_alpha;
    int64_t s2m_pairs = 0;
    for (int i = 0; i < N; i++)
      if (T.is_leaf(i) && T.pt_cnt0(i) > 0) {
        s2m_orow.push_back((int32_t)s2m_node.size());
        s2m_node.push_back(i);
        s2m_items.push_back(i);
        s2m_ptr.push_back((int32_t)s2m_items.size());
        s2m_alpha.push_back((Real)std::pow(2.0, -ts * T.depth[i]));
        s2m_pairs += T.pt_cnt0(i) * ns_;
      }
    n_s2m_ = (int)s2m_node.size();
    h_.s2m_node = s2m_node;
    s2m_node_.upload(s2m_node, &total); s2m_ptr_.upload(s2m_ptr, &total); s2m_items_.upload(s2m_items, &total);
    s2m_orow_.upload(s2m_orow, &total); s2m_alpha_.upload(s2m_alpha, &total);
    chk_vs_ = (int64_t)std::max(1, n_s2m_) * K1_;
    chk_.alloc(chk_vs_ * nvb_, &total);

    // ---- target leaves and L2T rows (Down2TargetSetup, include/fmm_pts.txx:5427-5442)
    std::vector<int32_t> tleaf, l2t_seq;
    for (int i = 0; i < N; i++)
      if (T.is_leaf(i) && T.pt_cnt1(i) > 0) {
        l2t_seq.push_back((int32_t)l2t_ra.size());
        l2t_ra.push_back(i);
        l2t_rc.push_back(N + (int32_t)tleaf.size());  // rows after the up block of eq_
        l2t_alpha.push_back((Real)std::pow(2.0, -ts * T.depth[i]));
        tleaf.push_back(i);
      }
    n_l2t_ = n_tleaf_ = (int)tleaf.size();
    l2t_ra_.upload(l2t_ra, &total); l2t_rc_.upload(l2t_rc, &total); l2t_alpha_.upload(l2t_alpha, &total);
    l2t_seq_.upload(l2t_seq, &total);
    // ctmp_ rows: the check-to-equivalent GEMMs' leaves, and the parents of one level of the
    // rank-factored U2U / D2D
    int max_level_rows = 0;
    for (int l = 0; l <= max_depth_; l++) {
      int na = 0, nb = 0;
      for (int i = T.level_begin[l]; i < T.level_begin[l + 1]; i++) {
        if (T.is_leaf(i)) continue;
        na += T.pt_cnt0(i) > 0;
        nb += T.pt_cnt1(i) > 0;
      }
      max_level_rows = std::max(max_level_rows, std::max(na, nb));
    }
    {
      std::vector<int32_t> seq(std::max(1, max_level_rows));
      for (size_t i = 0; i < seq.size(); i++) seq[i] = (int32_t)i;
      lvl_seq_.upload(seq, &total);
    }
    ctmp_vs_ = (int64_t)std::max(1, std::max(std::max(n_s2m_, n_l2t_), max_level_rows)) * std::max(1, std::max(ru_, rd_));
    ctmp_.alloc(ctmp_vs_ * nvb_, &total);
    t_leaf_.upload(tleaf, &total);
    eq_vs_ = (int64_t)(N + n_l2t_) * K0_;
    dn_vs_ = (int64_t)N * K1_;
    eq_.alloc(eq_vs_ * nvb_, &total);
    dn_.alloc(dn_vs_ * nvb_, &total);

    // ---- M2M / L2L rows per level
    m2m_.clear(); l2l_.clear();
    int64_t m2m_rows = 0, l2l_rows = 0;
    for (int l = 0; l <= max_depth_; l++) {
      auto a = std::make_unique<LevelRows>(), b = std::make_unique<LevelRows>();
      std::vector<int32_t> ra, rc, ra2, rc2;
      for (int i = T.level_begin[l]; i < T.level_begin[l + 1]; i++) {
        if (T.is_leaf(i)) continue;
        if (T.pt_cnt0(i) > 0) { ra.push_back(T.first_child[i]); rc.push_back(i); }  // Up2UpSetup 2618-2619
        if (T.pt_cnt1(i) > 0) { ra2.push_back(i); rc2.push_back(T.first_child[i]); }  // Down2DownSetup 3777-3778
      }
      a->n = (int)ra.size(); a->ra.upload(ra, &total); a->rc.upload(rc, &total);
      b->n = (int)ra2.size(); b->ra.upload(ra2, &total); b->rc.upload(rc2, &total);
      m2m_rows += a->n; l2l_rows += b->n;
      m2m_.push_back(std::move(a)); l2l_.push_back(std::move(b));
    }
    {
      int64_t out = std::max((int64_t)n_s2m_ * std::max(ru_, K0_), (int64_t)n_l2t_ * std::max(rd_, K0_));
      for (auto& L : m2m_) out = std::max(out, (int64_t)L->n * K0_);
      for (auto& L : l2l_) out = std::max(out, (int64_t)L->n * 8 * K1_);
      gemm_work_elems_ = std::min<int64_t>(int64_t(1) << 25, std::max<int64_t>(1, 4 * out * nvb_));
      gemm_work_.alloc(gemm_work_elems_, &total);
    }

    const int nl_x = T.list_count(kX), nl_w = T.list_count(kW);
    // ---- X list S2L (X_ListSetup, include/fmm_pts.txx:4444-4743)
    std::vector<int32_t> x_node, x_ptr{0}, x_items, x_orow;
    for (int i = 0; i < N; i++) {
      if (T.pt_cnt1(i) == 0) continue;
      if (T.is_leaf(i) && T.pt_cnt1(i) <= ns_) continue;  // handled in the U list
      int cnt = 0;
      for (int k = 0; k < nl_x; k++) {
        int32_t s = T.lists[kX][(size_t)i * nl_x + k];
        if (s >= 0 && T.is_leaf(s) && T.pt_cnt0(s) > 0) { x_items.push_back(img(i, kX, k, s)); cnt++; }
      }
      if (!cnt) continue;
      x_node.push_back(i); x_orow.push_back(i);
      x_ptr.push_back((int32_t)x_items.size());
    }
    n_x_ = (int)x_node.size();
    x_node_.upload(x_node, &total); x_ptr_.upload(x_ptr, &total); x_items_.upload(x_items, &total); x_orow_.upload(x_orow, &total);
    h_.x_node = x_node; h_.x_ptr = x_ptr; h_.x_items = x_items;

    // ---- V list (V_ListSetup, include/fmm_pts.txx:3346-3595): parents holding sources / targets
    std::vector<int32_t> slot_of(N, -1), v_src_fc, v_trg_fc, v_partner;
    std::vector<Real> v_src_scal, v_trg_scal;
    for (int i = 0; i < N; i++)
      if (!T.is_leaf(i) && T.pt_cnt0(i) > 0) {
        slot_of[i] = (int32_t)v_src_fc.size();
        v_src_fc.push_back(T.first_child[i]);
        v_src_scal.push_back((Real)std::pow(2.0, ki.src_scal * (T.depth[i] + 1)));
        h_.vsrc_depth.push_back(T.depth[i]);
      }
    std::vector<int32_t> offs(O_.v1_offsets.begin(), O_.v1_offsets.end());
    n_off_ = (int)offs.size();
    int64_t v_pairs = 0;
    const int nl_v1 = T.list_count(kV1);
    const RelTable& v1t = rel_table(kV1);
    for (int i = 0; i < N; i++) {
      if (T.is_leaf(i) || T.pt_cnt1(i) == 0) continue;
      std::vector<int32_t> row(n_off_, -1);
      int cnt = 0;
      for (int k = 0; k < nl_v1; k++) {
        int32_t s = T.lists[kV1][(size_t)i * nl_v1 + k];
        if (s < 0 || T.pt_cnt0(s) == 0) continue;
        const auto& r = v1t.rel[k];
        int R = (r[0] + 1) + 3 * (r[1] + 1) + 9 * (r[2] + 1);
        int oi = (int)(std::find(offs.begin(), offs.end(), R) - offs.begin());
        row[oi] = slot_of[s];
        cnt++;
      }
      if (!cnt) continue;
      v_pairs += cnt;
      v_trg_fc.push_back(T.first_child[i]);
      v_trg_scal.push_back((Real)std::pow(2.0, ki.trg_scal * (T.depth[i] + 1)));
      h_.vtrg_depth.push_back(T.depth[i]);
      v_partner.insert(v_partner.end(), row.begin(), row.end());
    }
    n_vsrc_ = (int)v_src_fc.size();
    n_vtrg_ = (int)v_trg_fc.size();
    v_src_fc_.upload(v_src_fc, &total); v_trg_fc_.upload(v_trg_fc, &total);
    v_src_scal_.upload(v_src_scal, &total); v_trg_scal_.upload(v_trg_scal, &total);
    v_partner_.upload(v_partner, &total); v_offsets_.upload(offs, &total);
    h_.v_partner = v_partner; h_.v_src_fc = v_src_fc; h_.v_trg_fc = v_trg_fc; h_.offs = offs;
    build_tiles(v_partner, n_vtrg_, n_vsrc_, vt_, &total);
    if (split_) {
      // per slot and component; the adjoint's sources are the target slots
      vbound_.alloc(std::max(1, std::max(n_vsrc_ * e0_, n_vtrg_ * e1_)), &total);
      vrscale_.alloc(std::max(1, vt_.n), &total);
    }
    stats.v_tiles = vt_.n;
    stats.v_tile_partners = vt_.partners;
    stats.v_tile_max_partners = vt_.max_partners;
    stats.v_tile_staged = vt_.staged_delta;
    X_.alloc((int64_t)n_vsrc_ * M_dim_ * 16 * e0_, &total);
    Y_.alloc((int64_t)n_vtrg_ * M_dim_ * 16 * e1_, &total);
    {
      // 16-vector spectra when the plan batches 8 or more vectors and they fit in 64 GiB
      const int64_t bytes = (int64_t)(n_vsrc_ + n_vtrg_) * M_dim_ * 256 * (int64_t)sizeof(Real);
      vec16_ = std::is_same<Real, float>::value && e0_ == 1 && e1_ == 1 && nvb_ >= 8 && n_vtrg_ > 0 &&
               bytes <= (int64_t(64) << 30);
      if (vec16_) {
        Xv_.alloc((int64_t)n_vsrc_ * M_dim_ * 256, &total);
        Yv_.alloc((int64_t)n_vtrg_ * M_dim_ * 256, &total);
      }
    }

    // ---- far field (L2T own surface + W list M2T) and U list per target leaf
    std::vector<int32_t> far_ptr{0}, far_node, u_ptr{0}, u_items;
    std::vector<int64_t> far_den;
    std::vector<double> far_rad;
    const int nl_u0 = T.list_count(kU0), nl_u1 = T.list_count(kU1), nl_u2 = T.list_count(kU2);
    int64_t u_pairs = 0, far_pairs = 0, u_stream = 0;
    for (int r = 0; r < n_tleaf_; r++) {
      const int t = tleaf[r];
      far_node.push_back(t); far_den.push_back((int64_t)(N + r) * K0_); far_rad.push_back(kRad1);
      for (int k = 0; k < nl_w; k++) {  // W_ListSetup 4910-4913
        int32_t s = T.lists[kW][(size_t)t * nl_w + k];
        if (s < 0 || T.pt_cnt0(s) == 0) continue;
        if (T.is_leaf(s) && T.pt_cnt0(s) <= ns_) continue;
        far_node.push_back(img(t, kW, k, s)); far_den.push_back((int64_t)s * K0_); far_rad.push_back(kRad0);
      }
      far_pairs += (int64_t)(far_node.size() - far_ptr.back()) * ns_ * T.pt_cnt1(t);
      far_ptr.push_back((int32_t)far_node.size());
      auto add = [&](int32_t s, int type, int k) {
        if (s >= 0 && T.is_leaf(s) && T.pt_cnt0(s) > 0) {
          u_items.push_back(img(t, type, k, s));
          u_pairs += T.pt_cnt0(s) * T.pt_cnt1(t);
          u_stream += T.pt_cnt0(s);
        }
      };
      for (int k = 0; k < nl_u0; k++) add(T.lists[kU0][(size_t)t * nl_u0 + k], kU0, k);  // U_ListSetup 5197-5321
      for (int k = 0; k < nl_u1; k++) add(T.lists[kU1][(size_t)t * nl_u1 + k], kU1, k);
      for (int k = 0; k < nl_u2; k++) add(T.lists[kU2][(size_t)t * nl_u2 + k], kU2, k);
      if (T.pt_cnt1(t) <= ns_)
        for (int k = 0; k < nl_x; k++) add(T.lists[kX][(size_t)t * nl_x + k], kX, k);
      for (int k = 0; k < nl_w; k++) {
        int32_t s = T.lists[kW][(size_t)t * nl_w + k];
        if (s >= 0 && T.is_leaf(s) && T.pt_cnt0(s) <= ns_) add(s, kW, k);
      }
      u_ptr.push_back((int32_t)u_items.size());
    }
    far_ptr_.upload(far_ptr, &total); far_node_.upload(far_node, &total); far_den_.upload(far_den, &total);
    far_rad_.upload(far_rad, &total);
    u_ptr_.upload(u_ptr, &total); u_items_.upload(u_items, &total);
    if (bndry_ != kFreeSpace) {  // FMM_Pts::PeriodicBC and PostProcessing (include/fmm_pts.txx:2641-2657,5740-5774)
      const std::vector<double>& bc = O_.bc[bndry_];
      if (bc.size() != (size_t)K0_ * K1_) throw Error(kErrInternal, "periodic tree without its boundary operator");
      bc_.upload(conv(bc), &total);
      vol_ = bndry_ == kPXYZ && ki.id != 1;
      if (vol_) {
        std::vector<double> xa(3 * (size_t)n_trg_);
        for (int64_t k = 0; k < n_trg_; k++)
          for (int d = 0; d < 3; d++) xa[d * n_trg_ + k] = trg_xyz[3 * T.trg.perm[k] + d];
        txa_.upload(xa, &total);
      }
    }
    h_.tleaf = tleaf; h_.u_ptr = u_ptr; h_.u_items = u_items;
    h_.far_ptr = far_ptr; h_.far_node = far_node; h_.far_rad = far_rad;

    ev_.resize(kPhaseCount + 1);
    for (auto& e : ev_) HIP_CHECK(hipEventCreate(&e));
    phase_ms.assign(kPhaseCount, 0.0);

    stats.n_nodes = N;
    for (int i = 0; i < N; i++) stats.n_leaves += T.is_leaf(i);
    stats.max_depth = max_depth_;
    stats.s2m_rows = n_s2m_; stats.l2t_rows = n_l2t_; stats.m2m_rows = m2m_rows; stats.l2l_rows = l2l_rows;
    stats.x_rows = n_x_; stats.x_items = (int64_t)x_items.size();
    stats.v_src_slots = n_vsrc_; stats.v_trg_slots = n_vtrg_; stats.v_pairs = v_pairs;
    stats.u_rows = n_tleaf_; stats.u_items = (int64_t)u_items.size(); stats.u_pairs = u_pairs;
    stats.u_src_points_streamed = u_stream;
    stats.far_items = (int64_t)far_node.size(); stats.far_pairs = far_pairs; stats.s2m_pairs = s2m_pairs;
    stats.device_bytes = total;
    device_bytes_ = total;
  }

  // ---- adjoint evaluation: the transpose of the evaluation's linear map, phase by phase in
  // reverse order with the transposed translations.  For the symmetric kernels (K(x, y) =
  // K(y, x)^T: laplace_poten, stokes_vel) every transposed point interaction is the same kernel
  // with sources and targets exchanged, and the V-list M2L of offset R transposes to the one of
  // offset -R (the check and equivalent lattices of a node coincide at radius RAD0).
  struct HostLists {  // forward lists kept on the host to build the transposed ones
    std::vector<int32_t> tleaf, u_ptr, u_items, far_ptr, far_node, x_node, x_ptr, x_items, s2m_node;
    std::vector<double> far_rad;
    std::vector<int32_t> v_partner, v_src_fc, v_trg_fc, offs, vsrc_depth, vtrg_depth;
    // the node table (real nodes, then the periodic images as (node, shift code))
    std::vector<double> center;
    std::vector<float> nsz;
    std::vector<int32_t> sb, se, tb, te;
    std::vector<std::pair<int32_t, int>> imgs;
  } h_;
  int64_t device_bytes_ = 0;

  struct AdjointPlan {
    DBuf<Real> uc2ue0T, uc2ue1T, dc2de0T, dc2de1T, u2uT, d2dT, u2urT, d2drT;
    // node table of the transposed lists: the forward one plus images of target-side nodes (a
    // forward item at shift s becomes, transposed, its target seen from the source at -s)
    DBuf<double> center;
    DBuf<float> nsz;
    DBuf<int32_t> sb, se, tb, te;
    DBuf<Real> bcT;        // periodic boundary operator transposed (K1_ x K0_)
    DBuf<double> volpart;  // per-block partial sums of the volume-potential correction's transpose
    DBuf<Real> tden;     // adjoint input gathered into target order (k1 x n_trg per vector)
    DBuf<double> sval;   // adjoint output accumulated in source order (k0 x n_src per vector)
    DBuf<int32_t> iota;  // 0, 1, 2, ...: one item per row
    DBuf<int32_t> u_leaf, u_ptr, u_items;  // U list: source leaf <- target leaves
    int n_u = 0;
    DBuf<int32_t> w_node, w_ptr, w_items;  // W list: RAD0 surface of a source node <- target leaves
    int n_w = 0;
    DBuf<int32_t> x_leaf, x_ptr, x_node;   // X list: source leaf <- RAD0 check surfaces
    DBuf<int64_t> x_den;
    DBuf<double> x_rad;
    int n_x = 0;
    DBuf<int64_t> s2u_den;                 // S2U: source leaf <- its RAD1 check surface
    DBuf<double> s2u_rad;
    TileSet vt;                            // V list: source parents <- target parents
    DBuf<float> vrscale;                   // run scales of vt (fp16 Hadamard)
    DBuf<int32_t> v_fc;                    // first child of each transposed target slot
    DBuf<Real> fft_scal, ifft_scal;
    int n_vt = 0;
  };
  std::unique_ptr<AdjointPlan> adj_;

  static std::vector<Real> transposed(const std::vector<Real>& a, int rows, int cols) {
    std::vector<Real> t((size_t)rows * cols);
    for (int i = 0; i < rows; i++)
      for (int j = 0; j < cols; j++) t[(size_t)j * rows + i] = a[(size_t)i * cols + j];
    return t;
  }

  // CSR of the transposed relation: for rows r with items it, the rows of item -> list of r
  // Groups (key, value, payload) triples by key (stable): the rows out_row (keys in increasing
  // order), their item ranges out_ptr and the values and payloads in forward order.
  static void group_by_key(const std::vector<int32_t>& keys, const std::vector<int32_t>& vals,
                           const std::vector<int32_t>& pay, int n_keys, std::vector<int32_t>& out_row,
                           std::vector<int32_t>& out_ptr, std::vector<int32_t>& out_items, std::vector<int32_t>& out_pay) {
    std::vector<int32_t> start(n_keys + 1, 0);
    for (int32_t k : keys) start[k + 1]++;
    for (int k = 0; k < n_keys; k++) start[k + 1] += start[k];
    std::vector<int32_t> flat(keys.size()), flatp(keys.size());
    std::vector<int32_t> fill = start;
    for (size_t i = 0; i < keys.size(); i++) {
      flat[fill[keys[i]]] = vals[i];
      flatp[fill[keys[i]]++] = pay[i];
    }
    out_row.clear(); out_ptr.assign(1, 0); out_items.clear(); out_pay.clear();
    for (int k = 0; k < n_keys; k++) {
      if (start[k + 1] == start[k]) continue;
      out_row.push_back(k);
      out_items.insert(out_items.end(), flat.begin() + start[k], flat.begin() + start[k + 1]);
      out_pay.insert(out_pay.end(), flatp.begin() + start[k], flatp.begin() + start[k + 1]);
      out_ptr.push_back((int32_t)out_items.size());
    }
  }

  void build_adjoint() {
    auto a = std::make_unique<AdjointPlan>();
    int64_t total = 0;
    const int N = n_nodes_;
    // transposed operators, in the padded layouts of the forward ones
    auto conv = [](const std::vector<double>& v) { return std::vector<Real>(v.begin(), v.end()); };
    auto pad_cols = [&](const std::vector<double>& m, int rows, int r, int rp) {
      std::vector<Real> out((size_t)rows * rp, Real(0));
      for (int i = 0; i < rows; i++)
        for (int j = 0; j < r; j++) out[(size_t)i * rp + j] = (Real)m[(size_t)i * r + j];
      return out;
    };
    auto pad_rows = [&](const std::vector<double>& m, int r, int cols, int rp) {
      std::vector<Real> out((size_t)rp * cols, Real(0));
      for (int i = 0; i < r; i++)
        for (int j = 0; j < cols; j++) out[(size_t)i * cols + j] = (Real)m[(size_t)i * cols + j];
      return out;
    };
    a->uc2ue0T.upload(transposed(pad_cols(O_.uc2ue0, K1_, O_.rank_u, ru_), K1_, ru_), &total);  // ru x K1
    a->uc2ue1T.upload(transposed(pad_rows(O_.uc2ue1, O_.rank_u, K0_, ru_), ru_, K0_), &total);  // K0 x ru
    a->dc2de0T.upload(transposed(pad_cols(O_.dc2de0, K1_, O_.rank_d, rd_), K1_, rd_), &total);  // rd x K1
    a->dc2de1T.upload(transposed(pad_rows(O_.dc2de1, O_.rank_d, K0_, rd_), rd_, K0_), &total);  // K0 x rd
    if (u2u_fact_) a->u2urT.upload(transposed(conv(u2u_rank_factor(O_, ru_)), 8 * K0_, ru_), &total);  // ru x 8K0
    if (d2d_fact_) a->d2drT.upload(transposed(conv(d2d_rank_factor(O_, rd_)), rd_, 8 * K1_), &total);  // 8K1 x rd
    a->u2uT.upload(transposed(conv(O_.u2u), 8 * K0_, K0_), &total);                                     // K0 x 8K0
    a->d2dT.upload(transposed(conv(O_.d2d), K1_, 8 * K1_), &total);                                     // 8K1 x K1
    a->tden.alloc((int64_t)k1_ * n_trg_ * nvb_, &total);
    a->sval.alloc((int64_t)k0_ * n_src_ * nvb_, &total);
    // the adjoint's node table and the transposition of a list with images: a forward item v of
    // row node r (v a source node, or an image (s, shift)) becomes item adj(r, -shift) of row s
    std::vector<double> ac = h_.center;
    std::vector<float> ansz = h_.nsz;
    std::vector<int32_t> asb = h_.sb, ase = h_.se, atb = h_.tb, ate = h_.te;
    std::unordered_map<int64_t, int32_t> aimg;
    auto real_of = [&](int32_t v, int& code) -> int32_t {
      if (v < N) { code = 13; return v; }
      code = h_.imgs[v - N].second;
      return h_.imgs[v - N].first;
    };
    auto adj = [&](int32_t node, int code) -> int32_t {
      if (code == 13) return node;
      const int neg = 26 - code;  // the negated shift
      const int64_t key = (int64_t)node * 27 + neg;
      auto it = aimg.find(key);
      if (it != aimg.end()) return it->second;
      const int32_t v = (int32_t)ansz.size();
      const int sh[3] = {neg % 3 - 1, (neg / 3) % 3 - 1, neg / 9 - 1};
      for (int d = 0; d < 3; d++) ac.push_back(h_.center[3 * (size_t)node + d] + sh[d]);
      ansz.push_back(h_.nsz[node]);
      asb.push_back(h_.sb[node]); ase.push_back(h_.se[node]);
      atb.push_back(h_.tb[node]); ate.push_back(h_.te[node]);
      aimg.emplace(key, v);
      return v;
    };
    auto transpose = [&](const std::vector<int32_t>& row_key, const std::vector<int32_t>& ptr,
                         const std::vector<int32_t>& items, std::vector<int32_t>& out_row, std::vector<int32_t>& out_ptr,
                         std::vector<int32_t>& out_items, std::vector<int32_t>& out_real) {
      std::vector<int32_t> keys(items.size()), vals(items.size()), real(items.size());
      for (size_t r = 0; r + 1 < ptr.size(); r++)
        for (int32_t i = ptr[r]; i < ptr[r + 1]; i++) {
          int code = 13;
          keys[i] = real_of(items[i], code);
          vals[i] = adj(row_key[r], code);
          real[i] = row_key[r];
        }
      group_by_key(keys, vals, real, N, out_row, out_ptr, out_items, out_real);
    };
    {
      std::vector<int32_t> iota(std::max<int>(n_tleaf_, n_s2m_) + 1);
      for (size_t i = 0; i < iota.size(); i++) iota[i] = (int32_t)i;
      a->iota.upload(iota, &total);
    }
    // U list
    {
      std::vector<int32_t> row, ptr, items, real;
      transpose(h_.tleaf, h_.u_ptr, h_.u_items, row, ptr, items, real);
      a->n_u = (int)row.size();
      a->u_leaf.upload(row, &total); a->u_ptr.upload(ptr, &total); a->u_items.upload(items, &total);
    }
    // W list: the far items after each target leaf's own L2T surface
    {
      std::vector<int32_t> key, ptr{0}, items;
      for (size_t r = 0; r < h_.tleaf.size(); r++) {
        for (int32_t i = h_.far_ptr[r]; i < h_.far_ptr[r + 1]; i++)
          if (h_.far_rad[i] == kRad0) items.push_back(h_.far_node[i]);
        key.push_back(h_.tleaf[r]);
        ptr.push_back((int32_t)items.size());
      }
      std::vector<int32_t> row, tptr, titems, treal;
      transpose(key, ptr, items, row, tptr, titems, treal);
      a->n_w = (int)row.size();
      a->w_node.upload(row, &total); a->w_ptr.upload(tptr, &total); a->w_items.upload(titems, &total);
    }
    // X list
    {
      std::vector<int32_t> row, ptr, items, real;
      transpose(h_.x_node, h_.x_ptr, h_.x_items, row, ptr, items, real);
      a->n_x = (int)row.size();
      std::vector<int64_t> den(items.size());  // the check potentials of the real nodes
      for (size_t i = 0; i < items.size(); i++) den[i] = (int64_t)real[i] * K1_;
      a->x_leaf.upload(row, &total); a->x_ptr.upload(ptr, &total); a->x_node.upload(items, &total);
      a->x_den.upload(den, &total);
      a->x_rad.upload(std::vector<double>(items.size(), kRad0), &total);
    }
    // S2U: row r reads its own check potential, chk row r
    {
      std::vector<int64_t> den(n_s2m_);
      for (int r = 0; r < n_s2m_; r++) den[r] = (int64_t)r * K1_;
      a->s2u_den.upload(den, &total);
      a->s2u_rad.upload(std::vector<double>(n_s2m_, kRad1), &total);
    }
    // V list: new targets are the source parents, new sources the target parents; the pair
    // (t, R, s) of the evaluation becomes (s, -R, t), whose offset index is that of 26 - R
    {
      const int n_off = n_off_;
      std::vector<int> neg(n_off, -1);
      for (int o = 0; o < n_off; o++)
        for (int o2 = 0; o2 < n_off; o2++)
          if (h_.offs[o2] == 26 - h_.offs[o]) neg[o] = o2;
      for (int o = 0; o < n_off; o++)
        if (neg[o] < 0) throw Error(kErrInternal, "V-list offsets are not closed under negation");
      std::vector<int32_t> full((size_t)n_vsrc_ * n_off, -1);
      std::vector<char> has(n_vsrc_, 0);
      for (int t = 0; t < n_vtrg_; t++)
        for (int o = 0; o < n_off; o++) {
          const int32_t sl = h_.v_partner[(size_t)t * n_off + o];
          if (sl < 0) continue;
          full[(size_t)sl * n_off + neg[o]] = t;
          has[sl] = 1;
        }
      std::vector<int32_t> partner, fc;
      std::vector<Real> fscal(n_vtrg_), iscal;
      const KernelInfo& ki = O_.ker;
      for (int sl = 0; sl < n_vsrc_; sl++) {
        if (!has[sl]) continue;
        partner.insert(partner.end(), full.begin() + (size_t)sl * n_off, full.begin() + (size_t)(sl + 1) * n_off);
        fc.push_back(h_.v_src_fc[sl]);
        iscal.push_back((Real)std::pow(2.0, ki.trg_scal * (h_.vsrc_depth[sl] + 1)));
      }
      for (int t = 0; t < n_vtrg_; t++) fscal[t] = (Real)std::pow(2.0, ki.src_scal * (h_.vtrg_depth[t] + 1));
      a->n_vt = (int)fc.size();
      build_tiles(partner, a->n_vt, n_vtrg_, a->vt, &total);
      if (split_) a->vrscale.alloc(std::max(1, a->vt.n), &total);
      a->v_fc.upload(fc, &total);
      a->fft_scal.upload(fscal, &total);
      a->ifft_scal.upload(iscal, &total);
    }
    a->center.upload(ac, &total);
    a->nsz.upload(ansz, &total);
    a->sb.upload(asb, &total); a->se.upload(ase, &total);
    a->tb.upload(atb, &total); a->te.upload(ate, &total);
    if (bndry_ != kFreeSpace) {
      a->bcT.upload(transposed(conv(O_.bc[bndry_]), K0_, K1_), &total);
      if (vol_) a->volpart.alloc((int64_t)K::vol_adjoint_blocks() * 3 * nvb_, &total);
    }
    device_bytes_ += total;
    stats.device_bytes = device_bytes_;
    adj_ = std::move(a);
  }

  void adjoint_one(const double* d_g, double* d_out, double den_scale, double out_scale, int nv) {
    AdjointPlan& a = *adj_;
    const int ek = O_.ker.exp_kind;  // == the point kernel for the symmetric kernels
    const int64_t g_vs = (int64_t)k1_ * n_trg_, o_vs = (int64_t)k0_ * n_src_;
    const int sp = tuning.gemm_split_blocks;
    // scatter^T: the adjoint input in sorted target order (times the output scale)
    K::gather_density(st_, n_trg_, k1_, d_g, trg_perm_.p, out_scale, a.tden.p, nv, g_vs);
    K::fill(st_, eq_.p, eq_vs_ * nv);
    K::fill(st_, dn_.p, dn_vs_ * nv);
    K::fill64(st_, a.sval.p, o_vs * nv);
    // (volume-potential correction)^T: the targets' weighted sums onto the root's equivalent density
    if (vol_)
      K::vol_correction_adjoint(st_, O_.ker.pt_kind, n_trg_, txa_.p, a.tden.p, K0_, e0_, eq_.p, a.volpart.p, nv, g_vs, eq_vs_);
    // U list^T: target points -> source points of the leaves listing them
    K::p2p_points(st_, ek, a.n_u, a.u_leaf.p, a.u_ptr.p, a.u_items.p, a.sb.p, a.se.p, a.tb.p, a.te.p,
                  a.center.p, sx64_.p, n_src_, tx64_.p, a.tden.p, n_trg_, a.sval.p, true, nv, g_vs, o_vs);
    // (L2T P2P)^T: target points -> their leaf's downward equivalent surface (RAD1)
    K::p2p_surface(st_, ek, n_tleaf_, t_leaf_.p, a.iota.p, t_leaf_.p, a.tb.p, a.te.p, a.center.p, a.nsz.p, kRad1,
                   unit_surf_.p, ns_, tx_.p, a.tden.p, n_trg_, eq_.p, l2t_rc_.p, K0_, true, nv, g_vs, eq_vs_);
    // (W list M2T)^T: target points -> upward equivalent surfaces (RAD0) of their W sources
    K::p2p_surface(st_, ek, a.n_w, a.w_node.p, a.w_ptr.p, a.w_items.p, a.tb.p, a.te.p, a.center.p, a.nsz.p, kRad0,
                   unit_surf_.p, ns_, tx_.p, a.tden.p, n_trg_, eq_.p, a.w_node.p, K0_, true, nv, g_vs, eq_vs_);
    // D2T GEMMs^T: equivalent -> check potential of each target leaf
    K::gemm_rows(st_, n_l2t_, rd_, K0_, eq_.p, l2t_rc_.p, K0_, a.dc2de1T.p, ctmp_.p, l2t_seq_.p, rd_, nullptr, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, eq_vs_, ctmp_vs_);
    K::gemm_rows(st_, n_l2t_, K1_, rd_, ctmp_.p, l2t_seq_.p, rd_, a.dc2de0T.p, dn_.p, l2t_ra_.p, K1_, l2t_alpha_.p, true,
                 gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, dn_vs_);
    // D2D^T, bottom-up: parent += [8 children] x [D2D_0 .. D2D_7]^T
    for (int l = max_depth_; l >= 1; l--) {
      const LevelRows& L = *l2l_[l - 1];
      if (d2d_fact_ && L.n >= kFactorRows) {
        K::gemm_rows(st_, L.n, rd_, 8 * K1_, dn_.p, L.rc.p, K1_, a.d2drT.p, ctmp_.p, lvl_seq_.p, rd_, nullptr, false,
                     gemm_work_.p, gemm_work_elems_, sp, nv, dn_vs_, ctmp_vs_);
        K::gemm_rows(st_, L.n, K1_, rd_, ctmp_.p, lvl_seq_.p, rd_, a.dc2de0T.p, dn_.p, L.ra.p, K1_, nullptr, true,
                     gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, dn_vs_);
      } else {
        K::gemm_rows(st_, L.n, K1_, 8 * K1_, dn_.p, L.rc.p, K1_, a.d2dT.p, dn_.p, L.ra.p, K1_, nullptr, true, gemm_work_.p,
                     gemm_work_elems_, sp, nv, dn_vs_, dn_vs_);
      }
    }
    // (periodic boundary)^T: root equivalent density += root check potential x BC^T
    if (bndry_ != kFreeSpace)
      K::gemm_rows(st_, 1, K0_, K1_, dn_.p, lvl_seq_.p, K1_, a.bcT.p, eq_.p, lvl_seq_.p, K0_, nullptr, true, gemm_work_.p,
                   gemm_work_elems_, sp, nv, dn_vs_, eq_vs_);
    // V list^T: the M2L of offset -R from the target parents' children to the source parents'
    // (the spectra buffers swap roles: Y_ holds the new sources, X_ the new targets)
    for (int v = 0; v < nv; v++) {
      K::fft_up(st_, n_vtrg_, v_trg_fc_.p, dn_.p + v * dn_vs_, K1_, e1_, m_, lattice_.p, lattice_inv_.p, a.fft_scal.p, Y_.p,
                1, 0, split_ ? vbound_.p : nullptr);
      K::hadamard_tiled(st_, a.vt.n, M_dim_, e0_, e1_, n_off_, a.vt.t0.p, a.vt.u0.p, a.vt.usrc.p, a.vt.d0.p, a.vt.dent.p,
                        a.vt.uidx.p, a.vt.gmask.p, v_offsets_.p, V_.p, Y_.p, X_.p, tuning.hadamard_tpb,
                        split_ ? &hparts(a.vrscale) : nullptr);
      K::ifft_down(st_, a.n_vt, a.v_fc.p, X_.p, eq_.p + v * eq_vs_, K0_, e0_, m_, lattice_.p, lattice_inv_.p, ns_,
                   a.ifft_scal.p);
    }
    // (X list S2L)^T: downward check surfaces (RAD0) -> source points
    K::far_points(st_, ek, a.n_x, a.x_leaf.p, a.x_ptr.p, a.x_node.p, a.x_den.p, a.x_rad.p, a.sb.p, a.se.p,
                  a.center.p, a.nsz.p, unit_surf_.p, ns_, sx_.p, n_src_, dn_.p, a.sval.p, true, nv, dn_vs_, o_vs);
    // U2U^T, top-down: [8 children] += parent x [U2U_0 .. U2U_7]^T
    for (int l = 0; l <= max_depth_ - 1; l++) {
      const LevelRows& L = *m2m_[l];
      if (u2u_fact_ && L.n >= kFactorRows) {
        K::gemm_rows(st_, L.n, ru_, K0_, eq_.p, L.rc.p, K0_, a.uc2ue1T.p, ctmp_.p, lvl_seq_.p, ru_, nullptr, false,
                     gemm_work_.p, gemm_work_elems_, sp, nv, eq_vs_, ctmp_vs_);
        K::gemm_rows(st_, L.n, 8 * K0_, ru_, ctmp_.p, lvl_seq_.p, ru_, a.u2urT.p, eq_.p, L.ra.p, K0_, nullptr, true,
                     gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, eq_vs_);
      } else {
        K::gemm_rows(st_, L.n, 8 * K0_, K0_, eq_.p, L.rc.p, K0_, a.u2uT.p, eq_.p, L.ra.p, K0_, nullptr, true, gemm_work_.p,
                     gemm_work_elems_, sp, nv, eq_vs_, eq_vs_);
      }
    }
    // S2U GEMMs^T: upward equivalent -> check potential of each source leaf
    K::gemm_rows(st_, n_s2m_, ru_, K0_, eq_.p, s2m_node_.p, K0_, a.uc2ue1T.p, ctmp_.p, s2m_orow_.p, ru_, nullptr, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, eq_vs_, ctmp_vs_);
    K::gemm_rows(st_, n_s2m_, K1_, ru_, ctmp_.p, s2m_orow_.p, ru_, a.uc2ue0T.p, chk_.p, s2m_orow_.p, K1_, s2m_alpha_.p, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, chk_vs_);
    // (S2U P2P)^T: upward check surface (RAD1) -> the leaf's source points
    K::far_points(st_, ek, n_s2m_, s2m_node_.p, a.iota.p, s2m_node_.p, a.s2u_den.p, a.s2u_rad.p, a.sb.p, a.se.p,
                  a.center.p, a.nsz.p, unit_surf_.p, ns_, sx_.p, n_src_, chk_.p, a.sval.p, true, nv, chk_vs_, o_vs);
    // gather^T: back to the input order of the sources (times the density scale)
    K::scatter_output(st_, n_src_, k0_, a.sval.p, src_perm_.p, den_scale, d_out, nv, o_vs);
  }

 public:
  bool adjoint_supported() const override { return O_.ker.pt_kind == O_.ker.exp_kind; }
  void evaluate_adjoint(const double* d_g, double* d_out, int n_vec, double den_scale, double out_scale) override {
    if (!adjoint_supported())
      throw Error(kErrUnsupported, std::string("the adjoint evaluation needs a symmetric kernel (laplace_poten, stokes_vel); ") +
                                       O_.ker.name + " is not");
    if (n_vec < 1) throw Error(kErrInvalidArgument, "n_vec must be >= 1");
    if (!adj_) build_adjoint();
    for (int v = 0; v < n_vec; v += nvb_)
      adjoint_one(d_g + (int64_t)v * k1_ * n_trg_, d_out + (int64_t)v * k0_ * n_src_, den_scale, out_scale,
                  std::min(nvb_, n_vec - v));
  }

 private:
  void evaluate_one(const double* d_den, double* d_out, double den_scale, double out_scale, int nv) {
    const int pk = O_.ker.pt_kind, ek = O_.ker.exp_kind;
    const int64_t den_vs = (int64_t)k0_ * n_src_, val_vs = (int64_t)k1_ * n_trg_;
    const int sp = tuning.gemm_split_blocks;
    mark(0);
    K::gather_density(st_, n_src_, k0_, d_den, src_perm_.p, den_scale, sden_.p, nv, den_vs);
    K::fill(st_, eq_.p, eq_vs_ * nv);
    K::fill(st_, dn_.p, dn_vs_ * nv);
    mark(1);
    // S2U: sources -> upward check surface (RAD1), then check -> equivalent
    K::p2p_surface(st_, ek, n_s2m_, s2m_node_.p, s2m_ptr_.p, s2m_items_.p, src_begin_.p, src_end_.p, center_.p, node_size_.p,
                   kRad1, unit_surf_.p, ns_, sx_.p, sden_.p, n_src_, chk_.p, s2m_orow_.p, K1_, false, nv, den_vs, chk_vs_);
    mark(2);
    K::gemm_rows(st_, n_s2m_, ru_, K1_, chk_.p, s2m_orow_.p, K1_, uc2ue0_.p, ctmp_.p, s2m_orow_.p, ru_, s2m_alpha_.p, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, chk_vs_, ctmp_vs_);
    K::gemm_rows(st_, n_s2m_, K0_, ru_, ctmp_.p, s2m_orow_.p, ru_, uc2ue1_.p, eq_.p, s2m_node_.p, K0_, nullptr, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, eq_vs_);
    mark(3);
    // U2U, bottom-up: parent row = [8 children] x stacked U2U
    for (int l = max_depth_ - 1; l >= 0; l--) {
      const LevelRows& L = *m2m_[l];
      if (u2u_fact_ && L.n >= kFactorRows) {  // [8 children] -> rank coordinates -> parent
        K::gemm_rows(st_, L.n, ru_, 8 * K0_, eq_.p, L.ra.p, K0_, u2ur_.p, ctmp_.p, lvl_seq_.p, ru_, nullptr, false,
                     gemm_work_.p, gemm_work_elems_, sp, nv, eq_vs_, ctmp_vs_);
        K::gemm_rows(st_, L.n, K0_, ru_, ctmp_.p, lvl_seq_.p, ru_, uc2ue1_.p, eq_.p, L.rc.p, K0_, nullptr, false,
                     gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, eq_vs_);
      } else {
        K::gemm_rows(st_, L.n, K0_, 8 * K0_, eq_.p, L.ra.p, K0_, u2u_.p, eq_.p, L.rc.p, K0_, nullptr, false, gemm_work_.p,
                     gemm_work_elems_, sp, nv, eq_vs_, eq_vs_);
      }
    }
    // periodic images beyond the root's colleagues: root check potential += root equivalent x BC
    if (bndry_ != kFreeSpace)
      K::gemm_rows(st_, 1, K1_, K0_, eq_.p, lvl_seq_.p, K0_, bc_.p, dn_.p, lvl_seq_.p, K1_, nullptr, true, gemm_work_.p,
                   gemm_work_elems_, sp, nv, eq_vs_, dn_vs_);
    mark(4);
    // X list: source points -> downward check surface (RAD0)
    K::p2p_surface(st_, ek, n_x_, x_node_.p, x_ptr_.p, x_items_.p, src_begin_.p, src_end_.p, center_.p, node_size_.p, kRad0,
                   unit_surf_.p, ns_, sx_.p, sden_.p, n_src_, dn_.p, x_orow_.p, K1_, true, nv, den_vs, dn_vs_);
    mark(5);
    // V list: groups of 16 vectors through the 16-vector Hadamard, or one vector at a time
    // through the spectra buffers (with several passes the three sub-phases are timed per
    // pass and summed)
    const bool grouped = use_vec16(nv);
    const int n_pass = grouped ? (nv + 15) / 16 : nv;
    if (timing_ && n_pass > 1) ensure_vlist_events(n_pass);
    for (int g = 0; g < n_pass; g++) {
      const bool sub = timing_ && n_pass > 1;
      const int v0 = grouped ? 16 * g : g, gn = grouped ? std::min(16, nv - v0) : 1;
      if (sub) HIP_CHECK(hipEventRecord(vev_[4 * g], st_));
      for (int v = v0; v < v0 + gn; v++)
        K::fft_up(st_, n_vsrc_, v_src_fc_.p, eq_.p + v * eq_vs_, K0_, e0_, m_, lattice_.p, lattice_inv_.p, v_src_scal_.p,
                  grouped ? Xv_.p : X_.p, grouped ? 16 : 1, grouped ? v - v0 : 0,
                  split_ && !grouped ? vbound_.p : nullptr);
      if (sub) HIP_CHECK(hipEventRecord(vev_[4 * g + 1], st_));
      if (n_pass == 1) mark(6);
      if (grouped)
        K::hadamard_vec16(st_, n_vtrg_, M_dim_, v_partner_.p, Vb_.p, Xv_.p, Yv_.p, tuning.hadamard_vec_tpw);
      else
        K::hadamard_tiled(st_, vt_.n, M_dim_, e0_, e1_, n_off_, vt_.t0.p, vt_.u0.p, vt_.usrc.p, vt_.d0.p, vt_.dent.p,
                          vt_.uidx.p, vt_.gmask.p, v_offsets_.p, V_.p, X_.p, Y_.p, tuning.hadamard_tpb,
                          split_ ? &hparts(vrscale_) : nullptr);
      if (sub) HIP_CHECK(hipEventRecord(vev_[4 * g + 2], st_));
      if (n_pass == 1) mark(7);
      for (int v = v0; v < v0 + gn; v++)
        K::ifft_down(st_, n_vtrg_, v_trg_fc_.p, grouped ? Yv_.p : Y_.p, dn_.p + v * dn_vs_, K1_, e1_, m_, lattice_.p,
                     lattice_inv_.p, ns_, v_trg_scal_.p, grouped ? 16 : 1, grouped ? v - v0 : 0);
      if (sub) HIP_CHECK(hipEventRecord(vev_[4 * g + 3], st_));
    }
    if (n_pass > 1) {
      mark(6);
      mark(7);
    }
    mark(8);
    // D2D, top-down: 8 children rows += parent x [D2D_0 .. D2D_7]
    for (int l = 1; l <= max_depth_; l++) {
      const LevelRows& L = *l2l_[l - 1];
      if (d2d_fact_ && L.n >= kFactorRows) {  // parent check -> rank coordinates -> [8 children]
        K::gemm_rows(st_, L.n, rd_, K1_, dn_.p, L.ra.p, K1_, dc2de0_.p, ctmp_.p, lvl_seq_.p, rd_, nullptr, false,
                     gemm_work_.p, gemm_work_elems_, sp, nv, dn_vs_, ctmp_vs_);
        K::gemm_rows(st_, L.n, 8 * K1_, rd_, ctmp_.p, lvl_seq_.p, rd_, d2dr_.p, dn_.p, L.rc.p, K1_, nullptr, true,
                     gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, dn_vs_);
      } else {
        K::gemm_rows(st_, L.n, 8 * K1_, K1_, dn_.p, L.ra.p, K1_, d2d_.p, dn_.p, L.rc.p, K1_, nullptr, true, gemm_work_.p,
                     gemm_work_elems_, sp, nv, dn_vs_, dn_vs_);
      }
    }
    mark(9);
    // D2T: check -> equivalent
    K::gemm_rows(st_, n_l2t_, rd_, K1_, dn_.p, l2t_ra_.p, K1_, dc2de0_.p, ctmp_.p, l2t_seq_.p, rd_, l2t_alpha_.p, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, dn_vs_, ctmp_vs_);
    K::gemm_rows(st_, n_l2t_, K0_, rd_, ctmp_.p, l2t_seq_.p, rd_, dc2de1_.p, eq_.p, l2t_rc_.p, K0_, nullptr, false,
                 gemm_work_.p, gemm_work_elems_, sp, nv, ctmp_vs_, eq_vs_);
    mark(10);
    // equivalent surfaces -> targets (own downward equivalent + W list)
    K::far_points(st_, pk, n_tleaf_, t_leaf_.p, far_ptr_.p, far_node_.p, far_den_.p, far_rad_.p, trg_begin_.p, trg_end_.p,
                  center_.p, node_size_.p, unit_surf_.p, ns_, tx_.p, n_trg_, eq_.p, tval_.p, false, nv, eq_vs_, val_vs);
    mark(11);
    // U list: near-field P2P
    K::p2p_points(st_, pk, n_tleaf_, t_leaf_.p, u_ptr_.p, u_items_.p, trg_begin_.p, trg_end_.p, src_begin_.p, src_end_.p,
                  center_.p, tx64_.p, n_trg_, sx64_.p, sden_.p, n_src_, tval_.p, true, nv, den_vs, val_vs);
    if (vol_)  // minus the root's total equivalent density times the analytic volume potential
      K::vol_correction(st_, pk, n_trg_, txa_.p, eq_.p, K0_, e0_, tval_.p, nv, eq_vs_, val_vs);
    mark(12);
    // targets not in any leaf with targets are impossible; every target lies in exactly one leaf
    K::scatter_output(st_, n_trg_, k1_, tval_.p, trg_perm_.p, out_scale, d_out, nv, val_vs);
    mark(13);
    if (timing_) {
      HIP_CHECK(hipEventSynchronize(ev_[kPhaseCount]));
      for (int p = 0; p < kPhaseCount; p++) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev_[p], ev_[p + 1]));
        phase_ms[p] = ms;
      }
      if (n_pass > 1) {
        phase_ms[kPhFFT] = phase_ms[kPhHadamard] = phase_ms[kPhIFFT] = 0;
        for (int v = 0; v < n_pass; v++)
          for (int q = 0; q < 3; q++) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, vev_[4 * v + q], vev_[4 * v + q + 1]));
            phase_ms[kPhFFT + q] += ms;
          }
      }
    }
  }
};

}  // namespace

std::unique_ptr<DevicePlanBase> make_plan(const Octree& T, const Operators& O, const double* src_xyz, const double* trg_xyz,
                                          int precision, int n_vec_max, hipStream_t stream) {
  if (precision == 0) return std::make_unique<DevicePlan<double>>(T, O, src_xyz, trg_xyz, n_vec_max, stream);
  if (precision == 1) return std::make_unique<DevicePlan<float>>(T, O, src_xyz, trg_xyz, n_vec_max, stream);
  throw Error(kErrInvalidArgument, "precision must be 0 (fp64) or 1 (mixed fp32)");
}

}  // namespace pvfmm_amd
