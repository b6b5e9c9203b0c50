// Octree and interaction-list construction on the GPU (the setup step of PtFMM_CreateTree +
// FMM_Tree::SetupFMM: include/mpi_tree.txx:37-311 p2o / points2Octree, 314-492 refine and
// coarsen, 745-1066 balanceOctree, 1136-1255 SetColleagues; include/interac_list.txx:98-279
// BuildList).  It builds the same tree as the host builder in octree.cpp, with the same
// key-interval formulation mapped to data-parallel passes:
//   1. 90-bit Morton keys of every point set, one stable radix sort each (the point order);
//   2. refinement level by level: a node is split iff depth < 29 and its key interval holds
//      more than max_pts source + target points or more than max_pts tree-defining points
//      (SubdivCond, include/fmm_node.txx:62-95); child ranges by binary search;
//   3. 2:1 balance as the closure of {parents of leaves} under "parent of every same-level
//      neighbour", one sort + unique of candidate keys per level, deepest level first;
//   4. level-major node arrays, subtree point ranges, colleagues and the lists, one thread
//      per node with binary searches in the per-level sorted key arrays.
// Every step is integer arithmetic, so the result equals the host build bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "device_fmm.hpp"
#include "gpu_octree.hpp"

namespace pvfmm_amd {

using u128 = rocprim::uint128_t;

#define OCT_CHECK() HIP_CHECK(hipGetLastError())

namespace {

constexpr int kDepthLimit = kMaxDepth - 1;  // nodes at depth 29 are never split
constexpr u128 kNoKey = ~(u128)0;           // sorts after every node key

__host__ __device__ __forceinline__ uint64_t spread21(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}
__host__ __device__ __forceinline__ uint64_t compact21(uint64_t v) {
  v &= 0x1249249249249249ull;
  v = (v | v >> 2) & 0x10c30c30c30c30c3ull;
  v = (v | v >> 4) & 0x100f00f00f00f00full;
  v = (v | v >> 8) & 0x1f0000ff0000ffull;
  v = (v | v >> 16) & 0x1f00000000ffffull;
  v = (v | v >> 32) & 0x1fffff;
  return v;
}
// bit triplets (z, y, x) from the most significant level down: MortonId::operator<
// (include/mortonid.txx:118-130)
__host__ __device__ __forceinline__ u128 morton(uint32_t x, uint32_t y, uint32_t z) {
  const uint64_t lo = spread21(x) | spread21(y) << 1 | spread21(z) << 2;
  const uint64_t hi = spread21(x >> 21) | spread21(y >> 21) << 1 | spread21(z >> 21) << 2;
  return ((u128)hi << 63) | (u128)lo;
}
__device__ __forceinline__ void unmorton(u128 k, uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint64_t lo = (uint64_t)k & ((1ull << 63) - 1), hi = (uint64_t)(k >> 63);
  x = (uint32_t)(compact21(lo) | compact21(hi) << 21);
  y = (uint32_t)(compact21(lo >> 1) | compact21(hi >> 1) << 21);
  z = (uint32_t)(compact21(lo >> 2) | compact21(hi >> 2) << 21);
}
// first position in [a, b) with keys[pos] >= k
__device__ __forceinline__ int lower(const u128* __restrict__ keys, int a, int b, u128 k) {
  while (a < b) {
    const int m = a + ((b - a) >> 1);
    if (keys[m] < k) a = m + 1;
    else b = m;
  }
  return a;
}
// index of k in the sorted keys[0, n), or -1
__device__ __forceinline__ int find(const u128* __restrict__ keys, int n, u128 k) {
  const int p = lower(keys, 0, n, k);
  return (p < n && keys[p] == k) ? p : -1;
}
__device__ __forceinline__ u128 node_span(int d) { return (u128)1 << (3 * (kMaxDepth - d)); }

template <typename T>
struct DVec {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  DVec() = default;
  DVec(const DVec&) = delete;
  DVec& operator=(const DVec&) = delete;
  DVec(DVec&& o) noexcept : p(o.p), n(o.n), cap(o.cap) { o.p = nullptr; o.n = o.cap = 0; }
  DVec& operator=(DVec&& o) noexcept {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(cap, o.cap);
    return *this;
  }
  ~DVec() { if (p) (void)hipFree(p); }
  void resize(size_t m) {
    if (m > cap) {
      if (p) {
        HIP_CHECK(hipDeviceSynchronize());  // queued work may still read the old buffer
        (void)hipFree(p);
      }
      p = nullptr;
      HIP_CHECK(hipMalloc(&p, sizeof(T) * std::max<size_t>(m, 1)));
      cap = std::max<size_t>(m, 1);
    }
    n = m;
  }
};

// ----------------------------------------------------------------------------- kernels
// MortonId(x): floor(x * 2^30) per axis (include/mortonid.txx:26-38)
__global__ void k_point_keys(const double* __restrict__ X, int64_t n, u128* __restrict__ keys, int32_t* __restrict__ idx) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double s = 1073741824.0;
    keys[i] = morton((uint32_t)floor(X[3 * i] * s), (uint32_t)floor(X[3 * i + 1] * s), (uint32_t)floor(X[3 * i + 2] * s));
    idx[i] = (int32_t)i;
  }
}

struct Sorted {
  const u128* k;  // sorted keys of one point set
};

// frontier node i of depth d: corner key fk[i], point ranges r[6 i .. 6 i + 6) (src, trg, pt)
__global__ void k_split_flags(int n, const int32_t* __restrict__ r, int max_pts, int d, int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c_st = (int64_t)(r[6 * i + 1] - r[6 * i]) + (r[6 * i + 3] - r[6 * i + 2]);
  const int64_t c_pt = r[6 * i + 5] - r[6 * i + 4];
  flag[i] = (d < kDepthLimit && (c_st > max_pts || c_pt > max_pts)) ? 1 : 0;
}

__global__ void k_split_emit(int n, const u128* __restrict__ fk, const int32_t* __restrict__ r,
                             const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int d, Sorted s0,
                             Sorted s1, Sorted s2, u128* __restrict__ internal, u128* __restrict__ nk,
                             int32_t* __restrict__ nr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int p = pos[i];
  const u128 base = fk[i];
  internal[p] = base;
  const int sh = 3 * (kMaxDepth - 1 - d);  // key bits below one child of a depth-d node
  const u128* keys[3] = {s0.k, s1.k, s2.k};
  int b[3][9];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    b[s][0] = r[6 * i + 2 * s];
    b[s][8] = r[6 * i + 2 * s + 1];
    for (int k = 1; k < 8; k++) b[s][k] = lower(keys[s], b[s][k - 1], b[s][8], base + ((u128)k << sh));
  }
  for (int k = 0; k < 8; k++) {
    const int c = 8 * p + k;
    nk[c] = base + ((u128)k << sh);
#pragma unroll
    for (int s = 0; s < 3; s++) {
      nr[6 * c + 2 * s] = b[s][k];
      nr[6 * c + 2 * s + 1] = b[s][k + 1];
    }
  }
}

// internal nodes of depth d with at least one child that is not internal
__global__ void k_leaf_parent_flags(int n, const u128* __restrict__ internal, const u128* __restrict__ below, int n_below,
                                    int d, int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sh = 3 * (kMaxDepth - 1 - d);
  int f = 0;
  for (int k = 0; k < 8 && !f; k++)
    if (find(below, n_below, internal[i] + ((u128)k << sh)) < 0) f = 1;
  flag[i] = f;
}

__global__ void k_gather_flagged(int n, const u128* __restrict__ in, const int32_t* __restrict__ flag,
                                 const int32_t* __restrict__ pos, u128* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = in[i];
}

// the depth-(l-1) parents of the 27 same-level neighbours (self included) of each node of S_l;
// out of the unit cube: kNoKey (a periodic tree wraps every axis, as the reference's balance does)
__global__ void k_neighbour_parents(int n, const u128* __restrict__ S, int l, int periodic, u128* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * 27) return;
  const int i = (int)(e / 27), c = (int)(e % 27);
  uint32_t x, y, z;
  unmorton(S[i], x, y, z);
  const int64_t step = (int64_t)1 << (kMaxDepth - l), maxc = (int64_t)1 << kMaxDepth;
  int64_t nx = (int64_t)x + (c % 3 - 1) * step, ny = (int64_t)y + ((c / 3) % 3 - 1) * step,
          nz = (int64_t)z + (c / 9 - 1) * step;
  if (periodic) {
    nx = (nx + maxc) % maxc; ny = (ny + maxc) % maxc; nz = (nz + maxc) % maxc;
  }
  if (nx < 0 || ny < 0 || nz < 0 || nx >= maxc || ny >= maxc || nz >= maxc) {
    out[e] = kNoKey;
    return;
  }
  const uint32_t pmask = ~(uint32_t)((((uint64_t)1) << (kMaxDepth - (l - 1))) - 1);
  out[e] = morton((uint32_t)nx & pmask, (uint32_t)ny & pmask, (uint32_t)nz & pmask);
}

// nodes of level d in Morton order: inner flag from the balanced internal set S_d
__global__ void k_inner_flags(int n, const u128* __restrict__ keys, const u128* __restrict__ S, int n_s, int d,
                              int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (d < kDepthLimit && find(S, n_s, keys[i]) >= 0) ? 1 : 0;
}

__global__ void k_emit_level(int n, const u128* __restrict__ keys, const int32_t* __restrict__ flag,
                             const int32_t* __restrict__ pos, int d, int base, int next_base, u128* __restrict__ nkeys,
                             int32_t* __restrict__ first_child, int32_t* __restrict__ nparent, int32_t* __restrict__ npos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!flag[i]) {
    first_child[i] = -1;
    return;
  }
  const int p = pos[i];
  first_child[i] = next_base + 8 * p;
  const int sh = 3 * (kMaxDepth - 1 - d);
  for (int k = 0; k < 8; k++) {
    nkeys[8 * p + k] = keys[i] + ((u128)k << sh);
    nparent[8 * p + k] = base + i;
    npos[8 * p + k] = k;
  }
}

// per node: corner, subtree point ranges (key interval [key, key + span)) and the 27 colleagues
struct LevelTable {
  const u128* keys;   // all nodes, level-major; each level sorted
  const int32_t* lb;  // level_begin, max_depth + 2 entries
  int n_levels;
};

// colleagues wrap along the periodic axes of bndry (1: x; 2: x, y; 3: all)
__global__ void k_node_data(int N, LevelTable L, const uint8_t* __restrict__ depth, const u128* __restrict__ sk, int n_s,
                            const u128* __restrict__ tk, int n_t, int bndry, uint32_t* __restrict__ xyz,
                            int32_t* __restrict__ rng, int32_t* __restrict__ coll) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int d = depth[i];
  const u128 k = L.keys[i], k1 = k + node_span(d);
  uint32_t x, y, z;
  unmorton(k, x, y, z);
  xyz[3 * i] = x;
  xyz[3 * i + 1] = y;
  xyz[3 * i + 2] = z;
  rng[4 * i] = lower(sk, 0, n_s, k);
  rng[4 * i + 1] = lower(sk, 0, n_s, k1);
  rng[4 * i + 2] = lower(tk, 0, n_t, k);
  rng[4 * i + 3] = lower(tk, 0, n_t, k1);
  const int64_t step = (int64_t)1 << (kMaxDepth - d), maxc = (int64_t)1 << kMaxDepth;
  const int a = L.lb[d], nl = L.lb[d + 1] - a;
  for (int c = 0; c < 27; c++) {
    int64_t nx = (int64_t)x + (c % 3 - 1) * step, ny = (int64_t)y + ((c / 3) % 3 - 1) * step,
            nz = (int64_t)z + (c / 9 - 1) * step;
    if (bndry >= 1) nx = (nx + maxc) % maxc;
    if (bndry >= 2) ny = (ny + maxc) % maxc;
    if (bndry >= 3) nz = (nz + maxc) % maxc;
    int v = -1;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < maxc && ny < maxc && nz < maxc) {
      const int f = find(L.keys + a, nl, morton((uint32_t)nx, (uint32_t)ny, (uint32_t)nz));
      if (f >= 0) v = a + f;
    }
    coll[27 * i + c] = v;
  }
}

// Relative-coordinate lookup tables of the list types, coord_hash -> slot or -1
// (include/interac_list.txx:343-404)
struct ListOut {
  int32_t* out[kTypeCount];
  int count[kTypeCount];
  const int32_t* lut[kTypeCount];  // 2000 entries each
};

__device__ __forceinline__ void put(const ListOut& L, int t, int64_t n, int rx, int ry, int rz, int32_t v) {
  if (rx < -5 || rx > 5 || ry < -5 || ry > 5 || rz < -5 || rz > 5) return;
  const int idx = L.lut[t][((rz + 5) * 10 + (ry + 5)) * 10 + (rx + 5)];
  if (idx >= 0) L.out[t][n * L.count[t] + idx] = v;
}

// the lists of include/interac_list.txx:98-279 for nodes holding sources (S2U, U2U) or
// targets (the rest), as the host builder forms them from colleagues
__global__ void k_lists(int N, const int32_t* __restrict__ parent, const int32_t* __restrict__ first_child,
                        const int32_t* __restrict__ child_pos, const int32_t* __restrict__ rng,
                        const int32_t* __restrict__ coll, ListOut L) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const bool has_src = rng[4 * n + 1] > rng[4 * n], has_trg = rng[4 * n + 3] > rng[4 * n + 2];
  const bool leaf = first_child[n] < 0;
  const int32_t p = parent[n];
  const int pos = child_pos[n];
  const int bx = pos & 1, by = (pos >> 1) & 1, bz = (pos >> 2) & 1;
  if (has_src) {
    if (leaf) L.out[kS2U][n] = n;
    else
      for (int j = 0; j < 8; j++) put(L, kU2U, n, (j & 1) ? 1 : -1, (j & 2) ? 1 : -1, (j & 4) ? 1 : -1, first_child[n] + j);
  }
  if (!has_trg) return;
  if (leaf) L.out[kD2T][n] = n;
  if (p >= 0) {
    put(L, kD2D, n, bx ? 1 : -1, by ? 1 : -1, bz ? 1 : -1, p);
    for (int i = 0; i < 27; i++) {
      const int32_t pc = coll[27 * p + i];
      if (pc < 0 || first_child[pc] >= 0) continue;
      const int rx = (i % 3) * 4 - 4 - 2 * bx + 1, ry = ((i / 3) % 3) * 4 - 4 - 2 * by + 1, rz = (i / 9) * 4 - 4 - 2 * bz + 1;
      if (leaf) put(L, kU0, n, rx, ry, rz, pc);
      put(L, kX, n, rx, ry, rz, pc);
    }
  }
  for (int i = 0; i < 27; i++) {
    const int32_t c = coll[27 * n + i];
    if (c < 0) continue;
    const int dx = i % 3 - 1, dy = (i / 3) % 3 - 1, dz = i / 9 - 1;
    if (leaf) {
      if (first_child[c] < 0) {
        put(L, kU1, n, dx, dy, dz, c);
      } else {
        for (int j = 0; j < 8; j++) {
          const int rx = dx * 4 + ((j & 1) ? 2 : 0) - 1, ry = dy * 4 + ((j & 2) ? 2 : 0) - 1, rz = dz * 4 + ((j & 4) ? 2 : 0) - 1;
          put(L, kU2, n, rx, ry, rz, first_child[c] + j);
          put(L, kW, n, rx, ry, rz, first_child[c] + j);
        }
      }
    } else if (first_child[c] >= 0) {
      put(L, kV1, n, dx, dy, dz, c);
    }
  }
}

template <typename T>
__global__ void k_fill_value(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

// ----------------------------------------------------------------------------- library calls
struct Temp {
  DVec<char> buf;
  void* get(size_t bytes) {
    buf.resize(bytes);
    return buf.p;
  }
};

// exclusive scan of flag[0, n) into pos; returns the total
int scan_flags(hipStream_t s, Temp& tmp, const int32_t* flag, int32_t* pos, int n) {
  if (n == 0) return 0;
  size_t bytes = 0;
  HIP_CHECK(rocprim::exclusive_scan(nullptr, bytes, flag, pos, 0, (size_t)n, rocprim::plus<int32_t>(), s));
  HIP_CHECK(rocprim::exclusive_scan(tmp.get(bytes), bytes, flag, pos, 0, (size_t)n, rocprim::plus<int32_t>(), s));
  int32_t last[2] = {0, 0};
  HIP_CHECK(hipMemcpyAsync(&last[0], pos + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&last[1], flag + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return last[0] + last[1];
}

// sorted unique node keys of `in` (n keys; kNoKey entries dropped) into `out`
int sort_unique(hipStream_t s, Temp& tmp, DVec<u128>& in, int n, DVec<u128>& out) {
  if (n == 0) {
    out.resize(0);
    return 0;
  }
  DVec<u128> sorted;
  sorted.resize(n);
  size_t bytes = 0;
  HIP_CHECK(rocprim::radix_sort_keys(nullptr, bytes, in.p, sorted.p, (size_t)n, 0, 128, s));
  HIP_CHECK(rocprim::radix_sort_keys(tmp.get(bytes), bytes, in.p, sorted.p, (size_t)n, 0, 128, s));
  out.resize(n);
  DVec<size_t> cnt;
  cnt.resize(1);
  bytes = 0;
  HIP_CHECK(rocprim::unique(nullptr, bytes, sorted.p, out.p, cnt.p, (size_t)n, rocprim::equal_to<u128>(), s));
  HIP_CHECK(rocprim::unique(tmp.get(bytes), bytes, sorted.p, out.p, cnt.p, (size_t)n, rocprim::equal_to<u128>(), s));
  size_t m = 0;
  HIP_CHECK(hipMemcpyAsync(&m, cnt.p, sizeof(size_t), hipMemcpyDeviceToHost, s));
  u128 last = 0;
  HIP_CHECK(hipStreamSynchronize(s));
  if (m > 0) {
    HIP_CHECK(hipMemcpyAsync(&last, out.p + m - 1, sizeof(u128), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (last == kNoKey) m--;
  }
  out.n = m;
  return (int)m;
}

// sorted keys and the stable sort permutation of one point set
struct SortedSet {
  DVec<u128> keys;
  std::vector<int64_t> perm;
  int n = 0;
};

void sort_set(hipStream_t s, Temp& tmp, const double* X, int64_t n, SortedSet& S) {
  S.n = (int)n;
  S.keys.resize(n);
  S.perm.assign(n, 0);
  if (n == 0) return;
  DVec<double> dx;
  dx.resize(3 * n);
  HIP_CHECK(hipMemcpyAsync(dx.p, X, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  DVec<u128> k0;
  DVec<int32_t> i0, i1;
  k0.resize(n);
  i0.resize(n);
  i1.resize(n);
  hipLaunchKernelGGL(k_point_keys, dim3(std::min(blocks_for(n), 8192u)), dim3(256), 0, s, dx.p, n, k0.p, i0.p);
  OCT_CHECK();
  size_t bytes = 0;
  HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, k0.p, S.keys.p, i0.p, i1.p, (size_t)n, 0, 90, s));
  HIP_CHECK(rocprim::radix_sort_pairs(tmp.get(bytes), bytes, k0.p, S.keys.p, i0.p, i1.p, (size_t)n, 0, 90, s));
  std::vector<int32_t> p32(n);
  HIP_CHECK(hipMemcpyAsync(p32.data(), i1.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  for (int64_t i = 0; i < n; i++) S.perm[i] = p32[i];
}

}  // namespace

bool device_tree_enabled() {
  const char* e = std::getenv("PVFMM_TREE");
  if (e && std::string(e) == "host") return false;
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

void build_octree_device(Octree& T, const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt,
                         int64_t n_pt, int max_pts, int bndry) {
  if (max_pts <= 0) throw Error(kErrInvalidArgument, "max_pts must be positive");
  if (bndry < kFreeSpace || bndry > kPXYZ) throw Error(kErrInvalidArgument, "boundary type must be 0 (free), 1 (PX), 2 (PXY) or 3 (PXYZ)");
  if (n_src >= (1ll << 31) || n_trg >= (1ll << 31) || n_pt >= (1ll << 31))
    throw Error(kErrUnsupported, "more than 2^31 points");
  for (const auto& set : {std::make_pair(src, n_src), std::make_pair(trg, n_trg), std::make_pair(pt, n_pt)})
    for (int64_t i = 0; i < 3 * set.second; i++)
      if (!(set.first[i] >= 0.0 && set.first[i] < 1.0)) throw Error(kErrInvalidArgument, "point coordinates must lie in [0,1)");
  T = Octree();
  T.max_pts = max_pts;
  T.bndry = bndry;
  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  std::unique_ptr<std::remove_pointer<hipStream_t>::type, hipError_t (*)(hipStream_t)> guard(s, &hipStreamDestroy);
  Temp tmp;

  // 1. point keys and orders (the tree-defining set is the targets unless given separately)
  SortedSet S0, S1, S2;
  sort_set(s, tmp, src, n_src, S0);
  sort_set(s, tmp, trg, n_trg, S1);
  const bool pt_is_trg = (pt == trg && n_pt == n_trg);
  if (!pt_is_trg) sort_set(s, tmp, pt, n_pt, S2);
  const SortedSet& SP = pt_is_trg ? S1 : S2;
  T.src.n = n_src;
  T.trg.n = n_trg;
  T.src.perm = S0.perm;
  T.trg.perm = S1.perm;

  // 2. refinement by counts, level by level
  std::vector<DVec<u128>> internal(kMaxDepth + 1);
  std::vector<int> n_int(kMaxDepth + 1, 0);
  DVec<u128> fk, nk;
  DVec<int32_t> fr, nr, flag, pos;
  fk.resize(1);
  fr.resize(6);
  {
    const u128 root = 0;
    const int32_t r0[6] = {0, (int32_t)n_src, 0, (int32_t)n_trg, 0, (int32_t)n_pt};
    HIP_CHECK(hipMemcpyAsync(fk.p, &root, sizeof(u128), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(fr.p, r0, sizeof(r0), hipMemcpyHostToDevice, s));
  }
  int nf = 1;
  for (int d = 0; d < kDepthLimit && nf > 0; d++) {
    flag.resize(nf);
    pos.resize(nf);
    hipLaunchKernelGGL(k_split_flags, dim3(blocks_for(nf)), dim3(256), 0, s, nf, fr.p, max_pts, d, flag.p);
    OCT_CHECK();
    const int ni = scan_flags(s, tmp, flag.p, pos.p, nf);
    n_int[d] = ni;
    internal[d].resize(ni);
    nk.resize((size_t)8 * ni);
    nr.resize((size_t)48 * ni);
    if (ni > 0) {
      hipLaunchKernelGGL(k_split_emit, dim3(blocks_for(nf)), dim3(256), 0, s, nf, fk.p, fr.p, flag.p, pos.p, d,
                         Sorted{S0.keys.p}, Sorted{S1.keys.p}, Sorted{SP.keys.p}, internal[d].p, nk.p, nr.p);
      OCT_CHECK();
    }
    std::swap(fk, nk);
    std::swap(fr, nr);
    nf = 8 * ni;
  }

  // 3. 2:1 balance: parents of leaves, closed under parents of same-level neighbours
  std::vector<DVec<u128>> S(kMaxDepth + 1);
  std::vector<int> n_S(kMaxDepth + 1, 0);
  for (int d = 0; d < kMaxDepth; d++) {
    const int n = n_int[d];
    if (n == 0) continue;
    flag.resize(n);
    pos.resize(n);
    hipLaunchKernelGGL(k_leaf_parent_flags, dim3(blocks_for(n)), dim3(256), 0, s, n, internal[d].p, internal[d + 1].p,
                       n_int[d + 1], d, flag.p);
    OCT_CHECK();
    const int m = scan_flags(s, tmp, flag.p, pos.p, n);
    S[d].resize(m);
    if (m > 0) {
      hipLaunchKernelGGL(k_gather_flagged, dim3(blocks_for(n)), dim3(256), 0, s, n, internal[d].p, flag.p, pos.p, S[d].p);
      OCT_CHECK();
    }
    n_S[d] = m;
  }
  for (int l = kMaxDepth; l >= 1; l--) {
    if (n_S[l] == 0) continue;
    const int n = n_S[l], old = n_S[l - 1];
    DVec<u128> cand;
    cand.resize((size_t)27 * n + old);
    hipLaunchKernelGGL(k_neighbour_parents, dim3(blocks_for((int64_t)27 * n)), dim3(256), 0, s, n, S[l].p, l,
                       bndry != kFreeSpace ? 1 : 0, cand.p);
    OCT_CHECK();
    if (old > 0) HIP_CHECK(hipMemcpyAsync(cand.p + (size_t)27 * n, S[l - 1].p, sizeof(u128) * old, hipMemcpyDeviceToDevice, s));
    n_S[l - 1] = sort_unique(s, tmp, cand, 27 * n + old, S[l - 1]);
  }

  // 4. level-major node arrays: the children of each internal node in Morton order
  std::vector<u128> hkeys{(u128)0};
  T.level_begin.assign(1, 0);
  DVec<u128> lk, nlk;
  DVec<int32_t> fc, npar, npos;
  lk.resize(1);
  HIP_CHECK(hipMemcpyAsync(lk.p, &hkeys[0], sizeof(u128), hipMemcpyHostToDevice, s));
  T.parent.assign(1, -1);
  T.child_pos.assign(1, 0);
  int nl = 1, d = 0;
  std::vector<u128> all_keys;
  while (nl > 0) {
    const int base = (int)T.depth.size();
    flag.resize(nl);
    pos.resize(nl);
    hipLaunchKernelGGL(k_inner_flags, dim3(blocks_for(nl)), dim3(256), 0, s, nl, lk.p, S[d].p, n_S[d], d, flag.p);
    OCT_CHECK();
    const int ni = scan_flags(s, tmp, flag.p, pos.p, nl);
    fc.resize(nl);
    nlk.resize((size_t)8 * ni);
    npar.resize((size_t)8 * ni);
    npos.resize((size_t)8 * ni);
    hipLaunchKernelGGL(k_emit_level, dim3(blocks_for(nl)), dim3(256), 0, s, nl, lk.p, flag.p, pos.p, d, base, base + nl,
                       nlk.p, fc.p, npar.p, npos.p);
    OCT_CHECK();
    std::vector<u128> keys_h(nl);
    std::vector<int32_t> fc_h(nl), par_h(8 * ni), pos_h(8 * ni);
    HIP_CHECK(hipMemcpyAsync(keys_h.data(), lk.p, sizeof(u128) * nl, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(fc_h.data(), fc.p, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, s));
    if (ni) {
      HIP_CHECK(hipMemcpyAsync(par_h.data(), npar.p, sizeof(int32_t) * 8 * ni, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(pos_h.data(), npos.p, sizeof(int32_t) * 8 * ni, hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    all_keys.insert(all_keys.end(), keys_h.begin(), keys_h.end());
    T.first_child.insert(T.first_child.end(), fc_h.begin(), fc_h.end());
    T.depth.insert(T.depth.end(), (size_t)nl, (uint8_t)d);
    T.level_begin.push_back(base + nl);
    T.parent.insert(T.parent.end(), par_h.begin(), par_h.end());
    T.child_pos.insert(T.child_pos.end(), pos_h.begin(), pos_h.end());
    std::swap(lk, nlk);
    nl = 8 * ni;
    d++;
  }
  const int32_t N = T.n_nodes();
  T.max_depth = (int)T.level_begin.size() - 2;

  // 5. corners, subtree point ranges and colleagues
  DVec<u128> dkeys;
  DVec<int32_t> dlb, dpar, dfc, dpos, drng, dcoll;
  DVec<uint8_t> ddep;
  DVec<uint32_t> dxyz;
  dkeys.resize(N);
  dlb.resize(T.level_begin.size());
  ddep.resize(N);
  HIP_CHECK(hipMemcpyAsync(dkeys.p, all_keys.data(), sizeof(u128) * N, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dlb.p, T.level_begin.data(), sizeof(int32_t) * T.level_begin.size(), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(ddep.p, T.depth.data(), N, hipMemcpyHostToDevice, s));
  dxyz.resize((size_t)3 * N);
  drng.resize((size_t)4 * N);
  dcoll.resize((size_t)27 * N);
  hipLaunchKernelGGL(k_node_data, dim3(blocks_for(N)), dim3(256), 0, s, N,
                     LevelTable{dkeys.p, dlb.p, (int)T.level_begin.size() - 1}, ddep.p, S0.keys.p, S0.n, S1.keys.p, S1.n,
                     bndry, dxyz.p, drng.p, dcoll.p);
  OCT_CHECK();

  // 6. interaction lists
  dpar.resize(N);
  dfc.resize(N);
  dpos.resize(N);
  HIP_CHECK(hipMemcpyAsync(dpar.p, T.parent.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dfc.p, T.first_child.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dpos.p, T.child_pos.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
  const int types[] = {kS2U, kU2U, kD2D, kD2T, kU0, kU1, kU2, kW, kX, kV1};
  std::vector<DVec<int32_t>> dl(kTypeCount), dlut(kTypeCount);
  ListOut LO{};
  for (int t : types) {
    const int cnt = T.list_count(t);
    dl[t].resize((size_t)N * cnt);
    dlut[t].resize(rel_table(t).lut.size());
    HIP_CHECK(hipMemcpyAsync(dlut[t].p, rel_table(t).lut.data(), sizeof(int32_t) * rel_table(t).lut.size(),
                             hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_value<int32_t>, dim3(std::min(blocks_for((int64_t)N * cnt), 8192u)), dim3(256), 0, s, dl[t].p,
                       (int64_t)N * cnt, (int32_t)-1);
    OCT_CHECK();
    LO.out[t] = dl[t].p;
    LO.count[t] = cnt;
    LO.lut[t] = dlut[t].p;
  }
  hipLaunchKernelGGL(k_lists, dim3(blocks_for(N)), dim3(256), 0, s, N, dpar.p, dfc.p, dpos.p, drng.p, dcoll.p, LO);
  OCT_CHECK();

  // 7. back to the host structure
  std::vector<uint32_t> xyz((size_t)3 * N);
  std::vector<int32_t> rng((size_t)4 * N);
  HIP_CHECK(hipMemcpyAsync(xyz.data(), dxyz.p, sizeof(uint32_t) * xyz.size(), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(rng.data(), drng.p, sizeof(int32_t) * rng.size(), hipMemcpyDeviceToHost, s));
  for (int t : types) {
    T.lists[t].resize((size_t)N * T.list_count(t));
    HIP_CHECK(hipMemcpyAsync(T.lists[t].data(), dl[t].p, sizeof(int32_t) * T.lists[t].size(), hipMemcpyDeviceToHost, s));
  }
  HIP_CHECK(hipStreamSynchronize(s));
  T.x.resize(N); T.y.resize(N); T.z.resize(N);
  T.src_begin.resize(N); T.src_end.resize(N); T.trg_begin.resize(N); T.trg_end.resize(N);
  for (int32_t i = 0; i < N; i++) {
    T.x[i] = xyz[3 * i]; T.y[i] = xyz[3 * i + 1]; T.z[i] = xyz[3 * i + 2];
    T.src_begin[i] = rng[4 * i]; T.src_end[i] = rng[4 * i + 1];
    T.trg_begin[i] = rng[4 * i + 2]; T.trg_end[i] = rng[4 * i + 3];
  }
}

}  // namespace pvfmm_amd
