// Host octree construction.  Design: points become 90-bit Morton keys, are sorted once,
// and every node is a key interval; the adaptive refinement, the 2:1 balance and the
// neighbour search are all key arithmetic plus hash lookups, producing flat level-major
// arrays that are copied to the GPU as-is.
//
// Equivalence with the reference (checked bit-exactly against the oracle in the tests):
//  * RefineTree/CoarsenTree (include/mpi_tree.txx:314-492) converge to the unique tree in
//    which a node is split iff depth < 29 and its subtree holds more than max_pts
//    src+trg points or more than max_pts tree-defining points (SubdivCond,
//    include/fmm_node.txx:62-95, include/mpi_node.txx:79-105).
//  * balanceOctree (include/mpi_tree.txx:745-943) makes internal exactly the closure of
//    {parents of leaves} under "parent of every same-level neighbour (self included)".
//  * colleagues (include/mpi_tree.txx:1136-1255) are the same-level nodes at offsets in
//    {-1,0,1}^3; the lists follow include/interac_list.txx:98-279.
#include "octree.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <unordered_set>

#include <omp.h>

namespace pvfmm_amd {

using u128 = unsigned __int128;

namespace {

struct U128Hash {
  size_t operator()(const u128& k) const {
    uint64_t a = (uint64_t)k, b = (uint64_t)(k >> 64);
    uint64_t h = a * 0x9E3779B97F4A7C15ull ^ (b + 0x7F4A7C159E3779B9ull + (a << 6) + (a >> 2));
    return (size_t)(h ^ (h >> 31));
  }
};

// spread the low 21 bits of v to every third bit
inline uint64_t spread21(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}
// 90-bit Morton key; bit triplets (z,y,x) from the most significant level down, which is the
// order of MortonId::operator< (include/mortonid.txx:118-130).
inline u128 morton(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t lo = spread21(x) | spread21(y) << 1 | spread21(z) << 2;
  uint64_t hi = spread21(x >> 21) | spread21(y >> 21) << 1 | spread21(z >> 21) << 2;
  return ((u128)hi << 63) | (u128)lo;
}
// key of a node (corner + depth); nodes are packed as depth|x|y|z for hashing
inline u128 node_id(int d, uint32_t x, uint32_t y, uint32_t z) {
  return ((u128)d << 96) | ((u128)x << 64) | ((u128)y << 32) | (u128)z;
}

inline uint32_t to_int(double c) {
  // MortonId(x,y,z): floor(x * 2^30), include/mortonid.txx:26-38
  return (uint32_t)std::floor(c * (double)(1u << 30));
}

void sort_points(PointSet& P, const double* X, int64_t n) {
  P.n = n;
  for (int64_t i = 0; i < 3 * n; i++)
    if (!(X[i] >= 0.0 && X[i] < 1.0)) throw Error(kErrInvalidArgument, "point coordinates must lie in [0,1)");
  std::vector<std::pair<u128, int64_t>> key(n);
  #pragma omp parallel for
  for (int64_t i = 0; i < n; i++) {
    key[i] = {morton(to_int(X[3 * i]), to_int(X[3 * i + 1]), to_int(X[3 * i + 2])), i};
  }
  std::stable_sort(key.begin(), key.end(), [](const std::pair<u128, int64_t>& a, const std::pair<u128, int64_t>& b) { return a.first < b.first; });
  P.perm.resize(n); P.ix.resize(n); P.iy.resize(n); P.iz.resize(n);
  #pragma omp parallel for
  for (int64_t i = 0; i < n; i++) {
    int64_t j = key[i].second;
    P.perm[i] = j;
    P.ix[i] = to_int(X[3 * j]); P.iy[i] = to_int(X[3 * j + 1]); P.iz[i] = to_int(X[3 * j + 2]);
  }
}

struct KeyedSet {
  std::vector<u128> keys;  // sorted Morton keys
  // first position with key >= k
  int64_t lower(u128 k, int64_t a, int64_t b) const {
    return std::lower_bound(keys.begin() + a, keys.begin() + b, k) - keys.begin();
  }
};

KeyedSet keyed(const PointSet& P) {
  KeyedSet K;
  K.keys.resize(P.n);
  #pragma omp parallel for
  for (int64_t i = 0; i < P.n; i++) K.keys[i] = morton(P.ix[i], P.iy[i], P.iz[i]);
  return K;
}

}  // namespace

// ----------------------------------------------------------------------------- list tables
static RelTable make_table(int max_r, int min_r, int step) {
  // Same enumeration and class grouping as the reference's InitList, so that list slot i
  // means the same relative coordinate in both (include/interac_list.txx:367-404).
  auto chash = [](int a, int b, int c) {
    int v[3] = {std::abs(a), std::abs(b), std::abs(c)};
    std::sort(v, v + 3);
    return RelTable::hash(v[0], v[1], v[2]);
  };
  std::vector<std::array<int, 3>> all;
  for (int k = -max_r; k <= max_r; k += step)
    for (int j = -max_r; j <= max_r; j += step)
      for (int i = -max_r; i <= max_r; i += step)
        if (std::abs(i) >= min_r || std::abs(j) >= min_r || std::abs(k) >= min_r) all.push_back({i, j, k});
  // stable order by class hash keeps the loop order within a class
  std::stable_sort(all.begin(), all.end(), [&](const std::array<int, 3>& a, const std::array<int, 3>& b) {
    return chash(a[0], a[1], a[2]) < chash(b[0], b[1], b[2]);
  });
  RelTable T;
  T.rel = all;
  T.lut.assign(2000, -1);  // coord_hash of [-5,5]^3 stays below 1111
  for (size_t i = 0; i < all.size(); i++) T.lut[RelTable::hash(all[i][0], all[i][1], all[i][2])] = (int)i;
  return T;
}

const RelTable& rel_table(int type) {
  static const std::vector<RelTable> tables = [] {
    std::vector<RelTable> t(kTypeCount);
    t[kS2U] = make_table(0, 0, 1); t[kU2U] = make_table(1, 1, 2); t[kD2D] = make_table(1, 1, 2);
    t[kD2T] = make_table(0, 0, 1); t[kU0] = make_table(3, 3, 2); t[kU1] = make_table(1, 0, 1);
    t[kU2] = make_table(3, 3, 2); t[kV] = make_table(3, 2, 1); t[kV1] = make_table(1, 1, 1);
    t[kW] = make_table(5, 5, 2); t[kX] = make_table(5, 5, 2);
    for (int i = 0; i < 4; i++) t[i] = make_table(0, 0, 1);
    t[15] = make_table(0, 0, 1);
    return t;
  }();
  return tables.at(type);
}

// ----------------------------------------------------------------------------- tree build
void image_shift(const Octree& T, int32_t t, int32_t s, int type, int k, int* shift) {
  const std::array<int, 3>& rel = rel_table(type).rel[k];
  // in units of 2^-32 of the period: node sizes are at least 2^-29, so a quarter of one is whole
  const int64_t st = (int64_t)4 << (30 - T.depth[t]), ss = (int64_t)4 << (30 - T.depth[s]);
  const int64_t unit = type == kU1 || type == kV1 ? st : (type == kU0 || type == kX) ? st / 2 : st / 4;
  const uint32_t ct[3] = {T.x[t], T.y[t], T.z[t]}, cs[3] = {T.x[s], T.y[s], T.z[s]};
  for (int d = 0; d < 3; d++) {
    const int64_t diff = rel[d] * unit + (4 * (int64_t)ct[d] + st / 2) - (4 * (int64_t)cs[d] + ss / 2);
    shift[d] = (int)(diff / ((int64_t)4 << 30));
    if (diff != (int64_t)shift[d] * ((int64_t)4 << 30)) throw Error(kErrInternal, "list entry is not a periodic image");
  }
}

void build_octree(Octree& T, const double* src, int64_t n_src, const double* trg, int64_t n_trg,
                  const double* pt, int64_t n_pt, int max_pts, int bndry) {
  if (max_pts <= 0) throw Error(kErrInvalidArgument, "max_pts must be positive");
  if (bndry < kFreeSpace || bndry > kPXYZ) throw Error(kErrInvalidArgument, "boundary type must be 0 (free), 1 (PX), 2 (PXY) or 3 (PXYZ)");
  T = Octree();
  T.max_pts = max_pts;
  T.bndry = bndry;
  sort_points(T.src, src, n_src);
  sort_points(T.trg, trg, n_trg);
  PointSet P;
  sort_points(P, pt, n_pt);
  KeyedSet Ks = keyed(T.src), Kt = keyed(T.trg), Kp = keyed(P);

  // 1. adaptive refinement by counts, level by level
  struct Cell { uint32_t x, y, z; int d; };
  std::vector<std::unordered_set<u128, U128Hash>> internal(kMaxDepth + 1);
  std::vector<Cell> level{{0, 0, 0, 0}};
  std::vector<std::array<int64_t, 6>> rng{{0, n_src, 0, n_trg, 0, n_pt}};
  for (int d = 0; d < kMaxDepth - 1 && !level.empty(); d++) {
    std::vector<Cell> nxt;
    std::vector<std::array<int64_t, 6>> nrng;
    const int sh = 3 * (29 - d);  // key bits below one child of a depth-d node
    for (size_t i = 0; i < level.size(); i++) {
      const auto& r = rng[i];
      int64_t c_st = (r[1] - r[0]) + (r[3] - r[2]), c_pt = r[5] - r[4];
      if (!(c_st > max_pts || c_pt > max_pts)) continue;
      const Cell& c = level[i];
      internal[d].insert(node_id(d, c.x, c.y, c.z));
      u128 base = morton(c.x, c.y, c.z);
      const uint32_t h = 1u << (29 - d);
      int64_t bs[9], bt[9], bp[9];
      bs[0] = r[0]; bt[0] = r[2]; bp[0] = r[4];
      bs[8] = r[1]; bt[8] = r[3]; bp[8] = r[5];
      for (int k = 1; k < 8; k++) {
        u128 kk = base + ((u128)k << sh);
        bs[k] = Ks.lower(kk, r[0], r[1]);
        bt[k] = Kt.lower(kk, r[2], r[3]);
        bp[k] = Kp.lower(kk, r[4], r[5]);
      }
      for (int k = 0; k < 8; k++) {
        nxt.push_back({c.x + ((k & 1) ? h : 0), c.y + ((k & 2) ? h : 0), c.z + ((k & 4) ? h : 0), d + 1});
        nrng.push_back({bs[k], bs[k + 1], bt[k], bt[k + 1], bp[k], bp[k + 1]});
      }
    }
    level.swap(nxt);
    rng.swap(nrng);
  }

  // 2. 2:1 balance closure: parents of leaves, then parents of all same-level neighbours
  std::vector<std::unordered_set<u128, U128Hash>> S(kMaxDepth + 1);
  for (int d = 0; d < kMaxDepth; d++)
    for (const u128& k : internal[d]) {
      // an internal node with a leaf child is a parent of a leaf
      uint32_t x = (uint32_t)(k >> 64), y = (uint32_t)(k >> 32), z = (uint32_t)k;
      const uint32_t h = 1u << (29 - d);
      bool has_leaf = false;
      for (int c = 0; c < 8 && !has_leaf; c++)
        if (!internal[d + 1].count(node_id(d + 1, x + ((c & 1) ? h : 0), y + ((c & 2) ? h : 0), z + ((c & 4) ? h : 0)))) has_leaf = true;
      if (has_leaf) S[d].insert(k);
    }
  for (int l = kMaxDepth; l >= 1; l--) {
    const int64_t step = (int64_t)1 << (30 - l), maxc = (int64_t)1 << 30;
    const uint32_t pmask = ~(uint32_t)((((uint64_t)1) << (30 - (l - 1))) - 1);
    std::vector<u128> add;
    for (const u128& k : S[l]) {
      int64_t x = (uint32_t)(k >> 64), y = (uint32_t)(k >> 32), z = (uint32_t)k;
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int64_t nx = x + dx * step, ny = y + dy * step, nz = z + dz * step;
            if (bndry != kFreeSpace) {  // the periodic balance wraps every axis
              nx = (nx + maxc) % maxc; ny = (ny + maxc) % maxc; nz = (nz + maxc) % maxc;
            }
            if (nx < 0 || ny < 0 || nz < 0 || nx >= maxc || ny >= maxc || nz >= maxc) continue;
            add.push_back(node_id(l - 1, (uint32_t)nx & pmask, (uint32_t)ny & pmask, (uint32_t)nz & pmask));
          }
    }
    S[l - 1].insert(add.begin(), add.end());
  }

  // 3. level-major node arrays (children of internal nodes created in Morton order)
  T.level_begin.assign(1, 0);
  struct NodeTmp { uint32_t x, y, z; int32_t parent, pos; };
  std::vector<NodeTmp> cur{{0, 0, 0, -1, 0}};
  int d = 0;
  std::vector<int64_t> prng{0, n_src, 0, n_trg};
  while (!cur.empty()) {
    std::vector<NodeTmp> nxt;
    int32_t base = (int32_t)T.depth.size();
    const uint32_t h = d < 30 ? (1u << (29 - d)) : 0;
    for (size_t i = 0; i < cur.size(); i++) {
      const NodeTmp& c = cur[i];
      int32_t id = base + (int32_t)i;
      T.x.push_back(c.x); T.y.push_back(c.y); T.z.push_back(c.z);
      T.depth.push_back((uint8_t)d);
      T.parent.push_back(c.parent);
      T.child_pos.push_back(c.pos);
      bool inner = d < kMaxDepth - 1 && S[d].count(node_id(d, c.x, c.y, c.z));
      if (inner) {
        for (int k = 0; k < 8; k++) nxt.push_back({c.x + ((k & 1) ? h : 0), c.y + ((k & 2) ? h : 0), c.z + ((k & 4) ? h : 0), id, k});
        T.first_child.push_back(0);  // patched below
      } else {
        T.first_child.push_back(-1);
      }
    }
    // first_child of this level's internal nodes
    int32_t next_base = base + (int32_t)cur.size();
    int32_t cnt = 0;
    for (size_t i = 0; i < cur.size(); i++)
      if (T.first_child[base + i] == 0) { T.first_child[base + i] = next_base + 8 * cnt; cnt++; }
    T.level_begin.push_back(next_base);
    cur.swap(nxt);
    d++;
  }
  const int32_t N = T.n_nodes();
  T.max_depth = (int)T.level_begin.size() - 2;

  // 4. subtree point ranges
  T.src_begin.resize(N); T.src_end.resize(N); T.trg_begin.resize(N); T.trg_end.resize(N);
  T.src_begin[0] = 0; T.src_end[0] = n_src; T.trg_begin[0] = 0; T.trg_end[0] = n_trg;
  for (int32_t i = 0; i < N; i++) {
    if (T.is_leaf(i)) continue;
    const int dd = T.depth[i];
    const int sh = 3 * (29 - dd);
    u128 base = morton(T.x[i], T.y[i], T.z[i]);
    int64_t bs[9], bt[9];
    bs[0] = T.src_begin[i]; bs[8] = T.src_end[i]; bt[0] = T.trg_begin[i]; bt[8] = T.trg_end[i];
    for (int k = 1; k < 8; k++) {
      u128 kk = base + ((u128)k << sh);
      bs[k] = Ks.lower(kk, bs[0], bs[8]);
      bt[k] = Kt.lower(kk, bt[0], bt[8]);
    }
    for (int k = 0; k < 8; k++) {
      int32_t c = T.first_child[i] + k;
      T.src_begin[c] = bs[k]; T.src_end[c] = bs[k + 1]; T.trg_begin[c] = bt[k]; T.trg_end[c] = bt[k + 1];
    }
  }

  // 5. colleagues through a hash of (depth, corner)
  std::unordered_map<u128, int32_t, U128Hash> where;
  where.reserve(N * 2);
  for (int32_t i = 0; i < N; i++) where[node_id(T.depth[i], T.x[i], T.y[i], T.z[i])] = i;
  T.colleague.assign((size_t)N * 27, -1);
  #pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < N; i++) {
    const int dd = T.depth[i];
    const int64_t step = (int64_t)1 << (30 - dd), maxc = (int64_t)1 << 30;
    for (int c = 0; c < 27; c++) {
      int dx = c % 3 - 1, dy = (c / 3) % 3 - 1, dz = c / 9 - 1;
      int64_t nx = (int64_t)T.x[i] + dx * step, ny = (int64_t)T.y[i] + dy * step, nz = (int64_t)T.z[i] + dz * step;
      if (bndry >= kPX) nx = (nx + maxc) % maxc;
      if (bndry >= kPXY) ny = (ny + maxc) % maxc;
      if (bndry >= kPXYZ) nz = (nz + maxc) % maxc;
      if (nx < 0 || ny < 0 || nz < 0 || nx >= maxc || ny >= maxc || nz >= maxc) continue;
      auto it = where.find(node_id(dd, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz));
      if (it != where.end()) T.colleague[(size_t)i * 27 + c] = it->second;
    }
  }

  // 6. interaction lists for nodes holding sources (S2U, U2U) or targets (the rest),
  //    include/fmm_tree.txx:300-352 and include/interac_list.txx:98-279
  const int types[] = {kS2U, kU2U, kD2D, kD2T, kU0, kU1, kU2, kW, kX, kV1};
  for (int t : types) T.lists[t].assign((size_t)N * T.list_count(t), -1);
  #pragma omp parallel for schedule(dynamic, 64)
  for (int32_t n = 0; n < N; n++) {
    const bool has_src = T.pt_cnt0(n) > 0, has_trg = T.pt_cnt1(n) > 0;
    const bool leaf = T.is_leaf(n);
    const int32_t p = T.parent[n];
    const int pos = T.child_pos[n];
    const int bx = pos & 1, by = (pos >> 1) & 1, bz = (pos >> 2) & 1;
    auto slot = [&](int t, int rx, int ry, int rz, int32_t v) {
      int idx = rel_table(t).index(rx, ry, rz);
      if (idx >= 0) T.lists[t][(size_t)n * T.list_count(t) + idx] = v;
    };
    if (has_src) {
      if (leaf) T.lists[kS2U][(size_t)n] = n;
      else for (int j = 0; j < 8; j++) slot(kU2U, (j & 1) ? 1 : -1, (j & 2) ? 1 : -1, (j & 4) ? 1 : -1, T.first_child[n] + j);
    }
    if (!has_trg) continue;
    if (leaf) T.lists[kD2T][(size_t)n] = n;
    if (p >= 0) slot(kD2D, bx ? 1 : -1, by ? 1 : -1, bz ? 1 : -1, p);
    // coarser leaves among the parent's colleagues: U0 (adjacent, leaf targets) and X (not adjacent)
    if (p >= 0)
      for (int i = 0; i < 27; i++) {
        int32_t pc = T.colleague[(size_t)p * 27 + i];
        if (pc < 0 || !T.is_leaf(pc)) continue;
        int rx = (i % 3) * 4 - 4 - 2 * bx + 1, ry = ((i / 3) % 3) * 4 - 4 - 2 * by + 1, rz = (i / 9) * 4 - 4 - 2 * bz + 1;
        if (leaf) slot(kU0, rx, ry, rz, pc);
        slot(kX, rx, ry, rz, pc);
      }
    for (int i = 0; i < 27; i++) {
      int32_t c = T.colleague[(size_t)n * 27 + i];
      if (c < 0) continue;
      int dx = i % 3 - 1, dy = (i / 3) % 3 - 1, dz = i / 9 - 1;
      if (leaf) {
        if (T.is_leaf(c)) slot(kU1, dx, dy, dz, c);
        else
          for (int j = 0; j < 8; j++) {
            int rx = dx * 4 + ((j & 1) ? 2 : 0) - 1, ry = dy * 4 + ((j & 2) ? 2 : 0) - 1, rz = dz * 4 + ((j & 4) ? 2 : 0) - 1;
            slot(kU2, rx, ry, rz, T.first_child[c] + j);
            slot(kW, rx, ry, rz, T.first_child[c] + j);
          }
      } else if (!T.is_leaf(c)) {
        slot(kV1, dx, dy, dz, c);
      }
    }
  }
}

}  // namespace pvfmm_amd
