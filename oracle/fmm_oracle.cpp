// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the reference particle FMM (dmalhotra/pvfmm) used as the
// parity checker for the HIP evaluator and as the CPU baseline in bench.py.  Nothing in
// the product (pvfmm_amd/, include/) links or calls this file.
//
// What is restated here, and from where (paths relative to the reference checkout):
//   * MortonId arithmetic ............ include/mortonid.txx:14-151, src/mortonid.cpp:15-76
//   * p2oLocal ....................... include/mpi_tree.txx:37-99
//   * pointer tree from linear tree .. include/mpi_tree.txx:266-306
//   * Subdivide / Truncate ........... include/mpi_node.txx:107-235, src/tree_node.cpp:83-105
//   * SubdivCond ..................... include/fmm_node.txx:62-95, include/mpi_node.txx:79-105
//   * CoarsenTree / RefineTree ....... include/mpi_tree.txx:314-492 (np=1)
//   * balanceOctree / Balance21 ...... include/mpi_tree.txx:745-1066 (np=1)
//   * SetColleagues .................. include/mpi_tree.txx:1136-1255
//   * InitList / BuildList ........... include/interac_list.txx:21-433
//   * pt_cnt aggregation ............. include/fmm_pts.txx:1306-1420
//   * list setups (node filters) ..... include/fmm_pts.txx:2275-2637,3346-3797,4444-5737
//   * EvalListPts semantics .......... include/fmm_pts.txx:3990-4441
//   * FFT_UpEquiv / Hadamard / IFFT .. include/fmm_pts.txx:2689-2922,3249-3344,3597-3759
//   * kernels ........................ include/kernel.txx:1137-1364
//   * periodic trees and images ...... src/mortonid.cpp:15-45 (NbrList), include/mpi_tree.txx:
//                                      1136-1170 (root colleagues), include/fmm_pts.txx:4646-4656,
//                                      4934-4940,5228-5340 (coord_shift)
//   * PeriodicBC / PostProcessing .... include/fmm_pts.txx:2641-2657,5740-5774 (the BC_Type
//                                      operator itself is precomputed in __init__.py)
// The reference's sctl::approx_rsqrt is replaced by an exact 1/sqrt (SCTL is not
// vendored); OpenMP partitioning and summation order differ, so agreement with the
// reference is at rounding level, not bitwise, for floating point results.  Tree and
// interaction lists are integer data and are restated exactly.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <set>
#include <vector>
#include <omp.h>

namespace orc {

constexpr int MAX_DEPTH = 30;     // include/pvfmm_common.hpp:29
constexpr double RAD0 = 1.05;     // include/pvfmm_common.hpp:33
constexpr double RAD1 = 2.95;     // include/pvfmm_common.hpp:34

// ---------------------------------------------------------------- MortonId
struct MortonId {
  uint32_t x = 0, y = 0, z = 0;
  uint8_t depth = 0;

  static uint32_t mask_for(int d) { return ~((((uint32_t)1) << (MAX_DEPTH - d)) - 1); }

  // include/mortonid.txx:41-55
  static MortonId from_coord(const double* c, int d) {
    const uint32_t max_int = ((uint32_t)1) << MAX_DEPTH;
    MortonId m;
    m.depth = (uint8_t)d;
    m.x = (uint32_t)std::floor(c[0] * max_int);
    m.y = (uint32_t)std::floor(c[1] * max_int);
    m.z = (uint32_t)std::floor(c[2] * max_int);
    uint32_t mk = mask_for(d);
    m.x &= mk; m.y &= mk; m.z &= mk;
    return m;
  }
  // include/mortonid.txx:78-99
  MortonId NextId() const {
    MortonId m = *this;
    uint32_t mask = ((uint32_t)1) << (MAX_DEPTH - depth);
    int i;
    for (i = depth; i >= 0; i--) {
      m.x = (m.x ^ mask); if (m.x & mask) break;
      m.y = (m.y ^ mask); if (m.y & mask) break;
      m.z = (m.z ^ mask); if (m.z & mask) break;
      mask = (mask << 1);
    }
    if (i < 0) i = 0;
    m.depth = (uint8_t)i;
    return m;
  }
  // include/mortonid.txx:101-110
  MortonId getAncestor(int lvl) const {
    MortonId m = *this;
    m.depth = (uint8_t)lvl;
    uint32_t mask = (((uint32_t)1) << MAX_DEPTH) - (((uint32_t)1) << (MAX_DEPTH - lvl));
    m.x &= mask; m.y &= mask; m.z &= mask;
    return m;
  }
  MortonId getDFD(int lvl = MAX_DEPTH) const { MortonId m = *this; m.depth = (uint8_t)lvl; return m; }
  // include/mortonid.txx:118-130
  bool operator<(const MortonId& m) const {
    if (x == m.x && y == m.y && z == m.z) return depth < m.depth;
    uint32_t x_ = x ^ m.x, y_ = y ^ m.y, z_ = z ^ m.z;
    if ((z_ > x_ || ((z_ ^ x_) < x_ && (z_ ^ x_) < z_)) && (z_ > y_ || ((z_ ^ y_) < y_ && (z_ ^ y_) < z_)))
      return z < m.z;
    if (y_ > x_ || ((y_ ^ x_) < x_ && (y_ ^ x_) < y_)) return y < m.y;
    return x < m.x;
  }
  bool operator==(const MortonId& m) const { return x == m.x && y == m.y && z == m.z && depth == m.depth; }
  bool operator!=(const MortonId& m) const { return !(*this == m); }
  bool operator>(const MortonId& m) const { return m < *this; }
  bool isAncestor(const MortonId& o) const { return o.depth > depth && o.getAncestor(depth) == *this; }
  // src/mortonid.cpp:15-45; periodic wraps every axis (balanceOctree passes bndry != FreeSpace)
  void NbrList(std::vector<MortonId>& nbrs, int level, bool periodic = false) const {
    nbrs.clear();
    const int64_t maxCoord = ((int64_t)1) << MAX_DEPTH;
    uint32_t mask = (uint32_t)(maxCoord - (((int64_t)1) << (MAX_DEPTH - level)));
    int64_t pX = x & mask, pY = y & mask, pZ = z & mask;
    int64_t step = ((int64_t)1) << (MAX_DEPTH - level);
    for (int i = 0; i < 27; i++) {
      int64_t nx = pX + ((i / 1) % 3 - 1) * step;
      int64_t ny = pY + ((i / 3) % 3 - 1) * step;
      int64_t nz = pZ + ((i / 9) % 3 - 1) * step;
      if (periodic) {
        if (nx < 0) nx += maxCoord;
        if (nx >= maxCoord) nx -= maxCoord;
        if (ny < 0) ny += maxCoord;
        if (ny >= maxCoord) ny -= maxCoord;
        if (nz < 0) nz += maxCoord;
        if (nz >= maxCoord) nz -= maxCoord;
      }
      if (nx >= 0 && nx < maxCoord && ny >= 0 && ny < maxCoord && nz >= 0 && nz < maxCoord) {
        MortonId t; t.x = (uint32_t)nx; t.y = (uint32_t)ny; t.z = (uint32_t)nz; t.depth = (uint8_t)level;
        nbrs.push_back(t);
      }
    }
  }
  // src/mortonid.cpp:47-66
  void Children(MortonId* out) const {
    uint32_t mask = (uint32_t)((((uint64_t)1) << MAX_DEPTH) - (((uint64_t)1) << (MAX_DEPTH - depth)));
    uint32_t pX = x & mask, pY = y & mask, pZ = z & mask;
    uint32_t step = ((uint32_t)1) << (MAX_DEPTH - (depth + 1));
    for (int i = 0; i < 8; i++) {
      out[i].x = pX + step * ((i / 1) % 2);
      out[i].y = pY + step * ((i / 2) % 2);
      out[i].z = pZ + step * ((i / 4) % 2);
      out[i].depth = (uint8_t)(depth + 1);
    }
  }
};

// ---------------------------------------------------------------- interaction lists
// Mat_Type order follows include/precomp_mat.hpp:21-39
enum MatType { UC2UE0_T = 0, UC2UE1_T, DC2DE0_T, DC2DE1_T, S2U_T, U2U_T, D2D_T, D2T_T,
               U0_T, U1_T, U2_T, V_T, W_T, X_T, V1_T, BC_T, TYPE_COUNT };

struct InteracList {
  std::vector<std::vector<int>> rel;  // rel[t][3*i+d]
  std::vector<std::vector<int>> hash_lut;
  static int coord_hash(const int* c) { const int n = 5; return ((c[2] + n) * (2 * n) + (c[1] + n)) * (2 * n) + (c[0] + n); }
  static int class_hash(const int* c_) {  // include/interac_list.txx:350-362
    int c[3] = {std::abs(c_[0]), std::abs(c_[1]), std::abs(c_[2])};
    if (c[1] > c[0] && c[1] > c[2]) std::swap(c[0], c[1]);
    if (c[0] > c[2]) std::swap(c[0], c[2]);
    if (c[0] > c[1]) std::swap(c[0], c[1]);
    return coord_hash(c);
  }
  // include/interac_list.txx:368-404
  void InitList(int max_r, int min_r, int step, int t) {
    const int HASH = 2000;
    hash_lut[t].assign(HASH, -1);
    std::vector<int> size(HASH, 0), disp(HASH, 0);
    int count = 0;
    for (int k = -max_r; k <= max_r; k += step)
      for (int j = -max_r; j <= max_r; j += step)
        for (int i = -max_r; i <= max_r; i += step)
          if (std::abs(i) >= min_r || std::abs(j) >= min_r || std::abs(k) >= min_r) {
            int c[3] = {i, j, k};
            size[class_hash(c)]++;
            count++;
          }
    for (int i = 1; i < HASH; i++) disp[i] = disp[i - 1] + size[i - 1];
    rel[t].assign(3 * count, 0);
    for (int k = -max_r; k <= max_r; k += step)
      for (int j = -max_r; j <= max_r; j += step)
        for (int i = -max_r; i <= max_r; i += step)
          if (std::abs(i) >= min_r || std::abs(j) >= min_r || std::abs(k) >= min_r) {
            int c[3] = {i, j, k};
            int& idx = disp[class_hash(c)];
            rel[t][3 * idx + 0] = i; rel[t][3 * idx + 1] = j; rel[t][3 * idx + 2] = k;
            hash_lut[t][coord_hash(c)] = idx;
            idx++;
          }
  }
  // include/interac_list.txx:21-57
  void Initialize() {
    rel.assign(TYPE_COUNT, {});
    hash_lut.assign(TYPE_COUNT, {});
    InitList(0, 0, 1, UC2UE0_T); InitList(0, 0, 1, UC2UE1_T);
    InitList(0, 0, 1, DC2DE0_T); InitList(0, 0, 1, DC2DE1_T);
    InitList(0, 0, 1, S2U_T); InitList(1, 1, 2, U2U_T); InitList(1, 1, 2, D2D_T); InitList(0, 0, 1, D2T_T);
    InitList(3, 3, 2, U0_T); InitList(1, 0, 1, U1_T); InitList(3, 3, 2, U2_T);
    InitList(3, 2, 1, V_T); InitList(1, 1, 1, V1_T); InitList(5, 5, 2, W_T); InitList(5, 5, 2, X_T);
    InitList(0, 0, 1, BC_T);
  }
  int count(int t) const { return (int)rel[t].size() / 3; }
};

// ---------------------------------------------------------------- tree
struct Node {
  Node* parent = nullptr;
  Node* child[8] = {};
  bool leaf = true, ghost = false;
  int depth = 0, p2n = 0;
  double coord[3] = {0, 0, 0};
  MortonId mid;
  std::vector<int64_t> pts[3];  // 0: src, 1: trg, 2: pt_coord (tree-defining set); indices into input arrays
  Node* colleague[27] = {};
  int64_t pt_cnt[2] = {0, 0};
  std::vector<Node*> lst[TYPE_COUNT];
  int id = -1;
};

struct Tree {
  int max_pts = 0;
  int bndry = 0;  // BoundaryType (include/mpi_tree.hpp:22-29): 0 FreeSpace, 1 PX, 2 PXY, 3 PXYZ
  const double* X[3] = {nullptr, nullptr, nullptr};  // coordinate arrays (AoS) of the three point sets
  int64_t N[3] = {0, 0, 0};
  Node* root = nullptr;
  std::vector<Node*> nodes;  // preorder
  InteracList il;

  ~Tree() { destroy(root); }
  static void destroy(Node* n) {
    if (!n) return;
    for (int i = 0; i < 8; i++) destroy(n->child[i]);
    delete n;
  }

  // ---- traversal (include/tree.txx)
  static Node* PreorderNxt(Node* n) {
    if (!n->leaf) return n->child[0];
    while (n->parent) {
      if (n->p2n < 7) return n->parent->child[n->p2n + 1];
      n = n->parent;
    }
    return nullptr;
  }
  Node* PostorderFirst() { Node* n = root; while (!n->leaf) n = n->child[0]; return n; }
  static Node* PostorderNxt(Node* n) {
    if (!n->parent) return nullptr;
    if (n->p2n < 7) {
      Node* s = n->parent->child[n->p2n + 1];
      while (!s->leaf) s = s->child[0];
      return s;
    }
    return n->parent;
  }

  // ---- SubdivCond: include/fmm_node.txx:62-95 calling include/mpi_node.txx:79-105
  bool SubdivCond(Node* n) {
    if (n->depth >= MAX_DEPTH - 1) return false;
    if (!n->leaf) {
      for (int i = 0; i < 8; i++) if (!n->child[i]->leaf || n->child[i]->ghost) return true;
    } else if (n->ghost) return false;
    // MPI_Node part: pt_coord count
    if (!n->leaf) {
      size_t c = 0;
      for (int i = 0; i < 8; i++) c += n->child[i]->pts[2].size();
      if ((int64_t)c > max_pts) return true;
    } else if ((int64_t)n->pts[2].size() > max_pts) return true;
    // FMM_Node part: src + surf + trg count (no surf points here)
    size_t c = 0;
    if (!n->leaf) {
      for (int i = 0; i < 8; i++) c += n->child[i]->pts[0].size() + n->child[i]->pts[1].size();
    } else c = n->pts[0].size() + n->pts[1].size();
    return (int64_t)c > max_pts;
  }

  // ---- MPI_Node::Subdivide: include/mpi_node.txx:107-189, child coords as in TreeNode::Initialize
  void Subdivide(Node* n) {
    if (!n->leaf) return;
    n->leaf = false;
    double s = std::pow(0.5, n->depth + 1);
    for (int i = 0; i < 8; i++) {
      Node* c = new Node();
      c->parent = n; c->p2n = i; c->depth = n->depth + 1;
      c->coord[0] = n->coord[0] + ((i & 1) ? s : 0);
      c->coord[1] = n->coord[1] + ((i & 2) ? s : 0);
      c->coord[2] = n->coord[2] + ((i & 4) ? s : 0);
      c->ghost = n->ghost;
      MortonId cm[8]; n->mid.Children(cm); c->mid = cm[i];
      n->child[i] = c;
    }
    if (n->ghost) { for (auto& v : n->pts) v.clear(); return; }
    const double* c = n->coord;
    for (int j = 0; j < 3; j++) {
      std::vector<int64_t>& v = n->pts[j];
      int64_t npts = (int64_t)v.size();
      if (!npts) continue;
      const double* X_ = X[j];
      int64_t cdata[9];
      for (int i = 0; i < 9; i++) {
        int64_t pt1 = -1, pt2 = npts;
        while (pt2 - pt1 > 1) {
          int64_t pt3 = (pt1 + pt2) / 2;
          const double* p = &X_[3 * v[pt3]];
          int ch_id = (p[0] >= c[0] + s) * 1 + (p[1] >= c[1] + s) * 2 + (p[2] >= c[2] + s) * 4;
          if (ch_id < i) pt1 = pt3;
          if (ch_id >= i) pt2 = pt3;
        }
        cdata[i] = pt2;
      }
      for (int i = 0; i < 8; i++) n->child[i]->pts[j].assign(v.begin() + cdata[i], v.begin() + cdata[i + 1]);
      v.clear();
      v.shrink_to_fit();
    }
  }
  // ---- MPI_Node::Truncate: include/mpi_node.txx:191-235
  void Truncate(Node* n) {
    if (n->leaf) return;
    for (int i = 0; i < 8; i++) if (!n->child[i]->leaf) Truncate(n->child[i]);
    for (int j = 0; j < 3; j++) {
      std::vector<int64_t>& v = n->pts[j];
      v.clear();
      for (int i = 0; i < 8; i++) v.insert(v.end(), n->child[i]->pts[j].begin(), n->child[i]->pts[j].end());
    }
    for (int i = 0; i < 8; i++) { destroy(n->child[i]); n->child[i] = nullptr; }
    n->leaf = true;
  }

  // ---- MPI_Tree::FindNode: include/mpi_tree.txx:656-672
  Node* FindNode(const MortonId& key, bool subdiv, Node* start = nullptr) {
    Node* n = start ? start : root;
    while (n->mid < key && (!n->leaf || subdiv)) {
      if (n->leaf && !n->ghost) Subdivide(n);
      if (n->leaf) break;
      for (int j = 0; j < 8; j++) {
        if (n->child[j]->mid.NextId() > key) { n = n->child[j]; break; }
      }
    }
    return n;
  }

  // ---- p2oLocal: include/mpi_tree.txx:37-99
  static void p2oLocal(const std::vector<MortonId>& nodes, std::vector<MortonId>& leaves,
                       unsigned maxNumPts, unsigned maxDepth, bool complete) {
    std::vector<MortonId> leaves_lst;
    unsigned init_size = (unsigned)leaves.size();
    unsigned num_pts = (unsigned)nodes.size();
    MortonId curr_node = leaves[0];
    MortonId last_node = leaves[init_size - 1].NextId();
    MortonId next_node;
    unsigned curr_pt = 0, next_pt = curr_pt + maxNumPts;
    while (next_pt <= num_pts) {
      next_node = curr_node.NextId();
      while (next_pt < num_pts && next_node > nodes[next_pt] && curr_node.depth < maxDepth - 1) {
        curr_node = curr_node.getDFD(curr_node.depth + 1);
        next_node = curr_node.NextId();
      }
      leaves_lst.push_back(curr_node);
      curr_node = next_node;
      unsigned inc = maxNumPts;
      while (next_pt < num_pts && curr_node > nodes[next_pt]) {
        inc = inc << 1;
        next_pt += inc;
        if (next_pt > num_pts) { next_pt = num_pts; break; }
      }
      curr_pt = (unsigned)(std::lower_bound(nodes.begin() + curr_pt, nodes.begin() + next_pt, curr_node) - nodes.begin());
      if (curr_pt >= num_pts) break;
      next_pt = curr_pt + maxNumPts;
      if (next_pt > num_pts) next_pt = num_pts;
    }
    if (complete)
      while (curr_node < last_node) {
        while (curr_node.NextId() > last_node && curr_node.depth < maxDepth - 1)
          curr_node = curr_node.getDFD(curr_node.depth + 1);
        leaves_lst.push_back(curr_node);
        curr_node = curr_node.NextId();
      }
    leaves = leaves_lst;
  }

  // ---- MPI_Tree::Initialize at np=1: include/mpi_tree.txx:203-311
  void Initialize() {
    root = new Node();
    std::vector<MortonId> pt_mid(N[2]);
    for (int64_t i = 0; i < N[2]; i++) pt_mid[i] = MortonId::from_coord(&X[2][3 * i], MAX_DEPTH);
    std::sort(pt_mid.begin(), pt_mid.end());
    unsigned maxNumPts = (unsigned)max_pts;
    if ((int64_t)N[2] < (int64_t)maxNumPts) maxNumPts = (unsigned)N[2];  // points2Octree, mpi_tree.txx:121-124
    std::vector<MortonId> lin_oct(1);
    p2oLocal(pt_mid, lin_oct, maxNumPts, MAX_DEPTH, true);
    // SortScatterIndex + ScatterForward: every point set ordered by its depth-30 MortonId
    for (int j = 0; j < 3; j++) {
      std::vector<std::pair<MortonId, int64_t>> key(N[j]);
      for (int64_t i = 0; i < N[j]; i++) key[i] = {MortonId::from_coord(&X[j][3 * i], MAX_DEPTH), i};
      std::stable_sort(key.begin(), key.end(), [](const std::pair<MortonId, int64_t>& a, const std::pair<MortonId, int64_t>& b) { return a.first < b.first; });
      root->pts[j].resize(N[j]);
      for (int64_t i = 0; i < N[j]; i++) root->pts[j][i] = key[i].second;
    }
    // pointer tree from lin_oct (single thread)
    FindNode(lin_oct[0], true);
    size_t idx = 0, b = lin_oct.size();
    Node* n = root;
    while (n) {
      n->ghost = false;
      MortonId dn = n->mid;
      if (idx < b && dn.isAncestor(lin_oct[idx])) {
        if (n->leaf) Subdivide(n);
      } else if (idx < b && dn == lin_oct[idx]) {
        if (!n->leaf) Truncate(n);
        idx++;
      } else {
        Truncate(n);
        n->ghost = true;
      }
      n = PreorderNxt(n);
    }
  }

  // ---- CoarsenTree at np=1, omp_p=1: include/mpi_tree.txx:314-400
  void CoarsenTree() {
    std::vector<Node*> leaf_nodes;
    Node* n = PostorderFirst();
    while (n) { if (n->leaf && !n->ghost) break; Truncate(n); n->ghost = true; n = PostorderNxt(n); }
    while (n) { if (n->leaf && n->ghost) break; if (n->leaf) leaf_nodes.push_back(n); n = PreorderNxt(n); }
    while (n) { Truncate(n); n->ghost = true; n = PreorderNxt(n); }
    if (leaf_nodes.empty()) return;
    MortonId mid0 = leaf_nodes[0]->mid;
    Node* n_ = leaf_nodes[0];
    while (n_) {
      MortonId n_mid = n_->mid;
      if (!n_->leaf && !n_mid.isAncestor(mid0.getDFD()))
        if (!SubdivCond(n_)) Truncate(n_);
      n_ = PostorderNxt(n_);
    }
    n_ = FindNode(mid0, false, root);
    while (n_->depth > 0) {
      n_ = n_->parent;
      if (!SubdivCond(n_)) Truncate(n_);
      else break;
    }
  }

  // ---- RefineTree at np=1: include/mpi_tree.txx:403-492
  void RefineTree() {
    CoarsenTree();
    std::vector<Node*> leaf_nodes;
    for (Node* n = PostorderFirst(); n; n = PostorderNxt(n))
      if (n->leaf && !n->ghost) leaf_nodes.push_back(n);
    for (int l = 0; l < MAX_DEPTH; l++) {
      std::vector<Node*> nxt;
      for (Node* ln : leaf_nodes) {
        if (ln->leaf && !ln->ghost) {
          if (SubdivCond(ln)) Subdivide(ln);
          if (!ln->leaf) for (int k = 0; k < 8; k++) nxt.push_back(ln->child[k]);
        }
      }
      leaf_nodes.swap(nxt);
    }
    CoarsenTree();
    CoarsenTree();
  }

  // ---- balanceOctree at np=1, omp_p=1: include/mpi_tree.txx:745-943
  static void balanceOctree(std::vector<MortonId>& in, std::vector<MortonId>& out, unsigned maxDepth, bool periodic) {
    if (in.size() == 1) { out = in; return; }
    std::vector<std::set<MortonId>> nodes(maxDepth + 1);
    for (size_t i = 0; i < in.size();) {
      size_t d = in[i].depth;
      if (d == 0) { i++; continue; }
      MortonId pnode = in[i].getAncestor((int)d - 1);
      nodes[d - 1].insert(pnode);
      while (i < in.size() && d == in[i].depth && pnode == in[i].getAncestor((int)d - 1)) i++;
    }
    std::vector<MortonId> nbrs;
    for (unsigned l = maxDepth; l >= 1; l--) {
      std::set<MortonId> nbrs_parent;
      for (const MortonId& node : nodes[l]) {
        node.NbrList(nbrs, (int)l, periodic);
        for (const MortonId& nb : nbrs) nbrs_parent.insert(nb.getAncestor((int)l - 1));
      }
      nodes[l - 1].insert(nbrs_parent.begin(), nbrs_parent.end());
    }
    for (unsigned l = 1; l <= maxDepth; l++)
      for (const MortonId& node : nodes[l]) nodes[l - 1].erase(node.getAncestor(node.depth - 1));
    in.clear();
    for (unsigned l = 0; l <= maxDepth; l++) in.insert(in.end(), nodes[l].begin(), nodes[l].end());
    out = in;
    std::sort(out.begin(), out.end());
    {  // lineariseList, include/mpi_tree.txx:676-743
      std::vector<MortonId> tmp;
      if (!out.empty()) {
        for (size_t i = 0; i + 1 < out.size(); i++)
          if (!out[i].isAncestor(out[i + 1]) && out[i] != out[i + 1]) tmp.push_back(out[i]);
        tmp.push_back(out.back());
      }
      out.swap(tmp);
    }
    if (!out.empty()) {  // add children, include/mpi_tree.txx:856-927
      MortonId nxt_mid;  // (0,0,0,0)
      std::vector<MortonId> out1;
      MortonId ch[8];
      for (size_t i = 0; i < out.size(); i++) {
        while (nxt_mid.getDFD() < out[i]) {
          while (nxt_mid.isAncestor(out[i])) nxt_mid = nxt_mid.getAncestor(nxt_mid.depth + 1);
          out1.push_back(nxt_mid);
          nxt_mid = nxt_mid.NextId();
        }
        out[i].Children(ch);
        for (int j = 0; j < 8; j++) out1.push_back(ch[j]);
        nxt_mid = out[i].NextId();
      }
      while (nxt_mid.depth > 0) { out1.push_back(nxt_mid); nxt_mid = nxt_mid.NextId(); }
      out.swap(out1);
    }
  }

  // ---- Balance21 at np=1: include/mpi_tree.txx:945-1066
  void Balance21() {
    std::vector<MortonId> in, out;
    for (Node* n = root; n; n = PreorderNxt(n)) if (n->leaf && !n->ghost) in.push_back(n->mid);
    balanceOctree(in, out, MAX_DEPTH, bndry != 0);
    FindNode(out[0], true);
    FindNode(out[0], true);
    size_t node_iter = 0;
    Node* n = FindNode(out[node_iter], false);
    while (n && node_iter < out.size()) {
      n->ghost = false;
      MortonId dn = n->mid;
      if (dn.isAncestor(out[node_iter]) && dn != out[node_iter]) {
        if (n->leaf) Subdivide(n);
      } else if (dn == out[node_iter]) {
        node_iter++;
      } else {
        Truncate(n);
        n->ghost = true;
      }
      n = PreorderNxt(n);
    }
    while (n) { Truncate(n); n->ghost = true; n = PreorderNxt(n); }
  }

  // ---- SetColleagues: include/mpi_tree.txx:1136-1255.  The root is its own colleague at
  // offset 0 and, along each periodic axis (PX: x; PXY: x, y; PXYZ: all), at offsets -1 and +1.
  void SetColleagues() {
    for (int i = 0; i < 27; i++) root->colleague[i] = nullptr;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if ((dx && bndry < 1) || (dy && bndry < 2) || (dz && bndry < 3)) continue;
          root->colleague[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)] = root;
        }
    std::vector<std::vector<Node*>> by_depth(MAX_DEPTH + 1);
    for (Node* n = PreorderNxt(root); n; n = PreorderNxt(n)) by_depth[n->depth].push_back(n);
    for (int d = 0; d <= MAX_DEPTH; d++)
      for (Node* node : by_depth[d]) {
        for (int i = 0; i < 27; i++) node->colleague[i] = nullptr;
        Node* p = node->parent;
        int l = node->p2n;
        for (int i = 0; i < 27; i++) {
          Node* t1 = p->colleague[i];
          if (t1 && !t1->leaf)
            for (int j = 0; j < 8; j++) {
              Node* t2 = t1->child[j];
              bool flag = true;
              int a = 1, b = 1, new_indx = 0;
              for (int k = 0; k < 3; k++) {
                int diff = (((i / b) % 3) - 1) * 2 + ((j / a) % 2) - ((l / a) % 2);
                if (-1 > diff || diff > 1) flag = false;
                new_indx += (diff + 1) * b;
                a *= 2; b *= 3;
              }
              if (flag) node->colleague[new_indx] = t2;
            }
        }
      }
  }

  // ---- BuildList: include/interac_list.txx:98-279
  void BuildList(Node* n, int t) {
    std::vector<Node*>& L = n->lst[t];
    L.assign(il.count(t), nullptr);
    int rc[3];
    auto put = [&](Node* x) { int idx = il.hash_lut[t][InteracList::coord_hash(rc)]; if (idx >= 0) L[idx] = x; };
    switch (t) {
      case S2U_T: case D2T_T:
        if (!n->ghost && n->leaf) L[0] = n;
        break;
      case U2U_T:
        if (n->ghost || n->leaf) return;
        for (int j = 0; j < 8; j++) {
          rc[0] = -1 + ((j & 1) ? 2 : 0); rc[1] = -1 + ((j & 2) ? 2 : 0); rc[2] = -1 + ((j & 4) ? 2 : 0);
          int idx = il.hash_lut[t][InteracList::coord_hash(rc)];
          if (idx >= 0 && !n->child[j]->ghost) L[idx] = n->child[j];
        }
        break;
      case D2D_T: {
        if (n->ghost || !n->parent) return;
        int p2n = n->p2n;
        rc[0] = -1 + ((p2n & 1) ? 2 : 0); rc[1] = -1 + ((p2n & 2) ? 2 : 0); rc[2] = -1 + ((p2n & 4) ? 2 : 0);
        put(n->parent);
        break;
      }
      case U0_T: case X_T: {
        if (n->ghost || !n->parent) return;
        if (t == U0_T && !n->leaf) return;
        Node* p = n->parent;
        int p2n = n->p2n;
        for (int i = 0; i < 27; i++) {
          Node* pc = p->colleague[i];
          if (pc && pc->leaf) {
            rc[0] = (i % 3) * 4 - 4 - ((p2n & 1) ? 2 : 0) + 1;
            rc[1] = ((i / 3) % 3) * 4 - 4 - ((p2n & 2) ? 2 : 0) + 1;
            rc[2] = ((i / 9) % 3) * 4 - 4 - ((p2n & 4) ? 2 : 0) + 1;
            put(pc);
          }
        }
        break;
      }
      case U1_T:
        if (n->ghost || !n->leaf) return;
        for (int i = 0; i < 27; i++) {
          Node* c = n->colleague[i];
          if (c && c->leaf) { rc[0] = (i % 3) - 1; rc[1] = ((i / 3) % 3) - 1; rc[2] = ((i / 9) % 3) - 1; put(c); }
        }
        break;
      case U2_T: case W_T:
        if (n->ghost || !n->leaf) return;
        for (int i = 0; i < 27; i++) {
          Node* c = n->colleague[i];
          if (c && !c->leaf)
            for (int j = 0; j < 8; j++) {
              rc[0] = (i % 3) * 4 - 4 + ((j & 1) ? 2 : 0) - 1;
              rc[1] = ((i / 3) % 3) * 4 - 4 + ((j & 2) ? 2 : 0) - 1;
              rc[2] = ((i / 9) % 3) * 4 - 4 + ((j & 4) ? 2 : 0) - 1;
              put(c->child[j]);
            }
        }
        break;
      case V_T: {
        if (n->ghost || !n->parent) return;
        Node* p = n->parent;
        int p2n = n->p2n;
        for (int i = 0; i < 27; i++) {
          Node* pc = p->colleague[i];
          if (pc && !pc->leaf)
            for (int j = 0; j < 8; j++) {
              rc[0] = (i % 3) * 2 - 2 + ((j & 1) ? 1 : 0) - ((p2n & 1) ? 1 : 0);
              rc[1] = ((i / 3) % 3) * 2 - 2 + ((j & 2) ? 1 : 0) - ((p2n & 2) ? 1 : 0);
              rc[2] = ((i / 9) % 3) * 2 - 2 + ((j & 4) ? 1 : 0) - ((p2n & 4) ? 1 : 0);
              put(pc->child[j]);
            }
        }
        break;
      }
      case V1_T:
        if (n->ghost || n->leaf) return;
        for (int i = 0; i < 27; i++) {
          Node* c = n->colleague[i];
          if (c && !c->leaf) { rc[0] = (i % 3) - 1; rc[1] = ((i / 3) % 3) - 1; rc[2] = ((i / 9) % 3) - 1; put(c); }
        }
        break;
      default: break;
    }
  }

  // ---- pt_cnt (include/fmm_pts.txx:1306-1420) and list construction (include/fmm_tree.txx:300-352)
  void Finalize() {
    nodes.clear();
    for (Node* n = root; n; n = PreorderNxt(n)) { n->id = (int)nodes.size(); nodes.push_back(n); }
    for (Node* n : nodes) { n->pt_cnt[0] = n->pt_cnt[1] = 0; }
    for (int i = (int)nodes.size() - 1; i >= 0; i--) {
      Node* n = nodes[i];
      if (n->leaf) { n->pt_cnt[0] = (int64_t)n->pts[0].size(); n->pt_cnt[1] = (int64_t)n->pts[1].size(); }
      else for (int k = 0; k < 8; k++) { n->pt_cnt[0] += n->child[k]->pt_cnt[0]; n->pt_cnt[1] += n->child[k]->pt_cnt[1]; }
    }
    SetColleagues();
    const int src_types[] = {S2U_T, U2U_T};
    const int trg_types[] = {D2D_T, D2T_T, U0_T, U1_T, U2_T, W_T, X_T, V1_T};
    for (Node* n : nodes) {
      for (auto& L : n->lst) L.clear();
      if (!n->ghost && n->pt_cnt[0]) for (int t : src_types) BuildList(n, t);
      if (!n->ghost && n->pt_cnt[1]) for (int t : trg_types) BuildList(n, t);
      if (!n->ghost && n->pt_cnt[1] && n->parent) BuildList(n, V_T);  // V (child level) list, for inspection only
    }
  }
};

// ---------------------------------------------------------------- kernels (FP64, exact 1/sqrt)
// k = 0: laplace_poten (1x1, 1/(4 pi)), include/kernel.txx:1276-1287
// k = 1: laplace_grad  (1x3, 1/(4 pi)), include/kernel.txx:1316-1330, r = x_trg - x_src
// k = 2: stokes_vel    (3x3, 1/(8 pi)), include/kernel.txx:1351-1364
struct Ker { int id; int k0, k1; };
static inline Ker ker_of(int id) {
  if (id == 0) return {0, 1, 1};
  if (id == 1) return {1, 1, 3};
  return {2, 3, 3};
}
// Accumulate sources (coords already in a common frame) into ns x k1 targets.
static void p2p(const Ker& K, const double* sx, int64_t ns, const double* sv, const double* tx, int64_t nt, double* tv) {
  const double pi = 3.14159265358979323846;
  for (int64_t t = 0; t < nt; t++) {
    double acc[3] = {0, 0, 0};
    for (int64_t s = 0; s < ns; s++) {
      double r[3] = {tx[3 * t] - sx[3 * s], tx[3 * t + 1] - sx[3 * s + 1], tx[3 * t + 2] - sx[3 * s + 2]};
      double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (!(r2 > 0)) continue;
      double rinv = 1.0 / std::sqrt(r2);
      if (K.id == 0) acc[0] += sv[s] * rinv;
      else if (K.id == 1) { double f = rinv * rinv * rinv * sv[s]; acc[0] += r[0] * f; acc[1] += r[1] * f; acc[2] += r[2] * f; }
      else {
        const double* f = &sv[3 * s];
        double ip = (f[0] * r[0] + f[1] * r[1] + f[2] * r[2]) * rinv * rinv;
        for (int d = 0; d < 3; d++) acc[d] += rinv * (f[d] + r[d] * ip);
      }
    }
    double scale = (K.id == 2) ? 1.0 / (8 * pi) : 1.0 / (4 * pi);
    for (int d = 0; d < K.k1; d++) tv[K.k1 * t + d] += acc[d] * scale;
  }
}

// ---------------------------------------------------------------- surfaces, include/fmm_pts.txx:54-97
static std::vector<double> surface(int p, const double* c, double alpha, int depth) {
  size_t n_ = 6 * (p - 1) * (p - 1) + 2;
  std::vector<double> coord(n_ * 3);
  coord[0] = coord[1] = coord[2] = -1.0;
  size_t cnt = 1;
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) {
      coord[cnt * 3] = -1; coord[cnt * 3 + 1] = (2 * (i + 1) - p + 1) / (double)(p - 1); coord[cnt * 3 + 2] = (2 * j - p + 1) / (double)(p - 1); cnt++;
    }
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) {
      coord[cnt * 3] = (2 * i - p + 1) / (double)(p - 1); coord[cnt * 3 + 1] = -1; coord[cnt * 3 + 2] = (2 * (j + 1) - p + 1) / (double)(p - 1); cnt++;
    }
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) {
      coord[cnt * 3] = (2 * (i + 1) - p + 1) / (double)(p - 1); coord[cnt * 3 + 1] = (2 * j - p + 1) / (double)(p - 1); coord[cnt * 3 + 2] = -1; cnt++;
    }
  for (size_t i = 0; i < (n_ / 2) * 3; i++) coord[cnt * 3 + i] = -coord[i];
  double r = 0.5 * std::pow(0.5, depth);
  double b = alpha * r;
  for (size_t i = 0; i < n_; i++)
    for (int d = 0; d < 3; d++) coord[i * 3 + d] = (coord[i * 3 + d] + 1) * b + c[d];
  return coord;
}
// anchored at the box's lower corner c; radius RAD0 or RAD1 (include/fmm_pts.txx:99-141)
static std::vector<double> surf_at(int m, const double* c, int depth, double rad) {
  double r = 0.5 * std::pow(0.5, depth);
  double cc[3] = {c[0] - r * (rad - 1.0), c[1] - r * (rad - 1.0), c[2] - r * (rad - 1.0)};
  return surface(m, cc, rad, depth);
}

// ---------------------------------------------------------------- 3-D real DFT (separable, FFTW conventions)
// Forward r2c on an n^3 grid stored x-fastest; the x axis is halved (n/2+1), output layout [z][y][x'].
struct DFT3 {
  int n, nh;
  std::vector<double> cs, sn;  // cos/sin(2 pi k/n)
  explicit DFT3(int n_) : n(n_), nh(n_ / 2 + 1), cs(n_), sn(n_) {
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; k++) { cs[k] = std::cos(2 * pi * k / n); sn[k] = std::sin(2 * pi * k / n); }
  }
  // in: n^3 reals; out: n*n*nh complex (interleaved)
  void r2c(const double* in, double* out, std::vector<double>& w) const {
    size_t M = (size_t)n * n * nh;
    w.assign(2 * M, 0.0);
    // x-axis (real -> half complex)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++) {
        const double* row = &in[((size_t)z * n + y) * n];
        bool nz = false;
        for (int x = 0; x < n; x++) if (row[x] != 0) { nz = true; break; }
        if (!nz) continue;
        for (int k = 0; k < nh; k++) {
          double re = 0, im = 0;
          for (int x = 0; x < n; x++) {
            if (row[x] == 0) continue;
            int e = (k * x) % n;
            re += row[x] * cs[e]; im -= row[x] * sn[e];
          }
          size_t o = ((size_t)z * n + y) * nh + k;
          w[2 * o] = re; w[2 * o + 1] = im;
        }
      }
    axis(w, out, 1, -1);
    w.assign(out, out + 2 * M);
    axis(w, out, 2, -1);
  }
  // complex transform along axis a (1: y, 2: z) of an [z][y][x'] array; sign -1 forward, +1 backward
  void axis(const std::vector<double>& src, double* dst, int a, int sign) const {
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int k = 0; k < nh; k++) {
          double re = 0, im = 0;
          for (int j = 0; j < n; j++) {
            size_t o = (a == 1) ? (((size_t)z * n + j) * nh + k) : (((size_t)j * n + y) * nh + k);
            double xr = src[2 * o], xi = src[2 * o + 1];
            if (xr == 0 && xi == 0) continue;
            int e = (((a == 1) ? y : z) * j) % n;
            double c = cs[e], s = sign * sn[e];
            re += xr * c - xi * s; im += xr * s + xi * c;
          }
          size_t o = ((size_t)z * n + y) * nh + k;
          dst[2 * o] = re; dst[2 * o + 1] = im;
        }
  }
  // Backward c2r (unnormalised) from [z][y][x'] complex to n^3 reals (x fastest)
  void c2r(const double* in, double* out, std::vector<double>& w) const {
    size_t M = (size_t)n * n * nh;
    std::vector<double> w2(2 * M);
    w.assign(in, in + 2 * M);
    axis(w, w2.data(), 2, +1);
    axis(w2, w.data(), 1, +1);
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          double s = 0;
          for (int k = 0; k < nh; k++) {
            size_t o = ((size_t)z * n + y) * nh + k;
            double re = w[2 * o], im = w[2 * o + 1];
            double mult = (k == 0 || (n % 2 == 0 && k == n / 2)) ? 1.0 : 2.0;
            int e = (k * x) % n;
            s += mult * (re * cs[e] - im * sn[e]);
          }
          out[((size_t)z * n + y) * n + x] = s;
        }
  }
};

// ---------------------------------------------------------------- FMM evaluation
struct Operators {
  int m = 0, nsurf = 0;
  Ker K{0, 1, 1};   // kernel of the targets (s2t, m2t, l2t)
  Ker Km{0, 1, 1};  // kernel of the expansions (s2m, s2l, m2m, m2l, l2l)
  std::vector<double> uc2ue0, uc2ue1, dc2de0, dc2de1;  // level-0 factors
  std::vector<double> u2u[8], d2d[8];                  // by child position bits
  std::vector<double> v1[27];                          // by parent offset (dx+1)+3(dy+1)+9(dz+1); layout [(j+i*k1)*M_dim + f][s*8+t][2]
  double src_scal = 0, trg_scal = 1;                   // m2m/m2l/l2l exponents (kernels here are scale invariant)
  std::vector<double> bc;                              // periodic: root upward equiv -> root downward check (BC_Type)
};

static void matvec_acc(const double* x, int rows, const double* M, int cols, double* y) {
  // y[0:cols] += x[0:rows] * M[rows][cols]  (row-vector convention)
  for (int i = 0; i < rows; i++) {
    double xi = x[i];
    if (xi == 0) continue;
    const double* Mi = &M[(size_t)i * cols];
    for (int j = 0; j < cols; j++) y[j] += xi * Mi[j];
  }
}

struct FMM {
  Tree* T = nullptr;
  Operators* O = nullptr;
  std::vector<std::vector<double>> up, dn;   // upward equiv / downward check (then equiv), per node
  const double* src_val = nullptr;           // input order, k0 per point
  double* trg_val = nullptr;                 // input order, k1 per point
  double t_phase[12] = {0};

  int ns() const { return O->nsurf; }

  // Periodic image shift of source node s in list `type`, slot `k`, of target node t: the
  // source's expected centre (t's centre plus the slot's relative coordinate) minus its actual
  // centre; zero in free space.  Units of the relative coordinate: U1 the target size, U0/X
  // half of it, U2/W a quarter (include/fmm_pts.txx:4646-4656,4934-4940,5228-5340).
  void shift_of(const Node* t, const Node* s, int type, int k, double* sh) const {
    const double st = std::pow(0.5, t->depth), ss = std::pow(0.5, s->depth);
    const double unit = (type == U1_T) ? st : (type == U0_T || type == X_T) ? 0.5 * st : 0.25 * st;
    const int* rel = &T->il.rel[type][3 * k];
    for (int d = 0; d < 3; d++) sh[d] = rel[d] * unit + (t->coord[d] + 0.5 * st) - (s->coord[d] + 0.5 * ss);
  }

  void gather_pts(const Node* n, int set, std::vector<double>& xyz, const double* shift) const {
    const std::vector<int64_t>& v = n->pts[set];
    xyz.resize(3 * v.size());
    for (size_t i = 0; i < v.size(); i++)
      for (int d = 0; d < 3; d++) xyz[3 * i + d] = T->X[set][3 * v[i] + d] + (shift ? shift[d] : 0.0);
  }
  void gather_src_val(const Node* n, int k0, std::vector<double>& v) const {
    const std::vector<int64_t>& idx = n->pts[0];
    v.resize(k0 * idx.size());
    for (size_t i = 0; i < idx.size(); i++) for (int d = 0; d < k0; d++) v[k0 * i + d] = src_val[k0 * idx[i] + d];
  }

  void Run(int nthreads) {
    Operators& o = *O;
    const int m = o.m, n = ns();
    const int k0 = o.Km.k0, k1m = o.Km.k1;
    const size_t NN = T->nodes.size();
    up.assign(NN, {}); dn.assign(NN, {});
    // allocation rule: include/fmm_pts.txx:1316-1420 (children of nodes with pt_cnt>0, plus root)
    for (Node* nd : T->nodes) {
      bool up_alloc = (nd->depth == 0) || (nd->parent && nd->parent->pt_cnt[0]);
      bool dn_alloc = (nd->depth == 0) || (nd->parent && nd->parent->pt_cnt[1]);
      if (up_alloc) up[nd->id].assign((size_t)n * k0, 0.0);
      if (dn_alloc) dn[nd->id].assign((size_t)n * k1m, 0.0);
    }
    const int Nsrf = n;
    const std::vector<Node*>& nodes = T->nodes;
    int max_depth = 0;
    for (Node* nd : nodes) max_depth = std::max(max_depth, nd->depth);
    omp_set_num_threads(nthreads);
    double t0;

    // ---------------- S2U: include/fmm_pts.txx:2275-2599 (EvalListPts with M[2], M[3])
    t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < NN; i++) {
      Node* nd = nodes[i];
      if (!(nd->leaf && !nd->ghost && nd->pts[0].size())) continue;
      double zero[3] = {0, 0, 0};
      std::vector<double> surf = surf_at(m, zero, nd->depth, RAD1);
      double shift[3] = {-nd->coord[0], -nd->coord[1], -nd->coord[2]};
      std::vector<double> sx, sv;
      gather_pts(nd, 0, sx, shift);
      gather_src_val(nd, k0, sv);
      std::vector<double> chk((size_t)n * k1m, 0.0), tmp((size_t)n * k0, 0.0);
      p2p(o.Km, sx.data(), (int64_t)nd->pts[0].size(), sv.data(), surf.data(), n, chk.data());
      double sc = std::pow(2.0, -o.trg_scal * nd->depth);
      for (double& c : chk) c *= sc;
      std::vector<double> mid((size_t)n * k0, 0.0);
      matvec_acc(chk.data(), n * k1m, o.uc2ue0.data(), n * k0, mid.data());
      matvec_acc(mid.data(), n * k0, o.uc2ue1.data(), n * k0, tmp.data());
      double sc2 = std::pow(2.0, -o.src_scal * nd->depth);
      for (size_t j = 0; j < tmp.size(); j++) up[nd->id][j] += tmp[j] * sc2;
    }
    t_phase[0] = omp_get_wtime() - t0;

    // ---------------- U2U: include/fmm_pts.txx:2602-2637, levels max_depth-1 .. 0
    t0 = omp_get_wtime();
    for (int l = max_depth - 1; l >= 0; l--) {
      #pragma omp parallel for schedule(dynamic, 4)
      for (size_t i = 0; i < NN; i++) {
        Node* nd = nodes[i];
        if (nd->depth != l || nd->leaf || !nd->pt_cnt[0]) continue;
        for (int c = 0; c < 8; c++) {
          Node* ch = nd->child[c];
          if (!ch->pt_cnt[0]) continue;
          matvec_acc(up[ch->id].data(), n * k0, o.u2u[c].data(), n * k0, up[nd->id].data());
        }
      }
    }
    // ---------------- periodic boundary: root upward equivalent -> root downward check
    // (FMM_Pts::PeriodicBC, include/fmm_pts.txx:2641-2657, called from include/fmm_tree.txx:582-584)
    if (T->bndry != 0 && !o.bc.empty())
      matvec_acc(up[nodes[0]->id].data(), n * k0, o.bc.data(), n * k1m, dn[nodes[0]->id].data());
    t_phase[1] = omp_get_wtime() - t0;

    // ---------------- X list (S2L): include/fmm_pts.txx:4444-4743
    t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < NN; i++) {
      Node* t = nodes[i];
      if (!(t->pt_cnt[1] && !t->ghost)) continue;
      if (t->leaf && t->pt_cnt[1] <= Nsrf) continue;
      double zero[3] = {0, 0, 0};
      std::vector<double> surf = surf_at(m, zero, t->depth, RAD0);
      for (size_t k = 0; k < t->lst[X_T].size(); k++) {
        Node* s = t->lst[X_T][k];
        if (!s || !(s->leaf && s->pts[0].size())) continue;
        double shift[3];
        shift_of(t, s, X_T, (int)k, shift);
        for (int d = 0; d < 3; d++) shift[d] -= t->coord[d];
        std::vector<double> sx, sv;
        gather_pts(s, 0, sx, shift);
        gather_src_val(s, k0, sv);
        p2p(o.Km, sx.data(), (int64_t)s->pts[0].size(), sv.data(), surf.data(), n, dn[t->id].data());
      }
    }
    t_phase[2] = omp_get_wtime() - t0;

    // ---------------- W list (M2T): include/fmm_pts.txx:4746-5029
    t0 = omp_get_wtime();
    const Ker& K = o.K;
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < NN; i++) {
      Node* t = nodes[i];
      if (!(t->leaf && !t->ghost && t->pts[1].size())) continue;
      std::vector<double> tx, tv(K.k1 * t->pts[1].size(), 0.0);
      gather_pts(t, 1, tx, nullptr);
      for (size_t k = 0; k < t->lst[W_T].size(); k++) {
        Node* s = t->lst[W_T][k];
        if (!s || !s->pt_cnt[0]) continue;
        if (s->leaf && s->pt_cnt[0] <= Nsrf) continue;
        double sh[3];
        shift_of(t, s, W_T, (int)k, sh);
        double sc[3] = {s->coord[0] + sh[0], s->coord[1] + sh[1], s->coord[2] + sh[2]};
        std::vector<double> surf = surf_at(m, sc, s->depth, RAD0);
        p2p(K, surf.data(), n, up[s->id].data(), tx.data(), (int64_t)t->pts[1].size(), tv.data());
      }
      for (size_t j = 0; j < t->pts[1].size(); j++) for (int d = 0; d < K.k1; d++) trg_val[K.k1 * t->pts[1][j] + d] += tv[K.k1 * j + d];
    }
    t_phase[3] = omp_get_wtime() - t0;

    // ---------------- U list (P2P): include/fmm_pts.txx:5032-5424
    t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < NN; i++) {
      Node* t = nodes[i];
      if (!(t->leaf && !t->ghost && t->pts[1].size())) continue;
      std::vector<double> tx, tv(K.k1 * t->pts[1].size(), 0.0), sx, sv;
      gather_pts(t, 1, tx, nullptr);
      auto add = [&](Node* s, int ty, size_t k) {
        double sh[3];
        shift_of(t, s, ty, (int)k, sh);
        gather_pts(s, 0, sx, sh);
        gather_src_val(s, K.k0, sv);
        p2p(K, sx.data(), (int64_t)s->pts[0].size(), sv.data(), tx.data(), (int64_t)t->pts[1].size(), tv.data());
      };
      for (int ty : {U0_T, U1_T, U2_T})
        for (size_t k = 0; k < t->lst[ty].size(); k++) {
          Node* s = t->lst[ty][k];
          if (s && s->leaf && s->pts[0].size()) add(s, ty, k);
        }
      if (t->pt_cnt[1] <= Nsrf)
        for (size_t k = 0; k < t->lst[X_T].size(); k++) {
          Node* s = t->lst[X_T][k];
          if (s && s->leaf && s->pts[0].size()) add(s, X_T, k);
        }
      for (size_t k = 0; k < t->lst[W_T].size(); k++) {
        Node* s = t->lst[W_T][k];
        if (s && s->leaf && s->pts[0].size() && s->pt_cnt[0] <= Nsrf) add(s, W_T, k);
      }
      for (size_t j = 0; j < t->pts[1].size(); j++) for (int d = 0; d < K.k1; d++) trg_val[K.k1 * t->pts[1][j] + d] += tv[K.k1 * j + d];
    }
    t_phase[4] = omp_get_wtime() - t0;

    // ---------------- V list: include/fmm_pts.txx:3346-3759 (FFT_UpEquiv, VListHadamard, FFT_Check2Equiv)
    t0 = omp_get_wtime();
    VList(max_depth);
    t_phase[5] = omp_get_wtime() - t0;

    // ---------------- D2D: include/fmm_pts.txx:3762-3797, levels 1 .. max_depth
    t0 = omp_get_wtime();
    for (int l = 1; l <= max_depth; l++) {
      #pragma omp parallel for schedule(dynamic, 16)
      for (size_t i = 0; i < NN; i++) {
        Node* nd = nodes[i];
        if (nd->depth != l || !nd->pt_cnt[1] || !nd->parent->pt_cnt[1]) continue;
        matvec_acc(dn[nd->parent->id].data(), n * k1m, o.d2d[nd->p2n].data(), n * k1m, dn[nd->id].data());
      }
    }
    t_phase[6] = omp_get_wtime() - t0;

    // ---------------- D2T: include/fmm_pts.txx:5427-5737 (M[0], M[1] then k_l2t from the equiv surface)
    t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < NN; i++) {
      Node* t = nodes[i];
      if (!(t->leaf && !t->ghost && t->pts[1].size())) continue;
      std::vector<double> chk = dn[t->id];
      double sc = std::pow(2.0, -o.trg_scal * t->depth);
      for (double& c : chk) c *= sc;
      std::vector<double> mid((size_t)n * k0, 0.0), eq((size_t)n * k0, 0.0);
      matvec_acc(chk.data(), n * k1m, o.dc2de0.data(), n * k0, mid.data());
      matvec_acc(mid.data(), n * k0, o.dc2de1.data(), n * k0, eq.data());
      double sc2 = std::pow(2.0, -o.src_scal * t->depth);
      for (double& e : eq) e *= sc2;
      std::vector<double> surf = surf_at(m, t->coord, t->depth, RAD1);
      std::vector<double> tx, tv(K.k1 * t->pts[1].size(), 0.0);
      gather_pts(t, 1, tx, nullptr);
      p2p(K, surf.data(), n, eq.data(), tx.data(), (int64_t)t->pts[1].size(), tv.data());
      for (size_t j = 0; j < t->pts[1].size(); j++) for (int d = 0; d < K.k1; d++) trg_val[K.k1 * t->pts[1][j] + d] += tv[K.k1 * j + d];
    }
    t_phase[7] = omp_get_wtime() - t0;

    // ---------------- FMM_Pts::PostProcessing (include/fmm_pts.txx:5740-5774): with PXYZ and
    // a kernel whose M2L carries the volume potential of a uniform density (laplace_poten:
    // -|x|^2/6; stokes_vel: -(|x|^2 - x_j^2)/4 on the diagonal, include/kernel.txx:1289-1295,
    // 1456-1467), subtract the root's total equivalent density times that potential.
    if (T->bndry == 3 && !o.bc.empty() && K.id != 1) {
      double avg[3] = {0, 0, 0};
      const std::vector<double>& ue = up[nodes[0]->id];
      for (size_t i = 0; i < ue.size(); i += k0) for (int c = 0; c < k0; c++) avg[c] += ue[i + c];
      #pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < T->N[1]; i++) {
        const double* x = &T->X[1][3 * i];
        const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        if (K.id == 0) trg_val[i] -= avg[0] * (-r2 / 6);
        else for (int j = 0; j < 3; j++) trg_val[3 * i + j] -= avg[j] * (-(r2 - x[j] * x[j]) / 4);
      }
    }
  }

  void VList(int max_depth) {
    Operators& o = *O;
    const int m = o.m, n = ns(), n1 = 2 * m, n2 = n1 * n1, n3 = n1 * n2;
    const int M_dim = n1 * n1 * (n1 / 2 + 1);
    const int k0 = o.Km.k0, k1 = o.Km.k1;
    // maps, include/fmm_pts.txx:2703-2711 and 2787-2796
    double zero[3] = {0, 0, 0};
    std::vector<double> surf = surface(m, zero, (double)(m - 1), 0);
    std::vector<int> map_in(n), map_out(n);
    for (int i = 0; i < n; i++) {
      map_in[i] = (int)((size_t)(m - 1 - surf[3 * i] + 0.5) + (size_t)(m - 1 - surf[3 * i + 1] + 0.5) * n1 + (size_t)(m - 1 - surf[3 * i + 2] + 0.5) * n2);
      map_out[i] = (int)((size_t)(m * 2 - 0.5 - surf[3 * i]) + (size_t)(m * 2 - 0.5 - surf[3 * i + 1]) * n1 + (size_t)(m * 2 - 0.5 - surf[3 * i + 2]) * n2);
    }
    DFT3 dft(n1);
    const std::vector<Node*>& nodes = T->nodes;
    std::vector<int> src_par, trg_par;
    for (Node* nd : nodes) {
      if (nd->leaf) continue;
      if (nd->pt_cnt[0]) src_par.push_back(nd->id);
      if (nd->pt_cnt[1] && !nd->ghost) trg_par.push_back(nd->id);
    }
    // FFT_UpEquiv: spectra [comp][freq][child][re,im] per source parent
    std::vector<std::vector<double>> X(nodes.size());
    #pragma omp parallel for schedule(dynamic, 4)
    for (size_t ii = 0; ii < src_par.size(); ii++) {
      Node* P = nodes[src_par[ii]];
      std::vector<double> grid(n3), spec(2 * (size_t)M_dim), w;
      std::vector<double>& out = X[P->id];
      out.assign(2 * (size_t)M_dim * 8 * k0, 0.0);
      for (int c = 0; c < 8; c++) {
        Node* ch = P->child[c];
        double fs = std::pow(2.0, o.src_scal * (P->depth + 1));
        for (int i = 0; i < k0; i++) {
          std::fill(grid.begin(), grid.end(), 0.0);
          if (!up[ch->id].empty())
            for (int k = 0; k < n; k++) grid[map_in[k]] = up[ch->id][k0 * k + i] * fs;
          dft.r2c(grid.data(), spec.data(), w);
          for (int f = 0; f < M_dim; f++) {
            out[2 * (8 * ((size_t)M_dim * i + f) + c) + 0] = spec[2 * f];
            out[2 * (8 * ((size_t)M_dim * i + f) + c) + 1] = spec[2 * f + 1];
          }
        }
      }
    }
    // VListHadamard + FFT_Check2Equiv per target parent
    #pragma omp parallel for schedule(dynamic, 4)
    for (size_t ii = 0; ii < trg_par.size(); ii++) {
      Node* P = nodes[trg_par[ii]];
      if (P->depth + 1 > max_depth) continue;
      std::vector<double> Y(2 * (size_t)M_dim * 8 * k1, 0.0);
      for (int r = 0; r < 27; r++) {
        Node* S = P->colleague[r];
        if (r == 13 || !S || S->leaf || !S->pt_cnt[0]) continue;
        if (std::find(P->lst[V1_T].begin(), P->lst[V1_T].end(), S) == P->lst[V1_T].end()) continue;
        const std::vector<double>& Mr = o.v1[r];
        const std::vector<double>& Xs = X[S->id];
        for (int i = 0; i < k0; i++)
          for (int j = 0; j < k1; j++)
            for (int f = 0; f < M_dim; f++) {
              const double* M = &Mr[((size_t)(j + i * k1) * M_dim + f) * 128];
              const double* in = &Xs[2 * 8 * ((size_t)M_dim * i + f)];
              double* ot = &Y[2 * 8 * ((size_t)M_dim * j + f)];
              for (int s = 0; s < 8; s++) {
                double xr = in[2 * s], xi = in[2 * s + 1];
                for (int t = 0; t < 8; t++) {
                  double mr = M[2 * (s * 8 + t)], mi = M[2 * (s * 8 + t) + 1];
                  ot[2 * t] += mr * xr - mi * xi;
                  ot[2 * t + 1] += mr * xi + mi * xr;
                }
              }
            }
      }
      std::vector<double> spec(2 * (size_t)M_dim), grid(n3), w;
      for (int c = 0; c < 8; c++) {
        Node* ch = P->child[c];
        double is = std::pow(2.0, o.trg_scal * (P->depth + 1));
        for (int j = 0; j < k1; j++) {
          for (int f = 0; f < M_dim; f++) {
            spec[2 * f] = Y[2 * (8 * ((size_t)M_dim * j + f) + c)];
            spec[2 * f + 1] = Y[2 * (8 * ((size_t)M_dim * j + f) + c) + 1];
          }
          dft.c2r(spec.data(), grid.data(), w);
          if (!dn[ch->id].empty())
            for (int k = 0; k < n; k++) dn[ch->id][k1 * k + j] += grid[map_out[k]] * is;
        }
      }
    }
  }
};

// ---------------------------------------------------------------- direct sum
static void direct(const Ker& K, const double* sx, int64_t ns, const double* sv, const double* tx, int64_t nt, double* tv, int nthreads) {
  omp_set_num_threads(nthreads);
  const int64_t B = 64;
  #pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t0 = 0; t0 < nt; t0 += B) {
    int64_t t1 = std::min(nt, t0 + B);
    p2p(K, sx, ns, sv, tx + 3 * t0, t1 - t0, tv + K.k1 * t0);
  }
}

}  // namespace orc

// ================================================================ C interface for ctypes
using namespace orc;

struct OracleHandle {
  Tree tree;
  Operators ops;
  std::vector<double> coord_copy[3];
  int kernel = 0;
};

extern "C" {

// sets: src (n_src), trg (n_trg), pt (tree-defining points; pass trg again for PtFMM_CreateTree)
void* oracle_tree_create(const double* src, int64_t n_src, const double* trg, int64_t n_trg,
                         const double* pt, int64_t n_pt, int max_pts, int refine, int bndry) {
  OracleHandle* h = new OracleHandle();
  h->coord_copy[0].assign(src, src + 3 * n_src);
  h->coord_copy[1].assign(trg, trg + 3 * n_trg);
  h->coord_copy[2].assign(pt, pt + 3 * n_pt);
  for (int j = 0; j < 3; j++) { h->tree.X[j] = h->coord_copy[j].data(); }
  h->tree.N[0] = n_src; h->tree.N[1] = n_trg; h->tree.N[2] = n_pt;
  h->tree.max_pts = max_pts;
  h->tree.bndry = bndry;
  h->tree.il.Initialize();
  h->tree.Initialize();          // include/fmm_tree.txx:23-41
  if (refine) h->tree.RefineTree();  // include/fmm_tree.txx:44-69
  h->tree.Balance21();
  h->tree.Finalize();            // SetupFMM: colleagues, pt_cnt, lists
  return h;
}

void oracle_tree_destroy(void* p) { delete (OracleHandle*)p; }

int64_t oracle_tree_num_nodes(void* p) { return (int64_t)((OracleHandle*)p)->tree.nodes.size(); }

// per node (preorder): x, y, z, depth, is_leaf, parent id, n_src, n_trg, pt_cnt0, pt_cnt1
void oracle_tree_nodes(void* p, int64_t* out) {
  OracleHandle* h = (OracleHandle*)p;
  for (Node* n : h->tree.nodes) {
    int64_t* o = &out[10 * n->id];
    o[0] = n->mid.x; o[1] = n->mid.y; o[2] = n->mid.z; o[3] = n->depth; o[4] = n->leaf;
    o[5] = n->parent ? n->parent->id : -1;
    o[6] = (int64_t)n->pts[0].size(); o[7] = (int64_t)n->pts[1].size(); o[8] = n->pt_cnt[0]; o[9] = n->pt_cnt[1];
  }
}

// Points of each leaf in tree order: for set j (0 src, 1 trg) writes original indices.
void oracle_tree_point_order(void* p, int set, int64_t* out) {
  OracleHandle* h = (OracleHandle*)p;
  int64_t k = 0;
  for (Node* n : h->tree.nodes) if (n->leaf) for (int64_t i : n->pts[set]) out[k++] = i;
}

int oracle_list_count(void* p, int type) { return ((OracleHandle*)p)->tree.il.count(type); }
void oracle_list_rel(void* p, int type, int* out) {
  OracleHandle* h = (OracleHandle*)p;
  std::copy(h->tree.il.rel[type].begin(), h->tree.il.rel[type].end(), out);
}
// interaction list of one type: out[node][count] = node id or -1 (nodes without a list get all -1)
void oracle_tree_lists(void* p, int type, int32_t* out) {
  OracleHandle* h = (OracleHandle*)p;
  int c = h->tree.il.count(type);
  for (Node* n : h->tree.nodes) {
    const std::vector<Node*>& L = n->lst[type];
    for (int i = 0; i < c; i++) out[(int64_t)n->id * c + i] = (i < (int)L.size() && L[i]) ? L[i]->id : -1;
  }
}

// Operators from the Python side (level-0 factors, per-child U2U/D2D, V1 blocks by parent offset).
void oracle_set_operators(void* p, int kernel, int m, const double* uc2ue0, const double* uc2ue1,
                          const double* dc2de0, const double* dc2de1, const double* u2u, const double* d2d,
                          const double* v1) {
  OracleHandle* h = (OracleHandle*)p;
  Operators& o = h->ops;
  h->kernel = kernel;
  o.m = m; o.nsurf = 6 * (m - 1) * (m - 1) + 2;
  o.K = ker_of(kernel);
  o.Km = (kernel == 1) ? ker_of(0) : ker_of(kernel);
  int n = o.nsurf, k0 = o.Km.k0, k1 = o.Km.k1;
  size_t sq = (size_t)n * k0 * n * k0;
  o.uc2ue0.assign(uc2ue0, uc2ue0 + sq); o.uc2ue1.assign(uc2ue1, uc2ue1 + sq);
  o.dc2de0.assign(dc2de0, dc2de0 + sq); o.dc2de1.assign(dc2de1, dc2de1 + sq);
  for (int c = 0; c < 8; c++) { o.u2u[c].assign(u2u + c * sq, u2u + (c + 1) * sq); o.d2d[c].assign(d2d + c * sq, d2d + (c + 1) * sq); }
  int n1 = 2 * m;
  size_t M_dim = (size_t)n1 * n1 * (n1 / 2 + 1);
  size_t vsz = (size_t)k0 * k1 * M_dim * 128;
  for (int r = 0; r < 27; r++) { if (r == 13) continue; o.v1[r].assign(v1 + r * vsz, v1 + (r + 1) * vsz); }
  o.src_scal = 0.0; o.trg_scal = 1.0;  // potential-type expansion kernels (include/kernel.txx:72-199)
}

// Periodic boundary operator (n_surf*k0 x n_surf*k1 of the expansion kernel), from the
// Python precompute; without it a periodic tree evaluates only the wrapped lists.
void oracle_set_bc(void* p, const double* bc) {
  OracleHandle* h = (OracleHandle*)p;
  const Operators& o = h->ops;
  size_t sz = (size_t)o.nsurf * o.Km.k0 * o.nsurf * o.Km.k1;
  h->ops.bc.assign(bc, bc + sz);
}

// Evaluate: src_val (input order, k0 per source), trg_val (input order, k1 per target, overwritten)
void oracle_evaluate(void* p, const double* src_val, double* trg_val, int nthreads, double* phase_times) {
  OracleHandle* h = (OracleHandle*)p;
  FMM f;
  f.T = &h->tree; f.O = &h->ops; f.src_val = src_val; f.trg_val = trg_val;
  std::fill(trg_val, trg_val + (size_t)h->ops.K.k1 * h->tree.N[1], 0.0);
  f.Run(nthreads);
  if (phase_times) for (int i = 0; i < 8; i++) phase_times[i] = f.t_phase[i];
}

void oracle_direct(int kernel, const double* sx, int64_t ns, const double* sv, const double* tx, int64_t nt,
                   double* tv, int nthreads) {
  Ker K = ker_of(kernel);
  std::fill(tv, tv + (size_t)K.k1 * nt, 0.0);
  direct(K, sx, ns, sv, tx, nt, tv, nthreads);
}

// surface() of include/fmm_pts.txx:54-97, for tests and the Python operator precompute.
void oracle_surface(int p, const double* c, double alpha, int depth, double* out) {
  std::vector<double> s = surface(p, c, alpha, depth);
  std::copy(s.begin(), s.end(), out);
}

// 3-D r2c/c2r of the oracle's DFT (for the FFT-vs-numpy self-check).
void oracle_dft_r2c(int n, const double* in, double* out) { DFT3 d(n); std::vector<double> w; d.r2c(in, out, w); }
void oracle_dft_c2r(int n, const double* in, double* out) { DFT3 d(n); std::vector<double> w; d.c2r(in, out, w); }

}  // extern "C"
