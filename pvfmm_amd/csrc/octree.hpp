// Linear (level-major) adaptive octree with 2:1 balance and the reference's interaction
// lists, built on the host from sorted 90-bit Morton keys.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace pvfmm_amd {

// Interaction list types, numbered as Mat_Type in include/precomp_mat.hpp:21-39.
enum ListType : int {
  kS2U = 4, kU2U = 5, kD2D = 6, kD2T = 7, kU0 = 8, kU1 = 9, kU2 = 10, kV = 11, kW = 12, kX = 13, kV1 = 14,
  kTypeCount = 16
};

// Relative-coordinate table of one list type, in the reference's symmetry-class order
// (include/interac_list.txx:367-404).
struct RelTable {
  std::vector<std::array<int, 3>> rel;
  std::vector<int> lut;  // coord_hash -> index or -1
  static int hash(int x, int y, int z) { return ((z + 5) * 10 + (y + 5)) * 10 + (x + 5); }
  int index(int x, int y, int z) const {
    if (x < -5 || x > 5 || y < -5 || y > 5 || z < -5 || z > 5) return -1;
    return lut[hash(x, y, z)];
  }
};
const RelTable& rel_table(int type);

struct PointSet {
  int64_t n = 0;
  std::vector<int64_t> perm;       // sorted position -> input index
  std::vector<uint32_t> ix, iy, iz;  // 30-bit integer coordinates, sorted
};

// Boundary types of the reference (BoundaryType, include/mpi_tree.hpp:22-29).
enum Bndry : int { kFreeSpace = 0, kPX = 1, kPXY = 2, kPXYZ = 3 };

struct Octree {
  int max_pts = 0;
  int bndry = kFreeSpace;            // periodic axes: PX x; PXY x, y; PXYZ all (unit period)
  int max_depth = 0;                 // deepest level present
  // nodes in level-major order; siblings are contiguous (first_child .. first_child+7)
  std::vector<uint32_t> x, y, z;     // integer corner coordinates (units of 2^-30)
  std::vector<uint8_t> depth;
  std::vector<int32_t> parent, first_child, child_pos;
  std::vector<int64_t> src_begin, src_end, trg_begin, trg_end;  // subtree point ranges (sorted order)
  std::vector<int32_t> level_begin;  // nodes of level l: [level_begin[l], level_begin[l+1])
  std::vector<int32_t> colleague;    // n_nodes x 27, index (dz+1)*9+(dy+1)*3+(dx+1), -1 if absent
  PointSet src, trg;
  std::vector<int32_t> lists[kTypeCount];  // n_nodes x count(type), -1 = empty slot

  int32_t n_nodes() const { return (int32_t)depth.size(); }
  bool is_leaf(int32_t i) const { return first_child[i] < 0; }
  int64_t pt_cnt0(int32_t i) const { return src_end[i] - src_begin[i]; }
  int64_t pt_cnt1(int32_t i) const { return trg_end[i] - trg_begin[i]; }
  double size(int32_t i) const { return 1.0 / (double)(1u << depth[i]); }
  void corner(int32_t i, double* c) const {
    const double s = 1.0 / (double)(1u << 30);
    c[0] = x[i] * s; c[1] = y[i] * s; c[2] = z[i] * s;
  }
  int list_count(int type) const { return (int)rel_table(type).rel.size(); }
};

// Build the tree of PtFMM_CreateTree + SetupFMM for the given point sets.  `pt` is the set the
// reference stores in MPI_Node::pt_coord (the targets for PtFMM_CreateTree, include/pvfmm.txx:239).
// Coordinates must lie in [0,1)^3.  bndry (a Bndry value) selects periodic neighbours: the 2:1
// balance then wraps every axis (balanceOctree's periodic flag, include/mpi_tree.txx:745-967)
// and the colleagues wrap along the periodic axes (SetColleagues, include/mpi_tree.txx:1136-1170).
void build_octree(Octree& T, const double* src, int64_t n_src, const double* trg, int64_t n_trg,
                  const double* pt, int64_t n_pt, int max_pts, int bndry = kFreeSpace);

// Periodic image shift (whole periods per axis, in {-1, 0, 1}) of the source node s that sits in
// slot k of the list of type `type` of target node t: the slot's relative coordinate places the
// source's centre at t's centre plus rel * unit (unit: t's size for U1 and V1, half of it for U0
// and X, a quarter for U2 and W), and the shift is that place minus s's centre
// (coord_shift, include/fmm_pts.txx:4646-4656,4934-4940,5228-5340).  Zero in free space.
void image_shift(const Octree& T, int32_t t, int32_t s, int type, int k, int* shift);

}  // namespace pvfmm_amd
