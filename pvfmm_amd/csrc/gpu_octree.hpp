// Octree and interaction-list construction on the GPU (implemented in gpu_octree.hip).
#pragma once
#include <cstdint>

#include "octree.hpp"

namespace pvfmm_amd {

// The tree of build_octree (octree.hpp), built on the device: same nodes, point orders and
// lists, bit for bit, for every boundary type.  The colleague table and the sorted integer coordinates of the point
// sets, which only the host builder keeps, are left empty.
void build_octree_device(Octree& T, const double* src, int64_t n_src, const double* trg, int64_t n_trg, const double* pt,
                         int64_t n_pt, int max_pts, int bndry = kFreeSpace);

// Tree construction runs on the GPU when one is visible, unless the environment sets
// PVFMM_TREE=host; the host builder serves machines without one.
bool device_tree_enabled();

}  // namespace pvfmm_amd
