"""Structural checks of an octree given as the (n_nodes, 10) node table of the oracle and the
product (x, y, z in units of 2^-30, depth, is_leaf, parent, n_src, n_trg, pt_cnt0, pt_cnt1)."""
import numpy as np


def assert_minimal(nd, max_pts):
    """No leaf's parent could be coarsened: every internal node whose 8 children are all leaves
    holds more than max_pts source + target points (the split rule, include/fmm_node.txx:62-95
    with the tree-defining points = targets), or is touched by a leaf at least two levels finer,
    which a coarsened node would leave 2:1-unbalanced (include/mpi_tree.txx:745-1066)."""
    nd = np.asarray(nd)
    depth = nd[:, 3]
    leaf = nd[:, 4] == 1
    par = nd[:, 5]
    lo = nd[:, :3].astype(np.int64)
    size = np.left_shift(np.int64(1), (30 - depth).astype(np.int64))
    n_child = np.bincount(par[par >= 0], minlength=len(nd))
    n_leaf_child = np.bincount(par[(par >= 0) & leaf], minlength=len(nd))
    cand = np.where(~leaf & (n_child == 8) & (n_leaf_child == 8))[0]
    assert np.all(n_child[~leaf] == 8), "an internal node without 8 children"
    llo, lsz, ldep = lo[leaf], size[leaf], depth[leaf]
    checked = 0
    for p in cand:
        if nd[p, 8] + nd[p, 9] > max_pts:
            continue
        hi = lo[p] + size[p]
        touch = np.all((llo <= hi) & (lo[p] <= llo + lsz[:, None]), axis=1)
        assert np.any(touch & (ldep >= depth[p] + 2)), \
            f"node {p} (depth {depth[p]}, {nd[p, 8] + nd[p, 9]} points) could be coarsened"
        checked += 1
    return len(cand), checked
