"""CPU parity of the product's host side against the oracle: the octree (leaf Morton ids,
depths, point counts) and every interaction list must be bit-exact; the FP64 operators must
reproduce the physics (pseudo-inverse property, exact FFT-M2L == dense M2L, far-field
accuracy of S2M/U2U/D2D/L2T) at the level of the oracle's LAPACK-based operators."""
import numpy as np
import pytest

import oracle as O
import pvfmm_amd as P
import treecheck
from pvfmm_amd import workloads as W

RNG = np.random.default_rng(99)
LIST_TYPES = [O.S2U, O.U2U, O.D2D, O.D2T, O.U0, O.U1, O.U2, O.W, O.X, O.V1]


def _points(kind, n, rng):
    if kind == "uniform":
        return rng.random((n, 3))
    if kind == "sphere":
        d = rng.standard_normal((n, 3))
        return 0.5 + 0.4 * d / np.linalg.norm(d, axis=1)[:, None]
    if kind == "blob":
        return np.clip(rng.normal(0.3, 0.02, (n, 3)), 0, 1 - 1e-12)
    if kind == "cubed":
        return rng.random((n, 3)) ** 3
    if kind == "dup":  # duplicated points: leaves stop at depth 29
        x = rng.random((n // 2, 3))
        return np.concatenate([x, x])
    raise ValueError(kind)


def assert_same_tree(src, trg, max_pts, pt=None):
    ot = O.OracleTree(src, trg, max_pts, pt=pt)
    pt_ = P.PtFMMTree(src, trg, max_pts, pt=pt)
    on, pn = ot.nodes(), pt_.nodes()
    okey = [tuple(r[:4]) for r in on]
    pkey = [tuple(r[:4]) for r in pn]
    assert sorted(okey) == sorted(pkey)
    pidx = {k: i for i, k in enumerate(pkey)}
    perm = np.array([pidx[k] for k in okey])
    assert np.array_equal(on[:, 4], pn[perm, 4])        # leaf flags
    assert np.array_equal(on[:, 6:10], pn[perm, 6:10])  # point counts
    for t in LIST_TYPES:
        ol, pl = ot.lists(t), pt_.lists(t)
        assert ol.shape[1] == pl.shape[1]
        mapped = np.where(pl[perm] >= 0, perm.argsort()[np.maximum(pl[perm], 0)], -1)
        assert np.array_equal(ol, mapped), f"list type {t} differs"
    # same points in every leaf
    for which in (0, 1):
        oo, po = ot.point_order(which), pt_.point_order(which)
        cnt = 6 + which
        ocounts = on[on[:, 4] == 1][:, cnt]
        # leaves in oracle preorder == Morton order; product leaves sorted by Morton key too
        o_sets = np.split(oo, np.cumsum(ocounts)[:-1])
        leaves_p = pn[pn[:, 4] == 1]
        order = np.lexsort(np.array([_morton(r) for r in leaves_p]).T[::-1])
        p_counts = leaves_p[order][:, cnt]
        p_sets = np.split(po, np.cumsum(p_counts)[:-1])
        assert len(o_sets) == len(p_sets)
        for a, b in zip(o_sets, p_sets):
            assert set(a.tolist()) == set(b.tolist())
    return on


def _morton(row):
    """Sortable key of a node (x, y, z, depth) in MortonId order: interleaved bits, then depth."""
    x, y, z, d = (int(v) for v in row[:4])
    key = 0
    for b in range(29, -1, -1):
        key = (key << 3) | (((z >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((x >> b) & 1)
    return (key >> 60, key & ((1 << 60) - 1), d)


@pytest.mark.parametrize("seed", range(6))
def test_tree_and_lists_bit_exact(seed):
    rng = np.random.default_rng(1000 + seed)
    kinds = ["uniform", "sphere", "blob", "cubed", "uniform", "dup"]
    n = int(rng.integers(100, 5000))
    x = _points(kinds[seed], n, rng)
    max_pts = int(rng.choice([16, 40, 128] if kinds[seed] == "dup" else [4, 16, 40, 128]))
    trg = x if seed % 2 == 0 else _points("uniform", int(rng.integers(50, 3000)), rng)
    nd = assert_same_tree(x, trg, max_pts)
    treecheck.assert_minimal(nd, max_pts)
    treecheck.assert_minimal(P.PtFMMTree(x, trg, max_pts).nodes(), max_pts)


@pytest.mark.parametrize("name,n", [("adaptive-grad-2m", 1 << 21), ("ellipsoid-stokes-1m", 1 << 20),
                                    ("uniform-poten-64k", 1 << 16)])
def test_tree_and_lists_bit_exact_at_workload_parameters(name, n):
    """The benchmark point sets at full size (2^21 points of the adaptive sphere/blob/uniform
    mix, 2^20 on the ellipsoid surface, 65,536 uniform) at their own max_pts: tree, lists and
    leaf point sets bit-exact against the oracle, and the tree minimal."""
    c = W.WORKLOADS[name](n) if name != "uniform-poten-64k" else W.uniform_poten_64k()
    nd = assert_same_tree(c["src"], c["trg"], c["max_pts"])
    treecheck.assert_minimal(nd, c["max_pts"])


def test_tree_with_separate_defining_points():
    """The C-ABI context tree keeps its initial grid points in MPI_Node::pt_coord
    (src/pvfmm-wrapper.cpp:416-432); both builders must honour that extra split condition."""
    x = RNG.random((800, 3)) * 0.3
    nn = 6
    g = (np.arange(nn) + 0.5) / nn
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)[:, ::-1].copy()
    assert_same_tree(x, x, 128, pt=grid)


def test_tiny_inputs():
    assert_same_tree(np.array([[0.5, 0.5, 0.5]]), np.array([[0.25, 0.75, 0.1]]), 16)
    x = RNG.random((5, 3))
    assert_same_tree(x, x, 100)  # root only


def test_rejects_points_outside_unit_cube():
    with pytest.raises(Exception):
        P.PtFMMTree(np.array([[0.5, 1.0, 0.2]]), None, 10)


# ----------------------------------------------------------------------------- operators
def _ops(kernel, m, prec="fp64"):
    f = P.PtFMM(kernel, m, prec)
    ns = 6 * (m - 1) ** 2 + 2
    e = 3 if P.kernel_id(kernel) == 3 else 1
    K = ns * e
    a0 = f.operator("uc2ue0")
    ru = a0.size // K
    d0 = f.operator("dc2de0")
    rd = d0.size // K
    return dict(f=f, K=K, ns=ns, uc2ue0=a0.reshape(K, ru), uc2ue1=f.operator("uc2ue1").reshape(ru, K),
                dc2de0=d0.reshape(K, rd), dc2de1=f.operator("dc2de1").reshape(rd, K),
                u2u=f.operator("u2u").reshape(8, K, K), d2d=f.operator("d2d").reshape(K, 8, K).transpose(1, 0, 2))


@pytest.mark.parametrize("kernel,kid", [("laplace_poten", 0), ("stokes_vel", 2)])
def test_pinv_property(kernel, kid):
    m = 6
    o = _ops(kernel, m)
    uc = O.surf_at(m, [0, 0, 0], 0, O.RAD1)
    ue = O.surf_at(m, [0, 0, 0], 0, O.RAD0)
    M = O.kernel_matrix(kid, ue, uc)
    c = RNG.standard_normal(M.shape[0]) @ M
    e = (c @ o["uc2ue0"]) @ o["uc2ue1"]
    assert np.linalg.norm(e @ M - c) / np.linalg.norm(c) < 1e-8


@pytest.mark.parametrize("m", [6, 10])
def test_s2m_u2u_far_field_vs_oracle(m):
    """S2M then U2U then M2T at far points: the product's Jacobi-SVD factors must be at
    least as accurate as the oracle's LAPACK factors."""
    o = _ops("laplace_poten", m)
    ref = O.precompute("laplace_poten", m)
    far = np.array([[3.0, 2.5, -1.5], [0.5, 0.5, 4.0], [-2.0, 0.2, 0.3]])
    Ke = O.kernel_matrix(0, O.surf_at(m, [0, 0, 0], 0, O.RAD0), far).T
    for c in (0, 3, 7):
        bits = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1]) * 0.5
        src = bits + 0.5 * RNG.random((40, 3))
        q = RNG.standard_normal(40)
        chk = O.kernel_matrix(0, src, O.surf_at(m, bits, 1, O.RAD1)).T @ q
        exact = O.kernel_matrix(0, src, far).T @ q
        errs = []
        for a0, a1, u2u in ((o["uc2ue0"], o["uc2ue1"], o["u2u"][c]), (ref["uc2ue0"], ref["uc2ue1"], ref["u2u"][c])):
            e = ((chk @ a0) @ a1 * 0.5) @ u2u
            errs.append(np.linalg.norm(Ke @ e - exact) / np.linalg.norm(exact))
        assert errs[0] < max(3 * errs[1], 1e-9), errs


@pytest.mark.parametrize("m", [6, 10])
def test_d2d_l2t_vs_oracle(m):
    o = _ops("laplace_poten", m)
    ref = O.precompute("laplace_poten", m)
    src = RNG.random((60, 3)) + np.array([2.2, 0.1, -0.3])
    q = RNG.standard_normal(60)
    pchk = O.kernel_matrix(0, src, O.surf_at(m, [0, 0, 0], 0, O.RAD0)).T @ q
    for c in (0, 6):
        bits = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1]) * 0.5
        trg = bits + 0.5 * RNG.random((30, 3))
        exact = O.kernel_matrix(0, src, trg).T @ q
        errs = []
        for d2d, a0, a1 in ((o["d2d"][c], o["dc2de0"], o["dc2de1"]), (ref["d2d"][c], ref["dc2de0"], ref["dc2de1"])):
            eq = ((pchk @ d2d) * 0.5 @ a0) @ a1
            pot = O.kernel_matrix(0, O.surf_at(m, bits, 1, O.RAD1), trg).T @ eq
            errs.append(np.linalg.norm(pot - exact) / np.linalg.norm(exact))
        assert errs[0] < max(3 * errs[1], 1e-9), errs


@pytest.mark.parametrize("kernel,kid", [("laplace_poten", 0), ("stokes_vel", 2)])
def test_product_m2l_spectra_equal_dense(kernel, kid):
    """The product's M2L blocks (own DFT convention, no index reversal) applied through numpy
    FFTs reproduce the dense equivalent->check interaction exactly."""
    m = 6
    f = P.PtFMM(kernel, m)
    e = 3 if kid == 2 else 1
    n1, nh = 2 * m, m + 1
    Md = n1 * n1 * nh
    ns = 6 * (m - 1) ** 2 + 2
    V = f.operator("v1").reshape(27, Md, 8 * e, 8 * e, 2)
    V = V[..., 0] + 1j * V[..., 1]
    lat = np.rint(O.surface(m, [0, 0, 0], m - 1, 0)).astype(int)
    for R in (0, 10, 26):
        Rv = np.array([R % 3 - 1, (R // 3) % 3 - 1, R // 9 - 1])
        X = RNG.standard_normal((8, ns, e))
        Xg = np.zeros((8, e, n1, n1, n1))
        for c in range(8):
            for i in range(e):
                Xg[c, i, lat[:, 2], lat[:, 1], lat[:, 0]] = X[c, :, i]
        Xh = np.fft.rfftn(Xg, axes=(2, 3, 4)).reshape(8 * e, Md)
        Yh = np.einsum("frs,sf->rf", V[R], Xh)
        Y = np.fft.irfftn(Yh.reshape(8 * e, n1, n1, nh), s=(n1, n1, n1), axes=(1, 2, 3)) * n1 ** 3
        Y = Y.reshape(8, e, n1, n1, n1)
        y = np.stack([Y[t][:, lat[:, 2], lat[:, 1], lat[:, 0]].T for t in range(8)])  # [t][p][j]
        dense = np.zeros((8, ns, e))
        for t in range(8):
            bt = np.array([t & 1, (t >> 1) & 1, (t >> 2) & 1])
            chk = O.surf_at(m, bt * 1.0, 0, O.RAD0)
            for s in range(8):
                bs = np.array([s & 1, (s >> 1) & 1, (s >> 2) & 1]) + 2 * Rv
                if np.abs(bs - bt).max() < 2:
                    continue
                eq = O.surf_at(m, bs * 1.0, 0, O.RAD0)
                dense[t] += (X[s].ravel() @ O.kernel_matrix(kid, eq, chk)).reshape(ns, e)
        assert np.linalg.norm(y - dense) / np.linalg.norm(dense) < 1e-12


# ----------------------------------------------------------------------------- GPU tree build
def _tree_arrays(t):
    return t.nodes(), t.point_order(0), t.point_order(1), {ty: t.lists(ty) for ty in LIST_TYPES}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(range(6)) + ["adaptive-grad-2m", "ellipsoid-stokes-1m", "grid-pts"])
def test_device_tree_matches_host_and_oracle(gpu, monkeypatch, case):
    """The GPU tree builder (the default on a GPU) against the host builder (PVFMM_TREE=host):
    nodes, point orders and every list identical; and the device tree against the oracle as in
    the CPU tests (bit-exact lists and leaf point sets, minimal).  The cases are the seeded
    distributions of test_tree_and_lists_bit_exact (duplicated points included), the benchmark
    point sets at full size and a separate tree-defining point set."""
    import time

    pt = None
    if isinstance(case, int):
        rng = np.random.default_rng(1000 + case)
        kinds = ["uniform", "sphere", "blob", "cubed", "uniform", "dup"]
        n = int(rng.integers(100, 5000))
        x = _points(kinds[case], n, rng)
        max_pts = int(rng.choice([16, 40, 128] if kinds[case] == "dup" else [4, 16, 40, 128]))
        trg = x if case % 2 == 0 else _points("uniform", int(rng.integers(50, 3000)), rng)
    elif case == "grid-pts":
        x = RNG.random((800, 3)) * 0.3
        g = (np.arange(6) + 0.5) / 6
        pt = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)[:, ::-1].copy()
        trg, max_pts = x, 128
    else:
        c = W.WORKLOADS[case]()
        x, trg, max_pts = c["src"], c["trg"], c["max_pts"]
    monkeypatch.delenv("PVFMM_TREE", raising=False)
    t0 = time.perf_counter()
    td = P.PtFMMTree(x, trg, max_pts, pt=pt)
    t_dev = time.perf_counter() - t0
    monkeypatch.setenv("PVFMM_TREE", "host")
    t0 = time.perf_counter()
    th = P.PtFMMTree(x, trg, max_pts, pt=pt)
    t_host = time.perf_counter() - t0
    monkeypatch.delenv("PVFMM_TREE")
    print(f"{case}: {len(x)} points, {td.num_nodes()} nodes; tree build device {t_dev * 1e3:.1f} ms, host {t_host * 1e3:.1f} ms")
    nd, od0, od1, ld = _tree_arrays(td)
    nh, oh0, oh1, lh = _tree_arrays(th)
    assert np.array_equal(nd, nh)
    assert np.array_equal(od0, oh0) and np.array_equal(od1, oh1)
    for ty in LIST_TYPES:
        assert np.array_equal(ld[ty], lh[ty]), ty
    if len(x) <= (1 << 21):
        on = assert_same_tree(x, trg, max_pts, pt=pt)  # the device tree (default) against the oracle
        if pt is None:
            treecheck.assert_minimal(on, max_pts)
