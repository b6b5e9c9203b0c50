"""Pins the oracle (the CPU restatement of the reference) with what the mathematics fixes:
closed-form kernels, a brute-force direct sum, exact identities of the FFT M2L, the
pseudo-inverse property, geometric definitions of the tree and of every interaction list,
and convergence of the whole FMM in the multipole order.  The reference ships no golden
outputs (it has no test suite), so these checks are what anchors the oracle."""
import numpy as np
import pytest

import oracle as O
import treecheck

RNG = np.random.default_rng(1234)


# ----------------------------------------------------------------------------- kernels
def test_kernel_closed_forms():
    s = np.array([[0.1, 0.2, 0.3]])
    t = np.array([[0.4, 0.6, 0.3], [0.1, 0.2, 0.3]])  # second target coincides with the source
    d = t[0] - s[0]
    r = np.linalg.norm(d)
    # laplace potential 1/(4 pi r); r = 0 pairs contribute nothing (include/kernel.txx:1283)
    u = O.direct("laplace_poten", s, np.array([2.0]), t)
    assert np.allclose(u, [2.0 / (4 * np.pi * r), 0.0], rtol=1e-15, atol=0)
    # gradient kernel: q (x_t - x_s)/(4 pi r^3) = -grad phi (include/kernel.txx:1316-1330)
    g = O.direct("laplace_grad", s, np.array([2.0]), t).reshape(2, 3)
    assert np.allclose(g[0], 2.0 * d / (4 * np.pi * r ** 3), rtol=1e-14)
    assert np.all(g[1] == 0)
    # stokes velocity (f + r (f.r)/r^2)/(8 pi r)
    f = np.array([0.3, -1.0, 0.5])
    v = O.direct("stokes_vel", s, f, t).reshape(2, 3)
    assert np.allclose(v[0], (f / r + d * (f @ d) / r ** 3) / (8 * np.pi), rtol=1e-14)


def test_gradient_is_minus_grad_potential():
    s = RNG.random((30, 3))
    q = RNG.standard_normal(30)
    t = np.array([[1.7, 0.2, -0.4]])
    h = 1e-5
    g = O.direct("laplace_grad", s, q, t)
    fd = []
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd.append((O.direct("laplace_poten", s, q, t + e) - O.direct("laplace_poten", s, q, t - e))[0] / (2 * h))
    assert np.allclose(g, -np.array(fd), rtol=1e-7)


def test_direct_sum_translation_invariance():
    s = RNG.random((50, 3))
    q = RNG.standard_normal(50)
    t = RNG.random((20, 3)) + 2
    sh = np.array([0.37, -1.2, 5.0])
    assert np.allclose(O.direct(0, s, q, t), O.direct(0, s + sh, q, t + sh), rtol=1e-12)


# ----------------------------------------------------------------------------- DFT and operators
def test_oracle_dft_matches_numpy():
    n = 12
    x = np.zeros((n, n, n))
    x[:6, :6, :6] = RNG.standard_normal((6, 6, 6))
    out = np.zeros(2 * n * n * (n // 2 + 1))
    O.lib().oracle_dft_r2c(n, O._dp(np.ascontiguousarray(x)), O._dp(out))
    ref = np.fft.rfftn(x)
    got = out[0::2] + 1j * out[1::2]
    assert np.allclose(got, ref.ravel(), atol=1e-11)
    back = np.zeros(n ** 3)
    O.lib().oracle_dft_c2r(n, O._dp(np.ascontiguousarray(out)), O._dp(back))
    assert np.allclose(back / n ** 3, x.ravel(), atol=1e-12)


@pytest.mark.parametrize("m", [4, 6])
def test_pinv_factors_reproduce_check_potential(m):
    ops = O.precompute("laplace_poten", m)
    uc = O.surf_at(m, [0, 0, 0], 0, O.RAD1)
    ue = O.surf_at(m, [0, 0, 0], 0, O.RAD0)
    M = O.kernel_matrix(0, ue, uc)
    c = RNG.standard_normal(len(ue)) @ M  # a check potential produced by an equivalent density
    e = (c @ ops["uc2ue0"]) @ ops["uc2ue1"]
    assert np.linalg.norm(e @ M - c) / np.linalg.norm(c) < 1e-8


@pytest.mark.parametrize("m", [4, 6])
def test_m2l_fft_equals_dense(m):
    """FFT_UpEquiv -> Hadamard -> FFT_Check2Equiv (reference conventions) == dense M2L."""
    ops = O.precompute("laplace_poten", m)
    n1 = 2 * m
    nh = m + 1
    M_dim = n1 * n1 * nh
    surf = O.surface(m, [0, 0, 0], m - 1, 0)
    map_in = ((m - 1 - surf[:, 0] + 0.5).astype(int) + (m - 1 - surf[:, 1] + 0.5).astype(int) * n1
              + (m - 1 - surf[:, 2] + 0.5).astype(int) * n1 * n1)
    map_out = ((2 * m - 0.5 - surf[:, 0]).astype(int) + (2 * m - 0.5 - surf[:, 1]).astype(int) * n1
               + (2 * m - 0.5 - surf[:, 2]).astype(int) * n1 * n1)
    ns = len(surf)
    for R in [0, 4, 22]:  # corner, edge and face parent offsets
        Rv = np.array([R % 3 - 1, (R // 3) % 3 - 1, R // 9 - 1])
        X = RNG.standard_normal((8, ns))
        g = np.zeros((8, n1 ** 3))
        for c in range(8):
            g[c, map_in] = X[c]
        Xh = np.fft.rfftn(g.reshape(8, n1, n1, n1), axes=(1, 2, 3)).reshape(8, M_dim)
        V = ops["v1"][R].reshape(M_dim, 8, 8, 2)  # [f][s][t]
        V = V[..., 0] + 1j * V[..., 1]
        Yh = np.einsum("fst,sf->tf", V, Xh)
        Y = np.fft.irfftn(Yh.reshape(8, n1, n1, nh), s=(n1, n1, n1), axes=(1, 2, 3)).reshape(8, -1) * n1 ** 3
        y = Y[:, map_out]
        dense = np.zeros((8, ns))
        for t in range(8):
            bt = np.array([t & 1, (t >> 1) & 1, (t >> 2) & 1])
            chk = O.surf_at(m, bt * 1.0, 0, O.RAD0)
            for s in range(8):
                bs = np.array([s & 1, (s >> 1) & 1, (s >> 2) & 1]) + 2 * Rv
                if np.abs(bs - bt).max() < 2:
                    continue
                eq = O.surf_at(m, bs * 1.0, 0, O.RAD0)
                dense[t] += O.kernel_matrix(0, eq, chk).T @ X[s]
        assert np.linalg.norm(y - dense) / np.linalg.norm(dense) < 1e-12


# ----------------------------------------------------------------------------- tree invariants
def _morton_box(row):
    s = 2.0 ** -row[3]
    lo = row[:3] * 2.0 ** -30
    return lo, lo + s


def _check_tree_invariants(tr, src, trg, max_pts):
    nd = tr.nodes()
    leaves = nd[nd[:, 4] == 1]
    # leaves tile the unit cube
    vol = np.sum(8.0 ** -leaves[:, 3].astype(float))
    assert abs(vol - 1.0) < 1e-12
    # point counts: every point in exactly one leaf, counts match geometry
    assert leaves[:, 6].sum() == len(src) and leaves[:, 7].sum() == len(trg)
    for row in leaves[RNG.choice(len(leaves), min(40, len(leaves)), replace=False)]:
        lo, hi = _morton_box(row)
        ns = np.sum(np.all((src >= lo) & (src < hi), axis=1))
        nt = np.sum(np.all((trg >= lo) & (trg < hi), axis=1))
        assert ns == row[6] and nt == row[7]
    # split rule: leaf counts within max_pts (src+trg) unless at depth 29
    assert np.all((leaves[:, 6] + leaves[:, 7] <= max_pts) | (leaves[:, 3] == 29))
    # every node holding more than max_pts points (below depth 29) is split
    heavy = (nd[:, 8] + nd[:, 9] > max_pts) & (nd[:, 3] < 29)
    assert np.all(nd[heavy, 4] == 0)
    # 2:1 balance: adjacent leaves differ by at most one level
    lo = leaves[:, :3] * 2.0 ** -30
    size = 2.0 ** -leaves[:, 3]
    for i in RNG.choice(len(leaves), min(60, len(leaves)), replace=False):
        touch = np.all((lo <= lo[i] + size[i] + 1e-15) & (lo + size[:, None] >= lo[i] - 1e-15), axis=1)
        assert np.all(np.abs(leaves[touch, 3] - leaves[i, 3]) <= 1)
    # minimality: no leaf's parent could be coarsened without breaking the split rule or 2:1
    n_cand, n_balance = treecheck.assert_minimal(nd, max_pts)
    assert n_cand > 0
    return n_balance


def test_minimality_check_catches_over_refinement():
    """The minimality check must reject a tree with one needless split: a leaf of a light
    region subdivided into 8 leaves, far from any finer leaf."""
    x = RNG.random((400, 3)) * 0.5  # points only in the low octant; the other 7 stay leaves
    tr = O.OracleTree(x, x, 64)
    nd = tr.nodes()
    treecheck.assert_minimal(nd, 64)
    # split the empty leaf at the far corner of the cube (its neighbours are at most one level
    # finer by the 2:1 balance, so nothing justifies the split)
    leaves = np.where(nd[:, 4] == 1)[0]
    far = max(leaves, key=lambda i: tuple(nd[i, [2, 1, 0]]))
    assert nd[far, 9] == 0
    d = int(nd[far, 3]) + 1
    s = 1 << (30 - d)
    kids = [[nd[far, 0] + s * (c & 1), nd[far, 1] + s * ((c >> 1) & 1), nd[far, 2] + s * ((c >> 2) & 1),
             d, 1, far, 0, 0, 0, 0] for c in range(8)]
    bad = nd.copy()
    bad[far, 4] = 0
    bad = np.concatenate([bad, np.array(kids, dtype=np.int64)])
    with pytest.raises(AssertionError, match="could be coarsened"):
        treecheck.assert_minimal(bad, 64)


@pytest.mark.parametrize("kind,max_pts", [("uniform", 16), ("sphere", 32), ("blob", 8)])
def test_tree_invariants(kind, max_pts):
    n = 3000
    if kind == "uniform":
        x = RNG.random((n, 3))
    elif kind == "sphere":
        d = RNG.standard_normal((n, 3))
        x = 0.5 + 0.4 * d / np.linalg.norm(d, axis=1)[:, None]
    else:
        x = np.clip(RNG.normal(0.3, 0.03, (n, 3)), 0, 1 - 1e-9)
    tr = O.OracleTree(x, x, max_pts)
    n_balance = _check_tree_invariants(tr, x, x, max_pts)
    print(f"{kind}: {n_balance} light internal nodes kept by the 2:1 balance")


def _brute_lists(nd):
    """Interaction lists from their geometric definitions on a small tree (closed boxes touch
    when their intervals overlap in all three dimensions)."""
    lo = nd[:, :3] * 2.0 ** -30
    s = 2.0 ** -nd[:, 3].astype(float)
    par = nd[:, 5]
    leaf = nd[:, 4] == 1
    N = len(nd)
    hi = lo + s[:, None]
    touch = np.all((lo[:, None, :] <= hi[None, :, :] + 1e-15) & (lo[None, :, :] <= hi[:, None, :] + 1e-15), axis=2)
    same_depth = nd[:, 3][:, None] == nd[:, 3][None, :]
    coll = touch & same_depth  # colleagues, self included
    out = {}
    for i in range(N):
        if nd[i, 9] == 0:
            continue
        U, W, X, V1 = set(), set(), set(), set()
        if leaf[i]:
            U = set(np.where(leaf & touch[i])[0])
            for c in np.where(coll[i] & ~leaf)[0]:
                W |= {ch for ch in np.where(par == c)[0] if not touch[i, ch]}
        if par[i] >= 0:
            X = {pc for pc in np.where(coll[par[i]] & leaf)[0] if not touch[i, pc]}
        if not leaf[i]:
            V1 = {c for c in np.where(coll[i] & ~leaf)[0] if c != i}
        out[i] = (U, W, X, V1)
    return out


def test_lists_match_geometric_definitions():
    x = np.concatenate([RNG.random((300, 3)), np.clip(RNG.normal(0.2, 0.02, (300, 3)), 0, 0.999)])
    tr = O.OracleTree(x, x, 16)
    nd = tr.nodes()
    assert len(nd) < 2000
    brute = _brute_lists(nd)
    L = {t: tr.lists(t) for t in (O.U0, O.U1, O.U2, O.W, O.X, O.V1)}
    for i, (U, W, X, V1) in brute.items():
        got_u = set(L[O.U0][i][L[O.U0][i] >= 0]) | set(L[O.U1][i][L[O.U1][i] >= 0]) | set(L[O.U2][i][L[O.U2][i] >= 0])
        assert got_u == U, (i, got_u, U)
        assert set(L[O.W][i][L[O.W][i] >= 0]) == W
        assert set(L[O.X][i][L[O.X][i] >= 0]) == X
        assert set(L[O.V1][i][L[O.V1][i] >= 0]) == V1


def test_list_tables_follow_reference_counts():
    tr = O.OracleTree(RNG.random((100, 3)), RNG.random((100, 3)), 16)
    counts = {O.U0: 56, O.U1: 27, O.U2: 56, O.V: 316, O.V1: 26, O.W: 152, O.X: 152, O.U2U: 8, O.D2D: 8}
    for t, c in counts.items():
        assert len(tr.list_rel(t)) == c


# ----------------------------------------------------------------------------- whole FMM
@pytest.mark.parametrize("kernel,tol", [("laplace_poten", 1e-5), ("laplace_grad", 3e-5), ("stokes_vel", 2e-4)])
def test_fmm_vs_direct(kernel, tol):
    n = 3000
    x = RNG.random((n, 3))
    k0, k1 = O.KDIM[O.KERNELS[kernel]]
    q = RNG.standard_normal(n * k0)
    u = O.fmm(kernel, 6, 64, x, q)
    ud = O.direct(kernel, x, q, x)
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < tol


def test_fmm_converges_in_m():
    n = 2000
    x = RNG.random((n, 3))
    q = RNG.standard_normal(n)
    ud = O.direct(0, x, q, x)
    errs = [np.linalg.norm(O.fmm(0, m, 32, x, q) - ud) / np.linalg.norm(ud) for m in (4, 6, 8)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-7


def test_fmm_interaction_completeness_separate_targets():
    """Sources and targets in different sets; the FMM must cover every source once."""
    src = RNG.random((1500, 3))
    trg = RNG.random((700, 3))
    q = RNG.standard_normal(1500)
    u = O.fmm(0, 8, 32, src, q, trg)
    ud = O.direct(0, src, q, trg)
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-6
