"""Checks against the committed fixtures of tests/golden/ (written by make_golden.py from
closed forms and the oracle).  CPU: the oracle and the product's host side reproduce them;
GPU: the HIP evaluator and the GPU direct sum reproduce them."""
import os

import numpy as np
import pytest

import oracle as O
import pvfmm_amd as P

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = {0: "laplace_poten", 1: "laplace_grad", 2: "stokes_vel"}


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def test_oracle_kernels_match_kat():
    k = load("kernel_kat.npz")
    for kid, name in NAMES.items():
        u = O.direct(kid, k["src"], k[f"{name}_den"], k["trg"])
        assert rel(u, k[f"{name}_val"]) < 1e-14, name


def test_oracle_tree_matches_fixture():
    t = load("tree_small.npz")
    tr = O.OracleTree(t["points"], t["points"], int(t["max_pts"]))
    assert np.array_equal(tr.nodes(), t["nodes"])
    for ty, nm in ((O.U0, "U0"), (O.U1, "U1"), (O.U2, "U2"), (O.V1, "V1"), (O.W, "W"), (O.X, "X")):
        assert np.array_equal(tr.lists(ty), t[f"list_{nm}"]), nm
    assert np.array_equal(tr.point_order(0), t["point_order_src"])


def test_product_tree_matches_fixture():
    """Same leaves, counts and lists as the fixture (the product orders nodes level-major, the
    oracle in preorder: rows are matched by (x, y, z, depth))."""
    t = load("tree_small.npz")
    pt = P.PtFMMTree(t["points"], t["points"], int(t["max_pts"]))
    on, pn = t["nodes"], pt.nodes()
    okey = [tuple(r[:4]) for r in on]
    pidx = {tuple(r[:4]): i for i, r in enumerate(pn)}
    assert sorted(okey) == sorted(pidx)
    perm = np.array([pidx[k] for k in okey])
    assert np.array_equal(on[:, 4], pn[perm, 4])
    assert np.array_equal(on[:, 6:10], pn[perm, 6:10])
    inv = perm.argsort()
    for ty, nm in ((O.U0, "U0"), (O.U1, "U1"), (O.U2, "U2"), (O.V1, "V1"), (O.W, "W"), (O.X, "X")):
        pl = pt.lists(ty)[perm]
        mapped = np.where(pl >= 0, inv[np.maximum(pl, 0)], -1)
        assert np.array_equal(mapped, t[f"list_{nm}"]), nm


def _fft_m2l(V, X, m):
    """Apply per-frequency 8x8 blocks V[f][t][s] (product convention, 1/n1^3 folded in) to
    the children's surface densities X[8][ns] through numpy FFTs."""
    n1, nh = 2 * m, m + 1
    lat = np.rint(O.surface(m, [0, 0, 0], m - 1, 0)).astype(int)
    g = np.zeros((8, n1, n1, n1))
    for c in range(8):
        g[c, lat[:, 2], lat[:, 1], lat[:, 0]] = X[c]
    Xh = np.fft.rfftn(g, axes=(1, 2, 3)).reshape(8, -1)
    Yh = np.einsum("fts,sf->tf", V, Xh)
    Y = np.fft.irfftn(Yh.reshape(8, n1, n1, nh), s=(n1, n1, n1), axes=(1, 2, 3)) * n1 ** 3
    return np.stack([Y[t][lat[:, 2], lat[:, 1], lat[:, 0]] for t in range(8)])


def test_product_m2l_operators_match_kat():
    k = load("m2l_kat.npz")
    m = int(k["m"])
    n1 = 2 * m
    Md = n1 * n1 * (m + 1)
    V = P.PtFMM("laplace_poten", m).operator("v1").reshape(27, Md, 8, 8, 2)
    V = V[..., 0] + 1j * V[..., 1]
    for R in (0, 10, 22):
        y = _fft_m2l(V[R], k[f"R{R}_X"], m)
        assert rel(y, k[f"R{R}_Y"]) < 1e-12, R


def test_oracle_energy_gradients_match_fd():
    f = load("fd_energy.npz")
    x, q = f["x"], f["q"]
    phi = O.direct(0, x, q, x)
    assert abs(0.5 * q @ phi - f["energy"]) < 1e-12 * abs(f["energy"])
    assert np.abs(phi - f["dE_dq"]).max() < 1e-8 * np.abs(phi).max()
    dx = -q[:, None] * O.direct(1, x, q, x).reshape(-1, 3)
    assert np.abs(dx - f["dE_dx"]).max() < 1e-5 * np.abs(dx).max()


def test_oracle_fmm_matches_fixture():
    f = load("fmm_small.npz")
    u = O.fmm("laplace_grad", int(f["m"]), int(f["max_pts"]), f["x"], f["q"])
    assert rel(u, f["u"]) < 1e-12


@pytest.mark.gpu
def test_gpu_direct_sum_matches_kat(gpu):
    k = load("kernel_kat.npz")
    for kid, name in NAMES.items():
        u = P.direct(name, k["src"], k[f"{name}_den"], k["trg"])
        assert rel(u, k[f"{name}_val"]) < 1e-13, name


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_gpu_fmm_matches_fixture(gpu, prec, tol):
    f = load("fmm_small.npz")
    t = P.PtFMMTree(f["x"], f["x"], int(f["max_pts"]))
    t.setup(P.PtFMM("laplace_grad", int(f["m"]), prec))
    e = rel(t.evaluate(f["q"]), f["u"])
    print(f"{prec}: rel-L2 vs oracle fixture {e:.2e}")
    assert e < tol
