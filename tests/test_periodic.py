"""Periodic boundaries (BoundaryType PX, PXY, PXYZ; the reference's FMM_Pts::PeriodicBC,
include/fmm_pts.txx:790-1163,2641-2657,5740-5774, and the periodic tree of
include/mpi_tree.txx:745-1255).

What pins what:
  * the oracle's periodic Laplace FMM against an Ewald sum (tests/ewald.py, tin-foil, uniform
    background for a net charge): potential up to its additive constant, and the field; the
    error falls geometrically with m;
  * the oracle's periodic Stokes FMM against Hasimoto's Ewald sum (itself checked for
    independence of its splitting parameter), and by translation invariance of the periodic
    velocity field up to a uniform flow;
  * the oracle's closed-form cube integrals against brute-force quadrature and the identities
    tr S = 2 L, symmetry and div u = 0;
  * the oracle's periodic 2:1 balance by a brute-force check across the wrap;
  * the product's periodic tree and lists bit-exact against the oracle, its boundary operator
    (own Duffy quadrature and Jacobi pseudo-inverses) against the oracle's, and on the GPU its
    evaluations against the oracle on the same tree and against the Ewald sum.
"""
import numpy as np
import pytest

import ewald
import oracle as O
import pvfmm_amd as P
from test_tree_parity import LIST_TYPES, _points

BTYPES = ["px", "pxy", "pxyz"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _rel_mean_free(a, b):
    """relative L2 after removing the mean difference per component (FMM potentials are
    defined up to a constant)"""
    a2, b2 = np.asarray(a).reshape(len(a), -1), np.asarray(b).reshape(len(b), -1)
    d = a2 - b2
    return float(np.linalg.norm(d - d.mean(0)) / np.linalg.norm(b2 - b2.mean(0)))


_OPS = {}


def _oracle_ops(kernel, m, bndry, real):
    """the oracle's operators and boundary operator, computed once per module run"""
    key = (kernel, m, bndry, real)
    if key not in _OPS:
        ops = O.precompute(kernel, m, real)
        ops["bc"] = O.bc_matrix(kernel, m, bndry, real)
        _OPS[key] = ops
    return _OPS[key]


def _oracle_eval(kernel, x, q, m, max_pts, bndry, trg=None, bc=None, real="fp64"):
    """The oracle's periodic FMM; bc: a boundary operator to use instead of the oracle's own
    (the GPU parity tests pass the product's, whose own parity is tested separately)."""
    trg = x if trg is None else trg
    t = O.OracleTree(x, trg, max_pts, bndry=bndry)
    if bc is None:
        ops = _oracle_ops(kernel, m, bndry, real)
    else:
        ops = dict(O.precompute(kernel, m, real))
        ops["bc"] = bc
    t.set_operators(ops)
    return t.evaluate(np.ravel(q))


# ----------------------------------------------------------------------------- oracle pins
def test_cube_integrals_closed_form():
    from numpy.polynomial.legendre import leggauss

    g, w = leggauss(40)
    g, w = (g + 1) / 2, w / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel()
    y = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    for kid in (0, 2):
        k0, k1 = O.KDIM[kid]
        for x in ([-1.3, 0.4, 2.2], [2.5, -0.9, 0.3], [0.5, 0.5, -1.5]):
            x = np.array(x)
            brute = np.einsum("s,sab->ab", W, O.kernel_matrix(kid, y, x.reshape(1, 3)).reshape(len(y), k0, k1))
            assert _rel(O.cube_integral(kid, x, 0.0, 1.0)[0], brute) < 1e-13
    # inside the cube: Laplacian of the Laplace integral is -1, Stokes trace = 2 Laplace,
    # symmetric, divergence free
    x, h = np.array([1.2, 1.7, 0.9]), 1e-3
    f = lambda p: O.cube_integral(0, p, 0.0, 3.0)[0, 0, 0]  # noqa: E731
    lap = sum((f(x + h * e) - 2 * f(x) + f(x - h * e)) / h ** 2 for e in np.eye(3))
    assert abs(lap + 1) < 1e-5
    G, L = O.cube_integral(2, x, 0.0, 3.0)[0], f(x)
    assert abs(np.trace(G) / (2 * L) - 1) < 1e-14 and np.abs(G - G.T).max() < 1e-14
    h = 1e-4
    u = lambda p: O.cube_integral(2, p, 0.0, 3.0)[0]  # noqa: E731
    div = [sum((u(x + h * e)[a, b] - u(x - h * e)[a, b]) / (2 * h) for b, e in enumerate(np.eye(3))) for a in range(3)]
    assert np.abs(div).max() < 1e-8


def test_oracle_periodic_laplace_matches_ewald():
    rng = np.random.default_rng(3)
    x = rng.random((600, 3))
    q = rng.standard_normal(600)
    errs = []
    for m in (4, 6, 8):
        for qq in (q - q.mean(), q):  # neutral, and with a net charge (uniform background)
            e = _rel_mean_free(_oracle_eval("laplace_poten", x, qq, m, 20, "pxyz"), ewald.laplace(x, qq, x))
            errs.append(e)
    # m = 4, 6, 8: geometric convergence to the Ewald sum
    assert errs[0] < 1e-3 and errs[2] < 3e-5 and errs[4] < 1e-6 and errs[5] < 1e-6, errs
    assert errs[2] < 0.05 * errs[0] and errs[4] < 0.05 * errs[2], errs
    qn = q - q.mean()
    g = _oracle_eval("laplace_grad", x, qn, 8, 20, "pxyz").reshape(-1, 3)
    assert _rel(g, ewald.laplace(x, qn, x, grad=True)) < 2e-6


def test_stokes_ewald_sum_is_split_independent():
    """The Hasimoto sum must not depend on its splitting parameter: a wrong real-space term or
    self term would change with xi."""
    rng = np.random.default_rng(1)
    x = rng.random((200, 3))
    f = rng.standard_normal((200, 3))
    f -= f.mean(0)
    u5, u6 = ewald.stokes(x, f, x, xi=5.0, kmax=14), ewald.stokes(x, f, x, xi=6.0, kmax=16)
    assert _rel_mean_free(u5, u6) < 1e-11


def test_oracle_periodic_stokes_matches_ewald():
    """The oracle's periodic Stokes FMM against Hasimoto's Ewald sum (measured 1.8e-3 at m = 6,
    1.8e-4 at m = 8, 1.6e-6 at m = 10: the periodic Stokes lattice operator converges slowly at
    low orders)."""
    rng = np.random.default_rng(1)
    x = rng.random((300, 3))
    f = rng.standard_normal((300, 3))
    f -= f.mean(0)
    u = _oracle_eval("stokes_vel", x, f, 8, 20, "pxyz").reshape(-1, 3)
    assert _rel_mean_free(u, ewald.stokes(x, f, x)) < 5e-4


def test_oracle_periodic_stokes_translation_invariant():
    """Moving every point by the same vector (mod 1) moves the periodic velocity field with it,
    up to a uniform flow.  Measured: 7e-3 (m = 4), 5e-3 (m = 6), 2e-4 (m = 8), 2e-6 (m = 10);
    the low orders are simply less accurate for periodic Stokes (m = 6 is 4e-3 from m = 10)."""
    rng = np.random.default_rng(4)
    x = rng.random((500, 3))
    f = rng.standard_normal((500, 3))
    f -= f.mean(0)
    u1 = _oracle_eval("stokes_vel", x, f, 8, 20, "pxyz").reshape(-1, 3)
    u2 = _oracle_eval("stokes_vel", (x + np.array([0.31, 0.17, 0.62])) % 1.0, f, 8, 20, "pxyz").reshape(-1, 3)
    assert _rel_mean_free(u2, u1) < 1e-3


def _direct_lattice_sum(x, q, axes, Ns):
    """symmetric square-truncated image sums |n|_inf <= N for the three N of Ns (N, 2N, 4N),
    extrapolated by Richardson in 1/N and 1/N^2 (the tail after the dipole terms cancel in
    +-n pairs)"""
    S = []
    for N in Ns:
        r = np.arange(-N, N + 1, dtype=np.float64)
        g = np.stack(np.meshgrid(*([r] * len(axes)), indexing="ij"), -1).reshape(-1, len(axes))
        sh = np.zeros((len(g), 3))
        sh[:, axes] = g
        d = x[:, None, None, :] - x[None, None, :, :] - sh[None, :, None, :]
        dist = np.sqrt((d * d).sum(-1))
        with np.errstate(divide="ignore"):
            S.append((np.where(dist > 0, 1.0 / dist, 0.0) * q).sum((1, 2)) / (4 * np.pi))
    return (8 * S[2] - 6 * S[1] + S[0]) / 3


def test_partly_periodic_lattice_sums():
    """The PX and PXY references: the wire sum's tail formula converges (two near-image counts),
    the slab Ewald sum does not depend on its splitting, and both equal the plain symmetric image
    sum (Richardson-extrapolated; its own floor is about 1e-7 for the slab)."""
    rng = np.random.default_rng(2)
    x = rng.random((16, 3))
    q = rng.standard_normal(16)
    q -= q.mean()
    w = ewald.laplace_wire(x, q, x)
    assert np.abs(w - ewald.laplace_wire(x, q, x, n_near=40)).max() < 1e-11 * np.abs(w).max()
    assert np.abs(w - _direct_lattice_sum(x, q, [0], (500, 1000, 2000))).max() < 1e-9 * np.abs(w).max()
    s = ewald.laplace_slab(x, q, x)
    assert np.abs(s - ewald.laplace_slab(x, q, x, alpha=4.0, kmax=10)).max() < 1e-11 * np.abs(s).max()
    assert np.abs(s - _direct_lattice_sum(x, q, [0, 1], (16, 32, 64))).max() < 2e-6 * np.abs(s).max()


LATTICE = {"px": ewald.laplace_wire, "pxy": ewald.laplace_slab}


@pytest.mark.parametrize("bndry", ["px", "pxy"])
def test_oracle_partly_periodic_matches_lattice_sums(bndry):
    """The oracle's PX and PXY potentials against the lattice sums of the neutral charges: equal
    up to a constant (the corner fit leaves the constant free; the linear terms along the periodic
    axes it removes vanish for these sums).  Measured at m = 6, 8, 10: PX 4.8e-6, 1.3e-7, 7.0e-9;
    PXY 6.3e-6, 2.1e-7, 1.0e-8."""
    rng = np.random.default_rng(5)
    x = rng.random((300, 3))
    q = rng.standard_normal(300)
    q -= q.mean()
    ref = LATTICE[bndry](x, q, x)
    e6 = _rel_mean_free(_oracle_eval("laplace_poten", x, q, 6, 20, bndry), ref)
    e8 = _rel_mean_free(_oracle_eval("laplace_poten", x, q, 8, 20, bndry), ref)
    assert e6 < 1.5e-5 and e8 < 5e-7 and e8 < 0.1 * e6, (e6, e8)


@pytest.mark.parametrize("bndry", BTYPES)
def test_oracle_periodic_balance_across_the_wrap(bndry):
    """Leaves that touch across a periodic face (every axis wraps in the balance) differ by at
    most one level; checked by brute force over leaf pairs."""
    rng = np.random.default_rng(11)
    x = _points("cubed", 1200, rng)
    nd = O.OracleTree(x, x, 8, bndry=bndry).nodes()
    lv = nd[nd[:, 4] == 1]
    scale = float(1 << 30)
    lo = lv[:, :3] / scale
    size = 0.5 ** lv[:, 3]
    for i in range(len(lv)):
        for sh in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
            a0, a1 = lo[i] + sh, lo[i] + sh + size[i]
            touch = np.all((lo <= a1[None, :] + 1e-15) & (lo + size[:, None] >= a0[None, :] - 1e-15), axis=1)
            assert np.all(np.abs(lv[touch, 3] - lv[i, 3]) <= 1)


# ----------------------------------------------------------------------------- product, CPU
@pytest.mark.parametrize("bndry", BTYPES)
@pytest.mark.parametrize("kind", ["uniform", "blob", "cubed"])
def test_periodic_tree_and_lists_bit_exact(bndry, kind):
    rng = np.random.default_rng(hash((bndry, kind)) % (1 << 32))
    x = _points(kind, 3000, rng)
    trg = x if kind != "blob" else _points("uniform", 1500, rng)
    max_pts = 20
    ot = O.OracleTree(x, trg, max_pts, bndry=bndry)
    pt = P.PtFMMTree(x, trg, max_pts, bndry=bndry)
    on, pn = ot.nodes(), pt.nodes()
    okey, pkey = [tuple(r[:4]) for r in on], [tuple(r[:4]) for r in pn]
    assert sorted(okey) == sorted(pkey)
    pidx = {k: i for i, k in enumerate(pkey)}
    perm = np.array([pidx[k] for k in okey])
    assert np.array_equal(on[:, 4], pn[perm, 4]) and np.array_equal(on[:, 6:10], pn[perm, 6:10])
    for t in LIST_TYPES:
        ol, pl = ot.lists(t), pt.lists(t)
        mapped = np.where(pl[perm] >= 0, perm.argsort()[np.maximum(pl[perm], 0)], -1)
        assert np.array_equal(ol, mapped), f"list type {t} differs"


@pytest.mark.parametrize("bndry", BTYPES)
def test_boundary_operator_matches_oracle(bndry):
    """Two implementations (Duffy quadrature and Jacobi pseudo-inverses here; closed forms and
    LAPACK in the oracle) agree to 1e-11 at m = 6 and to 3.5e-8 at m = 8, where singular
    values near the 4 eps max_S cutoff of the order-10 factors decide differently; the potentials
    then differ by 1e-8, thirty times below the m = 8 truncation error against the Ewald sum."""
    f = P.PtFMM("laplace_poten", 6)
    M = f.operator("bc_" + bndry).reshape(152, 152)
    assert _rel(M, O.bc_matrix("laplace_poten", 6, bndry)) < 1e-9
    if bndry == "pxyz":
        M8 = P.PtFMM("laplace_poten", 8).operator("bc_pxyz").reshape(296, 296)
        assert _rel(M8, O.bc_matrix("laplace_poten", 8, "pxyz")) < 2e-7


def test_boundary_operator_stokes_matches_oracle():
    f = P.PtFMM("stokes_vel", 4)
    n = 3 * (6 * 9 + 2)
    assert _rel(f.operator("bc_pxyz").reshape(n, n), O.bc_matrix("stokes_vel", 4, "pxyz")) < 1e-9


def test_periodic_rejects_bad_types():
    with pytest.raises(ValueError):
        P.PtFMMTree(np.random.default_rng(0).random((10, 3)), max_pts=4, bndry="pz")


# ----------------------------------------------------------------------------- product, GPU
@pytest.mark.gpu
@pytest.mark.parametrize("bndry", BTYPES)
@pytest.mark.parametrize("kind", ["uniform", "cubed", "dup"])
def test_periodic_device_tree_matches_host(gpu, monkeypatch, bndry, kind):
    """The GPU tree builder wraps the balance and the colleagues as the host builder does:
    nodes, point orders and every list identical, for each periodic boundary type."""
    rng = np.random.default_rng(hash((bndry, kind, 1)) % (1 << 32))
    x = _points(kind, 4000, rng)
    trg = _points("uniform", 1500, rng)
    monkeypatch.delenv("PVFMM_TREE", raising=False)
    td = P.PtFMMTree(x, trg, 16, bndry=bndry)
    monkeypatch.setenv("PVFMM_TREE", "host")
    th = P.PtFMMTree(x, trg, 16, bndry=bndry)
    assert np.array_equal(td.nodes(), th.nodes())
    for which in (0, 1):
        assert np.array_equal(td.point_order(which), th.point_order(which))
    for t in LIST_TYPES:
        assert np.array_equal(td.lists(t), th.lists(t)), f"list type {t} differs"


CASES = [  # kernel, precision, m, bndry, fp64 bar vs oracle, mixed bar
    ("laplace_poten", "fp64", 8, "pxyz"),
    ("laplace_poten", "mixed", 8, "pxyz"),
    ("laplace_grad", "fp64", 8, "pxyz"),
    ("laplace_grad", "mixed", 8, "pxyz"),
    ("stokes_vel", "fp64", 6, "pxyz"),
    ("stokes_vel", "mixed", 6, "pxyz"),
    ("laplace_poten", "fp64", 6, "px"),
    ("laplace_poten", "fp64", 6, "pxy"),
    ("stokes_vel", "fp64", 4, "pxy"),
    ("stokes_vel", "mixed", 6, "px"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,prec,m,bndry", CASES)
def test_periodic_gpu_matches_oracle(gpu, kernel, prec, m, bndry):
    rng = np.random.default_rng(7)
    n = 6000
    x = np.concatenate([rng.random((n // 2, 3)), np.clip(rng.normal(0.1, 0.05, (n // 2, 3)) % 1.0, 0, 1 - 1e-12)])
    k0 = 3 if kernel == "stokes_vel" else 1
    q = rng.standard_normal((n, k0))
    print(f"periodic {kernel} {prec} m={m} {bndry} n={n}", flush=True)
    t = P.PtFMMTree(x, x, 40, bndry=bndry)
    f = P.PtFMM(kernel, m, precision=prec)
    t.setup(f)
    u = t.evaluate(q.ravel())
    # the oracle on the same tree with the product's boundary operator, and its own operators
    # at the precision's pseudo-inverse cutoff
    real = "fp64" if prec == "fp64" else "fp32"
    uo = _oracle_eval(kernel, x, q, m, 40, bndry, bc=f.operator("bc_" + bndry).reshape(t_bc_shape(kernel, m)), real=real)
    e = _rel(u, uo)
    print(f"  rel-L2 vs oracle {e:.3e}", flush=True)
    # FP64: rounding-level agreement (Stokes FP64 agrees to 1e-11 in free space as well)
    assert e < ((1e-10 if kernel == "stokes_vel" else 1e-12) if prec == "fp64" else 2e-5)


def t_bc_shape(kernel, m):
    n = (3 if kernel == "stokes_vel" else 1) * (6 * (m - 1) ** 2 + 2)
    return (n, n)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["root-leaf", "separate-targets", "stokes-targets"])
def test_periodic_gpu_edge_cases(gpu, case):
    """A tree that is only its root (every U-list partner is the root or one of its 26
    images), and target sets apart from the sources, against the oracle."""
    rng = np.random.default_rng(12)
    kernel, m = ("stokes_vel", 6) if case.startswith("stokes") else ("laplace_poten", 8)
    if case == "root-leaf":
        x, trg, max_pts = rng.random((40, 3)), None, 64
    else:
        x, trg, max_pts = rng.random((3000, 3)), rng.random((1200, 3)), 32
    k0 = 3 if kernel == "stokes_vel" else 1
    q = rng.standard_normal((len(x), k0))
    t = P.PtFMMTree(x, trg, max_pts, bndry="pxyz")
    f = P.PtFMM(kernel, m)
    t.setup(f)
    u = t.evaluate(q.ravel())
    uo = _oracle_eval(kernel, x, q, m, max_pts, "pxyz", trg=trg, bc=f.operator("bc_pxyz").reshape(t_bc_shape(kernel, m)))
    e = _rel(u, uo)
    print(f"periodic {case}: rel-L2 vs oracle {e:.3e}", flush=True)
    assert e < (1e-10 if kernel == "stokes_vel" else 1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,prec", [("laplace_poten", "mixed"), ("stokes_vel", "fp64")])
def test_periodic_density_batches_and_graph(gpu, kernel, prec):
    """Periodic plans evaluate several density vectors in one pass (the boundary product and the
    volume-potential correction take every vector) equal to one at a time, and periodic trees
    replay through a recorded HIP graph (PtFMMBatch) as they evaluate eagerly."""
    import torch

    rng = np.random.default_rng(13)
    n = 4000
    x = rng.random((n, 3))
    k0 = 3 if kernel == "stokes_vel" else 1
    q = rng.standard_normal((10, k0 * n))
    fmm = P.PtFMM(kernel, 6, prec)
    tb = P.PtFMMTree(x, x, 48, bndry="pxyz").setup(fmm, n_vec_max=8)
    t1 = P.PtFMMTree(x, x, 48, bndry="pxyz").setup(fmm, n_vec_max=1)
    ub = tb.evaluate(q)
    for i in range(10):
        r = _rel(ub[i], t1.evaluate(q[i]))
        assert r < (1e-12 if prec == "fp64" else 1e-6), (i, r)
    x2 = rng.random((2500, 3))
    t2 = P.PtFMMTree(x2, x2, 48, bndry="pxy").setup(fmm)
    dens = [torch.from_numpy(q[0]).to("cuda:0"), torch.from_numpy(rng.standard_normal(k0 * 2500)).to("cuda:0")]
    eager = [t1.evaluate_device(dens[0]).clone(), t2.evaluate_device(dens[1]).clone()]
    batch = P.PtFMMBatch([t1, t2])
    outs = [torch.empty_like(e) for e in eager]
    for rep in range(3):
        batch.evaluate_device(dens, outs)
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), rep
    assert batch.records() == 1


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_periodic_gpu_matches_ewald(gpu, prec):
    """End to end against the Ewald sum at m = 10: potential (net charge, so the volume-potential
    corrections act) up to its constant, and the field of a neutral set."""
    rng = np.random.default_rng(8)
    n = 2000
    x = rng.random((n, 3))
    q = rng.standard_normal(n)
    t = P.PtFMMTree(x, x, 32, bndry="pxyz")
    t.setup(P.PtFMM("laplace_poten", 10, precision=prec))
    e = _rel_mean_free(t.evaluate(q), ewald.laplace(x, q, x))
    qn = q - q.mean()
    tg = P.PtFMMTree(x, x, 32, bndry="pxyz")
    tg.setup(P.PtFMM("laplace_grad", 10, precision=prec))
    eg = _rel(tg.evaluate(qn).reshape(-1, 3), ewald.laplace(x, qn, x, grad=True))
    print(f"periodic {prec} vs Ewald: potential {e:.3e}, field {eg:.3e}", flush=True)
    # mixed: FP32 rounding measured against the potential's fluctuation (its mean removed)
    assert e < (1e-8 if prec == "fp64" else 1e-5) and eg < (1e-7 if prec == "fp64" else 5e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("bndry", ["px", "pxy"])
def test_partly_periodic_gpu_matches_lattice_sums(gpu, bndry, prec):
    """PX and PXY on the GPU at m = 10 against the wire and slab lattice sums, up to a constant
    (the oracle, on the CPU, reaches 1e-8 at this order)."""
    rng = np.random.default_rng(14)
    n = 1000
    x = rng.random((n, 3))
    q = rng.standard_normal(n)
    q -= q.mean()
    if bndry not in _LATTICE_REF:
        _LATTICE_REF[bndry] = LATTICE[bndry](x, q, x)
    t = P.PtFMMTree(x, x, 32, bndry=bndry)
    t.setup(P.PtFMM("laplace_poten", 10, precision=prec))
    e = _rel_mean_free(t.evaluate(q), _LATTICE_REF[bndry])
    print(f"{bndry} {prec} vs lattice sum: {e:.3e}", flush=True)
    # measured with 1500 points: FP64 4.9e-9 (PX), 4.6e-9 (PXY); mixed 4.1e-7, 6.9e-7
    assert e < (2e-8 if prec == "fp64" else 3e-6)


_LATTICE_REF = {}


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_periodic_stokes_gpu_matches_ewald(gpu, prec):
    """Periodic Stokes on the GPU against Hasimoto's Ewald sum at m = 10 (velocities up to a
    uniform flow), forces of zero sum."""
    rng = np.random.default_rng(10)
    n = 1500
    x = rng.random((n, 3))
    f = rng.standard_normal((n, 3))
    f -= f.mean(0)
    t = P.PtFMMTree(x, x, 32, bndry="pxyz")
    t.setup(P.PtFMM("stokes_vel", 10, precision=prec))
    e = _rel_mean_free(t.evaluate(f.ravel()).reshape(-1, 3), ewald.stokes(x, f, x))
    print(f"periodic stokes {prec} vs Ewald: {e:.3e}", flush=True)
    assert e < (1e-5 if prec == "fp64" else 1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,kid,out_exp", [("laplace_poten", 0, 1.0), ("laplace_grad", 1, 2.0), ("stokes_vel", 3, 1.0)])
def test_periodic_context_c_abi(gpu, kernel, kid, out_exp):
    """PVFMMCreateContextD with box_size > 0: coordinates in [0, box_size) (and outside it,
    wrapped by whole boxes), results equal the oracle's periodic FMM of the scaled points with
    the outputs scaled by box_size^-e, e the kernel's output exponent (src/pvfmm-wrapper.cpp:
    522-541: 1 for the potentials and velocities, 2 for the gradient)."""
    import ctypes

    from pvfmm_amd import ffi

    rng = np.random.default_rng(9)
    n, box = 5000, 2.0
    k0, k1 = (3, 3) if kernel == "stokes_vel" else (1, 3 if kernel == "laplace_grad" else 1)
    x = rng.random((n, 3)) * box
    x[:50] += box  # one box beyond: wrapped into [0, box)
    q = rng.standard_normal(k0 * n)
    L = ffi.lib()
    ctx = L.PVFMMCreateContextD(box, 64, 8, kid, 0)
    ffi.check()
    out = np.zeros(k1 * n)
    xf = x.ravel().copy()
    L.PVFMMEvalD(ffi.dptr(xf), ffi.dptr(q), None, n, ffi.dptr(xf), ffi.dptr(out), n, ctx, 1)
    ffi.check()
    p = ctypes.c_void_p(ctx)
    L.PVFMMDestroyContextD(ctypes.byref(p))
    xs = (x / box) % 1.0
    ot = O.OracleTree(xs, xs, 64, pt=_context_grid(64), bndry="pxyz")
    ops = O.precompute(kernel, 8)
    ops["bc"] = P.PtFMM(kernel, 8).operator("bc_pxyz").reshape(t_bc_shape(kernel, 8))  # the context's operators
    ot.set_operators(ops)
    uo = ot.evaluate(q) * box ** -out_exp
    # Stokes at m = 8: the product's Jacobi factors and the oracle's LAPACK ones differ at the
    # FP64 cutoff (free-space Stokes at m = 8 agrees to 3e-9 as well); measured 7.2e-9
    assert _rel(out, uo) < (5e-8 if kernel == "stokes_vel" else 1e-12)


def _context_grid(n):
    """The initial points of a context tree (src/pvfmm-wrapper.cpp:416-432)."""
    NN = int(np.ceil(n ** (1.0 / 3.0)))
    i = np.arange(NN ** 3)
    return np.stack([((i % NN) + 0.5) / NN, (((i // NN) % NN) + 0.5) / NN, ((i // (NN * NN)) + 0.5) / NN], 1)
