"""ORACLE -- TEST INFRASTRUCTURE ONLY.

FP64 CPU restatement of the reference (dmalhotra/pvfmm) particle FMM.  It is the parity
checker for the HIP evaluator in ``pvfmm_amd`` and the CPU baseline timed by ``bench.py``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import it;
the product never does.

Split:
  * ``fmm_oracle.cpp`` (built into ``liboracle.so``): tree, interaction lists and the
    list-by-list evaluation, restated from the reference sources (file:line in its header).
  * this module: the operator precompute of ``FMM_Pts::Precomp``
    (include/fmm_pts.txx:453-764) written with numpy/LAPACK, so that the product's own
    Jacobi-SVD pseudo-inverses and DFTs are checked against an independent implementation.

What pins each part (tests/test_oracle.py unless noted):
  * point kernels: closed forms, a finite-difference gradient, translation invariance;
  * DFT: numpy FFT; M2L via FFT == dense equivalent->check interaction (exact identity);
  * pseudo-inverses: the C2E*E2C projector property;
  * tree and lists: brute-force geometric list definitions, 2:1 balance and tiling
    invariants, reference table counts (include/interac_list.txx:39-56);
  * whole FMM: relative L2 against an FP64 direct sum, convergence in m;
  * periodic boundary (bc_matrix, cube_integral, the periodic tree): Ewald sums for Laplace
    (tests/ewald.py; potential, field, net charge), geometric convergence in m, translation
    invariance for Stokes, brute-force quadrature and identities for the closed-form cube
    integrals, a brute-force 2:1 check across the wrap (tests/test_periodic.py).
Bitwise agreement with the reference's floating point (sctl::approx_rsqrt, LAPACK SVD signs,
OpenMP summation order) is not pinned: SCTL is not vendored in the reference checkout, so no
reference build exists to compare with.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

RAD0 = 1.05  # include/pvfmm_common.hpp:33
RAD1 = 2.95  # include/pvfmm_common.hpp:34

# Mat_Type values, include/precomp_mat.hpp:21-39
UC2UE0, UC2UE1, DC2DE0, DC2DE1, S2U, U2U, D2D, D2T, U0, U1, U2, V, W, X, V1, BC = range(16)

KERNELS = {"laplace_poten": 0, "laplace_grad": 1, "stokes_vel": 2}
KDIM = {0: (1, 1), 1: (1, 3), 2: (3, 3)}


def build(force: bool = False) -> str:
    """Compile liboracle.so from fmm_oracle.cpp (g++, OpenMP)."""
    src = os.path.join(HERE, "fmm_oracle.cpp")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER
        L.oracle_tree_create.restype = vp
        L.oracle_tree_create.argtypes = [P(dp), i64, P(dp), i64, P(dp), i64, i32, i32, i32]
        L.oracle_set_bc.argtypes = [vp, P(dp)]
        L.oracle_tree_destroy.argtypes = [vp]
        L.oracle_tree_num_nodes.restype = i64
        L.oracle_tree_num_nodes.argtypes = [vp]
        L.oracle_tree_nodes.argtypes = [vp, P(i64)]
        L.oracle_tree_point_order.argtypes = [vp, i32, P(i64)]
        L.oracle_list_count.restype = i32
        L.oracle_list_count.argtypes = [vp, i32]
        L.oracle_list_rel.argtypes = [vp, i32, P(i32)]
        L.oracle_tree_lists.argtypes = [vp, i32, P(ctypes.c_int32)]
        L.oracle_set_operators.argtypes = [vp, i32, i32] + [P(dp)] * 7
        L.oracle_evaluate.argtypes = [vp, P(dp), P(dp), i32, P(dp)]
        L.oracle_direct.argtypes = [i32, P(dp), i64, P(dp), P(dp), i64, P(dp), i32]
        L.oracle_surface.argtypes = [i32, P(dp), dp, i32, P(dp)]
        L.oracle_dft_r2c.argtypes = [i32, P(dp), P(dp)]
        L.oracle_dft_c2r.argtypes = [i32, P(dp), P(dp)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def nthreads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ----------------------------------------------------------------------------- geometry
def surface(p: int, c, alpha: float, depth: int) -> np.ndarray:
    """surface() of include/fmm_pts.txx:54-97 -> (n_surf, 3)."""
    n = 6 * (p - 1) ** 2 + 2
    out = np.zeros(3 * n)
    c = _c(c)
    lib().oracle_surface(p, _dp(c), float(alpha), depth, _dp(out))
    return out.reshape(n, 3)


def surf_at(m, corner, depth, rad):
    """u/d check/equiv surfaces anchored at the box corner, include/fmm_pts.txx:99-141."""
    r = 0.5 * 0.5 ** depth
    c = np.asarray(corner, dtype=np.float64) - r * (rad - 1.0)
    return surface(m, c, rad, depth)


# ----------------------------------------------------------------------------- kernels
def kernel_matrix(kid: int, src: np.ndarray, trg: np.ndarray) -> np.ndarray:
    """Kernel::BuildMatrix (include/kernel.txx:1059-1135): rows src*k0+i, cols trg*k1+j."""
    src = np.asarray(src, dtype=np.float64).reshape(-1, 3)
    trg = np.asarray(trg, dtype=np.float64).reshape(-1, 3)
    k0, k1 = KDIM[kid]
    r = trg[None, :, :] - src[:, None, :]  # [s, t, 3], r = x_trg - x_src
    r2 = np.einsum("stk,stk->st", r, r)
    with np.errstate(divide="ignore"):
        rinv = np.where(r2 > 0, 1.0 / np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0)
    ns, nt = r2.shape
    if kid == 0:
        M = rinv / (4 * np.pi)
        return M
    if kid == 1:
        M = (r * (rinv ** 3)[..., None]) / (4 * np.pi)  # [s, t, 3]
        return M.reshape(ns, nt * 3)
    # stokes_vel: u_i = (f_i + r_i (f.r)/r^2)/r / (8 pi)
    M = np.zeros((ns, 3, nt, 3))
    for i in range(3):
        for j in range(3):
            M[:, i, :, j] = (rinv * ((i == j) + r[..., i] * r[..., j] * rinv * rinv)) / (8 * np.pi)
    return M.reshape(ns * 3, nt * 3)


def expansion_kernel(kid: int) -> int:
    """k_s2m/k_m2m/k_m2l/k_l2l: laplace_grad uses the potential kernel (include/kernel.txx:1338-1343)."""
    return 0 if kid == 1 else kid


# ----------------------------------------------------------------------------- operators
def _c2e(kid, equiv, check, eps):
    """SVD pseudo-inverse factors with the 4*eps*max_S cutoff (include/fmm_pts.txx:478-581),
    eps = machine epsilon of the evaluation type Real_t (include/fmm_pts.txx:503,555)."""
    M = kernel_matrix(kid, equiv, check)
    U, S, Vh = np.linalg.svd(M)
    Sinv = np.where(S > eps * S.max() * 4, 1.0 / S, 0.0)
    return (Vh.T * Sinv[None, :]), U.T.copy()


def precompute(kernel: str | int, m: int, real: str = "fp64") -> dict:
    """Level-0 operators of FMM_Pts::Precomp (include/fmm_pts.txx:453-764).

    real: the reference's Real_t whose machine epsilon sets the pseudo-inverse cutoff
    ("fp64", or "fp32" for the reference's single-precision build); the arithmetic is FP64
    either way.

    Returns dict with uc2ue0/1, dc2de0/1, u2u[8], d2d[8] (by child position bits x|y<<1|z<<2)
    and v1[27] (by parent offset (dx+1)+3(dy+1)+9(dz+1), 13 unused), V1 layout
    [(j + i*k1)*M_dim + f][(s*8 + t)*2 + re/im] with the 1/n3 factor folded in.
    """
    kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    ke = expansion_kernel(kid)
    k0, k1 = KDIM[ke]
    z = [0.0, 0.0, 0.0]
    uc = surf_at(m, z, 0, RAD1)
    ue = surf_at(m, z, 0, RAD0)
    dc = surf_at(m, z, 0, RAD0)
    de = surf_at(m, z, 0, RAD1)
    eps = float(np.finfo({"fp64": np.float64, "fp32": np.float32}[real]).eps)
    uc2ue0, uc2ue1 = _c2e(ke, ue, uc, eps)
    dc2de0, dc2de1 = _c2e(ke, de, dc, eps)
    u2u, d2d = [], []
    for c in range(8):
        bits = np.array([(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1], dtype=np.float64)
        ce = surf_at(m, bits * 0.5, 1, RAD0)  # child upward equivalent surface (depth 1)
        u2u.append((kernel_matrix(ke, ce, uc) @ uc2ue0) @ uc2ue1)
        cc = surf_at(m, bits * 0.5, 1, RAD0)  # child downward check surface (depth 1)
        d2d.append(dc2de0 @ (dc2de1 @ kernel_matrix(ke, de, cc)))
    # V: kernel on the (2m)^3 convolution grid, r2c FFT (include/fmm_pts.txx:671-718)
    n1 = 2 * m
    n3 = n1 ** 3
    nh = n1 // 2 + 1
    M_dim = n1 * n1 * nh
    a = RAD0  # r = 0.5^0
    g = (np.arange(n1) - m) * a / (m - 1)
    gk, gj, gi = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([gi.ravel(), gj.ravel(), gk.ravel()], axis=1)  # index i + n1 j + n2 k
    spec_cache = {}

    def vspec(rel):
        key = tuple(int(v) for v in rel)
        if key not in spec_cache:
            grid = base + np.asarray(rel, dtype=np.float64)[None, :]
            conv = kernel_matrix(ke, grid, np.zeros((1, 3)))  # [n3*k0][k1]
            conv = conv.reshape(n3, k0 * k1).T  # [k0k1][n3]
            sp = np.fft.rfftn(conv.reshape(k0 * k1, n1, n1, n1), axes=(1, 2, 3))
            spec_cache[key] = sp.reshape(k0 * k1, M_dim)
        return spec_cache[key]

    v1 = np.zeros((27, k0 * k1 * M_dim, 128))
    for r in range(27):
        if r == 13:
            continue
        R = np.array([r % 3 - 1, (r // 3) % 3 - 1, (r // 9) % 3 - 1])
        for t in range(8):
            for s in range(8):
                bt = np.array([t & 1, (t >> 1) & 1, (t >> 2) & 1])
                bs = np.array([s & 1, (s >> 1) & 1, (s >> 2) & 1])
                rel = 2 * R - bt + bs
                if np.abs(rel).max() < 2:
                    continue  # adjacent children: zero block (include/fmm_pts.txx:737-753)
                sp = vspec(rel) / n3
                v1[r, :, 2 * (s * 8 + t)] = sp.real.ravel()
                v1[r, :, 2 * (s * 8 + t) + 1] = sp.imag.ravel()
    return dict(kernel=kid, m=m, uc2ue0=uc2ue0, uc2ue1=uc2ue1, dc2de0=dc2de0, dc2de1=dc2de1,
                u2u=np.stack(u2u), d2d=np.stack(d2d), v1=v1, M_dim=M_dim)


# ----------------------------------------------------------------------------- periodic boundary
BC_LEVELS = 45  # PVFMM_BC_LEVELS, include/pvfmm_common.hpp:31
BNDRY = {"free": 0, "px": 1, "pxy": 2, "pxyz": 3}  # BoundaryType, include/mpi_tree.hpp:22-29


def u_equiv_surf(m, c, depth):
    return surf_at(m, c, depth, RAD0)


def u_check_surf(m, c, depth):
    return surf_at(m, c, depth, RAD1)


def d_check_surf(m, c, depth):
    return surf_at(m, c, depth, RAD0)


def d_equiv_surf(m, c, depth):
    return surf_at(m, c, depth, RAD1)


def _prism_inv_r(rel_lo, rel_hi):
    """Closed form of the integral of 1/|y - x| over a box, corners relative to x: the
    Newtonian potential of a rectangular prism (sum over the 8 corners with signs of
    X Y asinh(Z/rho_xy) + ... - X^2/2 atan(Y Z / (X R)) - ...).  rel_lo, rel_hi: (n, 3)."""
    out = np.zeros(len(rel_lo))
    for sx in (0, 1):
        for sy in (0, 1):
            for sz in (0, 1):
                X = np.where(sx, rel_hi[:, 0], rel_lo[:, 0])
                Y = np.where(sy, rel_hi[:, 1], rel_lo[:, 1])
                Z = np.where(sz, rel_hi[:, 2], rel_lo[:, 2])
                sign = (1 if sx else -1) * (1 if sy else -1) * (1 if sz else -1)
                R = np.sqrt(X * X + Y * Y + Z * Z)
                with np.errstate(divide="ignore", invalid="ignore"):
                    def ash(a, b, c):  # asinh(a / hypot(b, c)), 0 when b = c = 0
                        h = np.hypot(b, c)
                        return np.where(h > 0, np.arcsinh(a / np.where(h > 0, h, 1.0)), 0.0)

                    def at(a, b, c):  # a^2 / 2 * atan(b c / (a R)), 0 when a = 0
                        return np.where(a != 0, 0.5 * a * a * np.arctan(b * c / np.where(a != 0, a * R, 1.0)), 0.0)
                    F = X * Y * ash(Z, X, Y) + Y * Z * ash(X, Y, Z) + Z * X * ash(Y, Z, X) - at(X, Y, Z) - at(Y, Z, X) - at(Z, X, Y)
                out += sign * F
    return out


def _face_r_over_r(i, j, rel_lo, rel_hi):
    """sum over the two faces normal to axis j of n_j * integral over the face of
    (y_i - x_i) / |y - x| dA, corners relative to x.  Closed forms: in-plane component
    int int w / R dw dz = sum +- (z R + a^2 asinh(z / a)) / 2 with a^2 = w^2 + d^2; normal component
    d * int int 1 / R = d * sum +- (u asinh(v / hypot(u, d)) + v asinh(u / hypot(v, d)) - d atan(u v / (d R)))."""
    p, q = [a for a in range(3) if a != j]
    out = np.zeros(len(rel_lo))
    with np.errstate(divide="ignore", invalid="ignore"):
        for sj in (0, 1):
            d = rel_hi[:, j] if sj else rel_lo[:, j]
            nrm = 1.0 if sj else -1.0
            acc = np.zeros(len(rel_lo))
            for sp in (0, 1):
                for sq in (0, 1):
                    u = rel_hi[:, p] if sp else rel_lo[:, p]
                    v = rel_hi[:, q] if sq else rel_lo[:, q]
                    sign = (1 if sp else -1) * (1 if sq else -1)
                    R = np.sqrt(u * u + v * v + d * d)
                    if i == j:
                        hu, hv = np.hypot(u, d), np.hypot(v, d)
                        t1 = np.where(hu > 0, u * np.arcsinh(v / np.where(hu > 0, hu, 1.0)), 0.0)
                        t2 = np.where(hv > 0, v * np.arcsinh(u / np.where(hv > 0, hv, 1.0)), 0.0)
                        t3 = np.where(d != 0, d * np.arctan(u * v / np.where(d != 0, d * R, 1.0)), 0.0)
                        acc += sign * d * (t1 + t2 - t3)
                    else:
                        w, z = (u, v) if i == p else (v, u)  # w: the in-plane axis i, z: the other
                        a = np.hypot(w, d)  # int dw w / R = R, then int dz R with a^2 = w^2 + d^2
                        h = 0.5 * (z * R + np.where(a > 0, a * a * np.arcsinh(z / np.where(a > 0, a, 1.0)), 0.0))
                        acc += sign * h
            out += nrm * acc
    return out


def cube_integral(kernel, pts, lo, size) -> np.ndarray:
    """(n, k0, k1): integral over the cube [lo, lo + size]^3 of the expansion kernel from a
    uniform unit density (component a) to each point (component b), in closed form (the
    reference's cheb_integ of Chebyshev degree 0, include/cheb_utils.txx:1037-1081, as used by
    include/fmm_pts.txx:865-880,1009-1020).  Stokes: int S_ab = (2 delta_ab P - faces_b(a)) / (8 pi)
    from r_a r_b / r^3 = delta_ab / r - d_b((y_a - x_a) / r) and the divergence theorem."""
    kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    rel_lo = lo - pts
    rel_hi = lo + size - pts
    P = _prism_inv_r(rel_lo, rel_hi)
    if kid in (0, 1):
        return (P / (4 * np.pi)).reshape(-1, 1, 1)
    G = np.zeros((len(pts), 3, 3))
    for a in range(3):
        for b in range(3):
            G[:, a, b] = ((2.0 * P if a == b else 0.0) - _face_r_over_r(a, b, rel_lo, rel_hi)) / (8 * np.pi)
    return G


def vol_poten(kid, pts) -> np.ndarray:
    """(k0, n*k1): the kernels' analytic potential of a uniform unit density
    (laplace_vol_poten / stokes_vol_poten, include/kernel.txx:1289-1295,1456-1467)."""
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    r2 = (pts * pts).sum(1)
    if kid == 0:
        return (-r2 / 6).reshape(1, n)
    out = np.zeros((3, n, 3))
    for k in range(3):
        out[k, :, k] = -(r2 - pts[:, k] ** 2) / 4
    return out.reshape(3, n * 3)


def bc_matrix(kernel, m: int, bndry="pxyz", real: str = "fp64", upsample: int = 2) -> np.ndarray:
    """Periodic boundary operator of FMM_Pts::Precomp(BC_Type) (include/fmm_pts.txx:790-1163,
    the default build: no PVFMM_EXTENDED_BC): root upward equivalent density -> root downward
    check potential of every periodic image beyond the root's colleagues, (n_surf*k0, n_surf*k1)
    in the expansion kernel at order m.  Computed at order m + upsample and downsampled."""
    kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    bt = BNDRY[bndry] if isinstance(bndry, str) else int(bndry)
    ke = expansion_kernel(kid)
    k0, k1 = KDIM[ke]
    eps = float(np.finfo({"fp64": np.float64, "fp32": np.float32}[real]).eps)
    mo = m + upsample
    n = 6 * (mo - 1) ** 2 + 2
    K = lambda a, b: kernel_matrix(ke, a, b)  # noqa: E731

    def c2e(level, up):
        c = np.full(3, -(2.0 ** level) / 2)
        if up:
            return _c2e(ke, u_equiv_surf(mo, c, level), u_check_surf(mo, c, level), eps)
        return _c2e(ke, d_equiv_surf(mo, c, level), d_check_surf(mo, c, level), eps)

    uc2ue0, uc2ue1 = c2e(0, True)
    uc2ue0_, uc2ue1_ = c2e(-1, True)
    dc2de0, dc2de1 = c2e(0, False)
    dc2de0_, dc2de1_ = c2e(-1, False)
    has_vol = bt == 3 and kid != 1  # laplace_grad's M2L kernel carries no volume potential
    # zero-average projections (include/fmm_pts.txx:862-913)
    G = cube_integral(ke, u_check_surf(mo, np.zeros(3), 0), 0.0, 1.0)  # (n, a, b)
    M_s2c = np.transpose(G, (2, 0, 1)).reshape(k0, n * k1)  # [j][i*k1 + k] = int K_{k,j}(uc_i)
    M_s2e = (M_s2c @ uc2ue0) @ uc2ue1
    M_s2e = M_s2e / M_s2e.sum(1, keepdims=True)
    Z = np.eye(n * k0) - np.tile(M_s2e, (n, 1))
    C = np.eye(n * k1)
    for k in range(k1):
        C[k::k1, k::k1] -= 1.0 / n
    keep = lambda a, b, c: not ((bt == 1 and (b != 0 or c != 0)) or (bt == 2 and c != 0))  # noqa: E731
    # M2M: 8 unit boxes -> their parent of size 2 (include/fmm_pts.txx:917-941)
    tb = np.full(3, -1.0)
    chk = u_check_surf(mo, tb, -1)
    M2M = np.zeros((n * k0, n * k1))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                if keep(a, b, c):
                    M2M += K(u_equiv_surf(mo, tb + np.array([a, b, c]), 0), chk)
    M2M = (M2M @ uc2ue0_) @ uc2ue1_
    # L2L: parent of size 2 -> its lower-corner child (include/fmm_pts.txx:942-951)
    L2L = dc2de0_ @ (dc2de1_ @ K(d_equiv_surf(mo, tb, -1), d_check_surf(mo, tb, 0)))
    # M2L: unit boxes at offsets in [-2, 3]^3 not adjacent to the root (include/fmm_pts.txx:952-975)
    dnc = d_check_surf(mo, np.zeros(3), 0)
    M2L = np.zeros((n * k0, n * k1))
    for a in range(-2, 4):
        for b in range(-2, 4):
            for c in range(-2, 4):
                if not keep(a, b, c) or max(abs(a), abs(b), abs(c)) <= 1:
                    continue
                M2L += K(u_equiv_surf(mo, np.array([a, b, c], dtype=np.float64), 0), dnc)
    M2M = Z @ M2M @ Z
    L2L = C @ L2L @ C
    M2L = Z @ M2L @ C
    # one level up: the same map for a box of size 2 is 2^-(src_scal + trg_scal) = 1/2 of it
    # (Pr, Pc of include/fmm_pts.txx:980-993 for the potential-type expansion kernels)
    M = M2L.copy()
    for _ in range(1, BC_LEVELS):
        M = M2L + M2M @ (0.5 * M) @ L2L
    if has_vol:  # far field of the analytic volume potential (include/fmm_pts.txx:1001-1033)
        dc = d_check_surf(mo, np.ones(3), 0)
        G3 = cube_integral(ke, dc, 0.0, 3.0)
        M_near = np.transpose(G3, (2, 0, 1)).reshape(k0, n * k1)
        M_far = vol_poten(ke, dc) - M_near
        M += np.tile(M_far, (n, 1))
    # corner correction (include/fmm_pts.txx:1036-1138)
    enable = {3: [1, 1, 1, 1, 0, 0, 0, 0], 2: [1, 1, 1, 0, 0, 0, 0, 0], 1: [1, 1, 0, 0, 0, 0, 0, 0]}[bt]
    corners = np.array([[i % 2, (i // 2) % 2, (i // 4) % 2] for i in range(8)], dtype=np.float64)
    z3 = np.zeros(3)
    ue, de = u_equiv_surf(mo, z3, 0), d_equiv_surf(mo, z3, 0)
    cv = (M @ dc2de0) @ (dc2de1 @ K(de, corners))  # (n*k0, 8*k1)
    for kk in range(8):
        for j0 in (-1, 0, 1):
            for j1 in (-1, 0, 1):
                for j2 in (-1, 0, 1):
                    if (bt == 2 and j2 != 0) or (bt == 1 and (j1 != 0 or j2 != 0)):
                        continue
                    pt = corners[kk] - np.array([j0, j1, j2])
                    if np.all(np.abs(pt - 0.5) < 1):
                        continue
                    cv[:, kk * k1:(kk + 1) * k1] += K(ue, pt.reshape(1, 3))
    if has_vol:
        cv -= np.tile(vol_poten(ke, corners), (n, 1))
    basis = lambda p: np.array([np.ones(len(p)), p[:, 0], p[:, 1], p[:, 2], p[:, 0] * p[:, 1],  # noqa: E731
                                p[:, 0] * p[:, 2], p[:, 1] * p[:, 2], p[:, 0] * p[:, 1] * p[:, 2]])
    V = basis(corners)  # (coeff, corner)
    Vp = np.linalg.pinv(V)  # (corner, coeff)
    cvec = basis(dnc) * np.array(enable, dtype=np.float64)[:, None]  # (coeff, n)
    coeff = np.einsum("ikj,kc->icj", cv.reshape(n * k0, 8, k1), Vp)
    M = M - np.einsum("icj,cs->isj", coeff, cvec).reshape(n * k0, n * k1)
    if upsample:  # back to order m (include/fmm_pts.txx:1141-1160)
        rc = np.full(3, -0.5)
        M2Md = (K(u_equiv_surf(m, rc, 0), u_check_surf(mo, rc, 0)) @ uc2ue0) @ uc2ue1
        L2Ld = dc2de0 @ (dc2de1 @ K(d_equiv_surf(mo, rc, 0), d_check_surf(m, rc, 0)))
        M = M2Md @ M @ L2Ld
    return M


# ----------------------------------------------------------------------------- tree + FMM
class OracleTree:
    """PtFMM_CreateTree + FMM_Tree::SetupFMM restated (include/pvfmm.txx:215-245).

    bndry: "free" or a periodic BoundaryType ("px", "pxy", "pxyz"); periodic trees need the
    operators' "bc" entry (bc_matrix) for the lattice beyond the root's colleagues."""

    def __init__(self, src, trg, max_pts, pt=None, refine=True, bndry="free"):
        L = lib()
        self.src = _c(src).reshape(-1, 3)
        self.trg = _c(trg).reshape(-1, 3)
        self.pt = self.trg if pt is None else _c(pt).reshape(-1, 3)
        self.bndry = BNDRY[bndry] if isinstance(bndry, str) else int(bndry)
        self.h = L.oracle_tree_create(_dp(self.src), len(self.src), _dp(self.trg), len(self.trg),
                                      _dp(self.pt), len(self.pt), int(max_pts), int(refine), self.bndry)
        self.n_nodes = L.oracle_tree_num_nodes(self.h)
        self.ops = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().oracle_tree_destroy(self.h)
                self.h = None
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def nodes(self) -> np.ndarray:
        """(n_nodes, 10): x, y, z, depth, is_leaf, parent, n_src, n_trg, pt_cnt0, pt_cnt1 (preorder)."""
        out = np.zeros((self.n_nodes, 10), dtype=np.int64)
        lib().oracle_tree_nodes(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out

    def point_order(self, which: int) -> np.ndarray:
        n = len(self.src) if which == 0 else len(self.trg)
        out = np.zeros(n, dtype=np.int64)
        lib().oracle_tree_point_order(self.h, which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out

    def list_rel(self, t: int) -> np.ndarray:
        c = lib().oracle_list_count(self.h, t)
        out = np.zeros(3 * c, dtype=np.int32)
        lib().oracle_list_rel(self.h, t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out.reshape(c, 3)

    def lists(self, t: int) -> np.ndarray:
        c = lib().oracle_list_count(self.h, t)
        out = np.zeros((self.n_nodes, c), dtype=np.int32)
        lib().oracle_tree_lists(self.h, t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def set_operators(self, ops: dict):
        self.ops = {k: (np.ascontiguousarray(v, dtype=np.float64) if isinstance(v, np.ndarray) else v)
                    for k, v in ops.items()}
        o = self.ops
        lib().oracle_set_operators(self.h, o["kernel"], o["m"], _dp(o["uc2ue0"]), _dp(o["uc2ue1"]),
                                   _dp(o["dc2de0"]), _dp(o["dc2de1"]), _dp(o["u2u"]), _dp(o["d2d"]),
                                   _dp(o["v1"]))
        if "bc" in o:
            lib().oracle_set_bc(self.h, _dp(o["bc"]))

    def evaluate(self, src_val, threads=None, return_phases=False):
        kid = self.ops["kernel"]
        k0, k1 = KDIM[kid]
        sv = _c(src_val).ravel()
        assert sv.size == k0 * len(self.src)
        out = np.zeros(k1 * len(self.trg))
        ph = np.zeros(8)
        lib().oracle_evaluate(self.h, _dp(sv), _dp(out), int(threads or nthreads()), _dp(ph))
        if return_phases:
            names = ["S2U", "U2U", "X-List", "W-List", "U-List", "V-List", "D2D", "D2T"]
            return out, dict(zip(names, ph.tolist()))
        return out


def direct(kernel, src, src_val, trg, threads=None) -> np.ndarray:
    """FP64 O(N*M) direct sum (r = 0 pairs excluded), for error measurement."""
    kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    k0, k1 = KDIM[kid]
    src = _c(src).reshape(-1, 3)
    trg = _c(trg).reshape(-1, 3)
    sv = _c(src_val).ravel()
    out = np.zeros(k1 * len(trg))
    lib().oracle_direct(kid, _dp(src), len(src), _dp(sv), _dp(trg), len(trg), _dp(out), int(threads or nthreads()))
    return out


def fmm(kernel, m, max_pts, src, src_val, trg=None, threads=None):
    """Convenience: tree + operators + evaluation in one call."""
    trg = src if trg is None else trg
    t = OracleTree(src, trg, max_pts)
    t.set_operators(precompute(kernel, m))
    return t.evaluate(src_val, threads=threads)
