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
  * whole FMM: relative L2 against an FP64 direct sum, convergence in m.
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
        L.oracle_tree_create.argtypes = [P(dp), i64, P(dp), i64, P(dp), i64, i32, i32]
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


# ----------------------------------------------------------------------------- tree + FMM
class OracleTree:
    """PtFMM_CreateTree + FMM_Tree::SetupFMM restated (include/pvfmm.txx:215-245)."""

    def __init__(self, src, trg, max_pts, pt=None, refine=True):
        L = lib()
        self.src = _c(src).reshape(-1, 3)
        self.trg = _c(trg).reshape(-1, 3)
        self.pt = self.trg if pt is None else _c(pt).reshape(-1, 3)
        self.h = L.oracle_tree_create(_dp(self.src), len(self.src), _dp(self.trg), len(self.trg),
                                      _dp(self.pt), len(self.pt), int(max_pts), int(refine))
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
