"""Python interface of the MI355X particle FMM.

``FMMKernel``, ``KERNEL_DIMS`` and ``FMMParticleContext`` mirror
python/src/pvfmm/wrapper.py:28-47 and 92-179 of the reference (same names, argument meaning,
validation and ``ValueError`` behaviour).  ``PtFMM`` / ``PtFMMTree`` expose the tree-level
interface (FMM_Pts::Initialize, PtFMM_CreateTree, FMM_Tree::SetupFMM, PtFMM_Evaluate) that the
reference offers in C++ (include/pvfmm.hpp:199-313), including evaluation on device-resident
torch tensors.
"""
from __future__ import annotations

import ctypes
from enum import Enum
from typing import Optional

import numpy as np

from . import ffi


class FMMKernel(Enum):
    """Mirroring PVFMMKernel in include/pvfmm.h:21-28."""

    LaplacePotential = 0
    LaplaceGradient = 1
    StokesPressure = 2
    StokesVelocity = 3
    StokesVelocityGrad = 4
    BiotSavartPotential = 5


# python/src/pvfmm/wrapper.py:38-47
KERNEL_DIMS = {
    FMMKernel.LaplacePotential: (1, 1),
    FMMKernel.LaplaceGradient: (1, 3),
    FMMKernel.StokesPressure: (3, 1),
    FMMKernel.StokesVelocity: (3, 3),
    FMMKernel.StokesVelocityGrad: (3, 9),
    FMMKernel.BiotSavartPotential: (3, 3),
}

SUPPORTED_KERNELS = (FMMKernel.LaplacePotential, FMMKernel.LaplaceGradient, FMMKernel.StokesVelocity)


def _function_dtype(function_name: str, dtype) -> str:
    """python/src/pvfmm/wrapper.py:50-62."""
    if dtype == np.float64:
        return function_name + "D"
    if dtype == np.float32:
        return function_name + "F"
    raise ValueError("Invalid dtype, must be either float64 or float32")


def _comm_handle(comm) -> int:
    if comm is None:
        return 0
    for attr in ("handle", "py2f"):
        v = getattr(comm, attr, None)
        if v is not None:
            try:
                return int(v() if callable(v) else v)
            except Exception:  # noqa: BLE001 - any communicator object is accepted and ignored
                return 0
    return 0


class FMMParticleContext:
    """Particle FMM context (python/src/pvfmm/wrapper.py:92-179).

    ``comm`` is optional (the reference requires an mpi4py communicator); the evaluation runs
    on one GPU as the MPI_COMM_SELF share.  ``mixed_precision=True`` selects the FP32
    evaluator with FP64 operators and merge for a float64 context.
    """

    def __init__(self, box_size: float, max_points: int, multipole_order: int, kernel: FMMKernel,
                 comm=None, dtype=np.float64, mixed_precision: bool = False):
        self.kernel = kernel
        self.dtype = np.dtype(dtype)
        if multipole_order <= 0 or multipole_order % 2 != 0:
            raise ValueError("multipole order must be even and postive")
        fn = getattr(ffi.lib(), _function_dtype("PVFMMCreateContext", self.dtype))
        self._ptr = fn(float(box_size), max_points, multipole_order, int(self.kernel.value), _comm_handle(comm))
        if not self._ptr:
            ffi.check()
            raise ffi.PVFMMError(6, "context creation failed")
        if mixed_precision:
            ffi.check(ffi.lib().PVFMMSetMixedPrecision(self._ptr, 1))

    def __del__(self):
        if getattr(self, "_ptr", None):
            fn = getattr(ffi.lib(), _function_dtype("PVFMMDestroyContext", self.dtype))
            p = ctypes.c_void_p(self._ptr)
            fn(ctypes.byref(p))
            self._ptr = None

    def evaluate(self, src_pos: np.ndarray, sl_den: np.ndarray, dl_den: Optional[np.ndarray], trg_pos: np.ndarray,
                 setup: bool = True) -> np.ndarray:
        """python/src/pvfmm/wrapper.py:121-179; returns len(trg_pos) values like the reference."""
        if src_pos.dtype != self.dtype:
            raise ValueError(f"Source array had the wrong dtype: {src_pos.dtype}. "
                             f"This object was created with dtype {self.dtype}")
        source_length = len(src_pos)
        if source_length % 3 != 0:
            raise ValueError("Source arrays must have a length which is a multiple of 3")
        if sl_den is not None:
            if sl_den.dtype != self.dtype:
                raise ValueError(f"Source array had the wrong dtype: {src_pos.dtype}. "
                                 f"This object was created with dtype {self.dtype}")
            if len(sl_den) != source_length:
                raise ValueError("Source arrays must all be of the same length!")
        if dl_den is not None:
            if dl_den.dtype != self.dtype:
                raise ValueError(f"Source array had the wrong dtype: {src_pos.dtype}. "
                                 f"This object was created with dtype {self.dtype}")
            if len(dl_den) != source_length * 2:
                raise ValueError("Source arrays must all be of the same length!")
        n_src = source_length // 3
        target_length = len(trg_pos)
        if target_length % 3 != 0:
            raise ValueError("Target arrays must have a length which is a multiple of 3")
        n_trg = target_length // 3
        k1 = KERNEL_DIMS[self.kernel][1]
        # The reference sizes the output as len(trg_pos) (wrapper.py:166) whatever k1 is;
        # the same size is kept, large enough for k1 <= 3.
        trg_val = np.zeros(max(target_length, k1 * n_trg), dtype=self.dtype)
        src_pos = np.ascontiguousarray(src_pos)
        trg_pos = np.ascontiguousarray(trg_pos)
        sl = np.ascontiguousarray(sl_den) if sl_den is not None else None
        dl = np.ascontiguousarray(dl_den) if dl_den is not None else None
        ptr = ffi.dptr if self.dtype == np.float64 else ffi.fptr
        fn = getattr(ffi.lib(), _function_dtype("PVFMMEval", self.dtype))
        fn(ptr(src_pos), ptr(sl), ptr(dl), n_src, ptr(trg_pos), ptr(trg_val), n_trg, self._ptr, int(setup))
        ffi.check()
        return trg_val[:target_length]


# ----------------------------------------------------------------------------- tree-level API
_KERNEL_IDS = {"laplace_poten": 0, "laplace_grad": 1, "stokes_vel": 3}


def kernel_id(kernel) -> int:
    if isinstance(kernel, FMMKernel):
        return kernel.value
    if isinstance(kernel, str):
        return _KERNEL_IDS[kernel]
    return int(kernel)


def kernel_dims(kernel) -> tuple:
    return KERNEL_DIMS[FMMKernel(kernel_id(kernel))]


class PtFMM:
    """FMM_Pts::Initialize(m, comm, kernel) (include/fmm_pts.txx:223-331): operator precompute.

    precision: "fp64" or "mixed" (FP32 bulk evaluation, FP64 operators and merge).
    """

    def __init__(self, kernel, m: int, precision: str = "fp64", pinv_cutoff: float = 0.0):
        if m <= 0 or m % 2:
            raise ValueError("multipole order must be even and postive")
        self.kernel = kernel_id(kernel)
        self.m = m
        self.precision = precision
        p = {"fp64": 0, "mixed": 1}[precision]
        self._h = ffi.lib().pvfmm_amd_fmm_create(self.kernel, m, p, float(pinv_cutoff))
        if not self._h:
            ffi.check()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                ffi.lib().pvfmm_amd_fmm_destroy(self._h)
                self._h = None
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def operator(self, which: str) -> np.ndarray:
        idx = {"uc2ue0": 0, "uc2ue1": 1, "dc2de0": 2, "dc2de1": 3, "u2u": 4, "d2d": 5, "v1": 6,
               "bc_px": 7, "bc_pxy": 8, "bc_pxyz": 9}[which]
        n = ffi.lib().pvfmm_amd_fmm_operator(self._h, idx, None)
        if n < 0:
            ffi.check()
        out = np.zeros(n)
        ffi.lib().pvfmm_amd_fmm_operator(self._h, idx, ffi.dptr(out))
        return out


BOUNDARY = {"free": 0, "px": 1, "pxy": 2, "pxyz": 3, "periodic": 3}  # BoundaryType, include/mpi_tree.hpp:22-29


class PtFMMTree:
    """PtFMM_CreateTree + SetupFMM + PtFMM_Evaluate (include/pvfmm.txx:215-328).

    Coordinates are (n,3) arrays in [0,1)^3; densities and values are in input order with k0
    (k1) components per point.  bndry: "free" or a periodic boundary type ("px", "pxy",
    "pxyz" = "periodic"; unit period along the periodic axes).
    """

    def __init__(self, src, trg=None, max_pts: int = 100, pt=None, bndry: str = "free"):
        src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
        trg = src if trg is None else np.ascontiguousarray(trg, dtype=np.float64).reshape(-1, 3)
        self.n_src, self.n_trg = len(src), len(trg)
        if bndry not in BOUNDARY:
            raise ValueError(f"bndry must be one of {sorted(BOUNDARY)}")
        self.bndry = BOUNDARY[bndry]
        L = ffi.lib()
        if self.bndry:
            ptv = None if pt is None else np.ascontiguousarray(pt, dtype=np.float64).reshape(-1, 3)
            self._h = L.pvfmm_amd_tree_create_bndry(ffi.dptr(src), len(src), ffi.dptr(trg), len(trg),
                                                    ffi.dptr(ptv) if ptv is not None else None,
                                                    0 if ptv is None else len(ptv), int(max_pts), self.bndry)
        elif pt is None:
            self._h = L.pvfmm_amd_tree_create(ffi.dptr(src), len(src), ffi.dptr(trg), len(trg), int(max_pts))
        else:
            pt = np.ascontiguousarray(pt, dtype=np.float64).reshape(-1, 3)
            self._h = L.pvfmm_amd_tree_create_pts(ffi.dptr(src), len(src), ffi.dptr(trg), len(trg), ffi.dptr(pt),
                                                  len(pt), int(max_pts))
        if not self._h:
            ffi.check()
        self.fmm = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                ffi.lib().pvfmm_amd_tree_destroy(self._h)
                self._h = None
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    # -- introspection
    def num_nodes(self) -> int:
        return ffi.lib().pvfmm_amd_tree_num_nodes(self._h)

    def nodes(self) -> np.ndarray:
        out = np.zeros((self.num_nodes(), 10), dtype=np.int64)
        ffi.check(ffi.lib().pvfmm_amd_tree_nodes(self._h, out.ctypes.data_as(ffi.c_int64_p)))
        return out

    def point_order(self, which: int) -> np.ndarray:
        out = np.zeros(self.n_src if which == 0 else self.n_trg, dtype=np.int64)
        ffi.check(ffi.lib().pvfmm_amd_tree_point_order(self._h, which, out.ctypes.data_as(ffi.c_int64_p)))
        return out

    def lists(self, t: int) -> np.ndarray:
        c = ffi.lib().pvfmm_amd_list_count(t)
        out = np.zeros((self.num_nodes(), c), dtype=np.int32)
        ffi.check(ffi.lib().pvfmm_amd_tree_lists(self._h, t, out.ctypes.data_as(ffi.c_int32_p)))
        return out

    # -- setup / evaluation
    def setup(self, fmm: PtFMM, n_vec_max: int = 1):
        ffi.check(ffi.lib().pvfmm_amd_tree_setup(self._h, fmm._h, int(n_vec_max)))
        self.fmm = fmm
        return self

    def dims(self):
        return kernel_dims(self.fmm.kernel)

    def evaluate(self, den) -> np.ndarray:
        """Host arrays: den (n_vec, k0*n_src) or (k0*n_src,) -> values of the same leading shape."""
        k0, k1 = self.dims()
        den = np.ascontiguousarray(den, dtype=np.float64)
        n_vec = 1 if den.ndim == 1 else den.shape[0]
        if den.size != n_vec * k0 * self.n_src:
            raise ValueError("density array has the wrong size")
        out = np.zeros((n_vec, k1 * self.n_trg)) if den.ndim > 1 else np.zeros(k1 * self.n_trg)
        ffi.check(ffi.lib().pvfmm_amd_tree_evaluate(self._h, ffi.dptr(den), ffi.dptr(out), n_vec))
        return out

    def evaluate_device(self, den, out=None, sync: bool = True):
        """torch tensors on the GPU (float64, contiguous); returns the output tensor."""
        import torch

        k0, k1 = self.dims()
        if not (den.is_cuda and den.dtype == torch.float64 and den.is_contiguous()):
            raise ValueError("den must be a contiguous float64 GPU tensor")
        n_vec = 1 if den.dim() == 1 else den.shape[0]
        if den.numel() != n_vec * k0 * self.n_src:
            raise ValueError("density tensor has the wrong size")
        if out is None:
            shape = (k1 * self.n_trg,) if den.dim() == 1 else (n_vec, k1 * self.n_trg)
            out = torch.empty(shape, dtype=torch.float64, device=den.device)
        # order the library's stream after torch's current stream
        torch.cuda.current_stream(den.device).synchronize()
        ffi.check(ffi.lib().pvfmm_amd_tree_evaluate_device(self._h, ctypes.c_void_p(den.data_ptr()),
                                                           ctypes.c_void_p(out.data_ptr()), n_vec, 0 if sync else 1))
        return out

    def evaluate_adjoint(self, g) -> np.ndarray:
        """The transpose of evaluate: g (n_vec, k1*n_trg) or (k1*n_trg,) in target order ->
        A^T g of shape (n_vec, k0*n_src) or (k0*n_src,) in source order (symmetric kernels)."""
        k0, k1 = self.dims()
        g = np.ascontiguousarray(g, dtype=np.float64)
        n_vec = 1 if g.ndim == 1 else g.shape[0]
        if g.size != n_vec * k1 * self.n_trg:
            raise ValueError("adjoint input has the wrong size")
        out = np.zeros((n_vec, k0 * self.n_src)) if g.ndim > 1 else np.zeros(k0 * self.n_src)
        ffi.check(ffi.lib().pvfmm_amd_tree_evaluate_adjoint(self._h, ffi.dptr(g), ffi.dptr(out), n_vec))
        return out

    def evaluate_adjoint_device(self, g, out=None, sync: bool = True):
        """evaluate_adjoint on contiguous float64 GPU tensors; returns the output tensor."""
        import torch

        k0, k1 = self.dims()
        if not (g.is_cuda and g.dtype == torch.float64 and g.is_contiguous()):
            raise ValueError("g must be a contiguous float64 GPU tensor")
        n_vec = 1 if g.dim() == 1 else g.shape[0]
        if g.numel() != n_vec * k1 * self.n_trg:
            raise ValueError("adjoint input tensor has the wrong size")
        if out is None:
            shape = (k0 * self.n_src,) if g.dim() == 1 else (n_vec, k0 * self.n_src)
            out = torch.empty(shape, dtype=torch.float64, device=g.device)
        torch.cuda.current_stream(g.device).synchronize()
        ffi.check(ffi.lib().pvfmm_amd_tree_evaluate_adjoint_device(self._h, ctypes.c_void_p(g.data_ptr()),
                                                                   ctypes.c_void_p(out.data_ptr()), n_vec,
                                                                   0 if sync else 1))
        return out

    def stream_handle(self) -> int:
        return ffi.lib().pvfmm_amd_tree_stream(self._h) or 0

    # -- launch-shape tuning
    TUNING_KEYS = ("hadamard_tiles_per_block", "gemm_split_blocks", "hadamard_vec_targets_per_wave")

    def autotune(self, den=None) -> dict:
        """Times the launch-shape candidates on this tree with density vectors on the GPU (one
        seeded N(0,1) vector when none is given; a (n_vec, k0*n_src) tensor tunes the batched
        pass of n_vec <= n_vec_max vectors), keeps the fastest and returns it."""
        import torch

        k0, k1 = self.dims()
        if den is None:
            g = torch.Generator(device="cuda:0").manual_seed(0)
            den = torch.randn(k0 * self.n_src, dtype=torch.float64, device="cuda:0", generator=g)
        n_vec = 1 if den.dim() == 1 else den.shape[0]
        if not (den.is_cuda and den.dtype == torch.float64 and den.is_contiguous() and den.numel() == n_vec * k0 * self.n_src):
            raise ValueError("den must be contiguous float64 GPU density vectors of k0*n_src values")
        out = torch.empty(n_vec * k1 * self.n_trg, dtype=torch.float64, device=den.device)
        torch.cuda.current_stream(den.device).synchronize()
        p = (ctypes.c_int32 * 3)()
        ffi.check(ffi.lib().pvfmm_amd_tree_autotune(self._h, ctypes.c_void_p(den.data_ptr()),
                                                    ctypes.c_void_p(out.data_ptr()), n_vec, p, 3))
        return dict(zip(self.TUNING_KEYS, list(p)))

    def tuning(self) -> dict:
        p = (ctypes.c_int32 * 3)()
        ffi.check(ffi.lib().pvfmm_amd_tree_get_tuning(self._h, p, 3))
        return dict(zip(self.TUNING_KEYS, list(p)))

    def set_tuning(self, **kw):
        cur = self.tuning()
        for k, v in kw.items():
            if k not in cur:
                raise ValueError(f"unknown tuning key {k}")
            cur[k] = int(v)
        p = (ctypes.c_int32 * 3)(*[cur[k] for k in self.TUNING_KEYS])
        ffi.check(ffi.lib().pvfmm_amd_tree_set_tuning(self._h, p, 3))
        return self

    def set_timing(self, on: bool = True):
        ffi.check(ffi.lib().pvfmm_amd_tree_set_timing(self._h, int(on)))

    def phase_times(self) -> dict:
        buf = np.zeros(32)
        n = ffi.lib().pvfmm_amd_tree_phase_times(self._h, ffi.dptr(buf), 32)
        return {ffi.lib().pvfmm_amd_phase_name(i).decode(): float(buf[i]) for i in range(n)}

    def stats(self) -> dict:
        names = ["n_nodes", "n_leaves", "max_depth", "s2m_rows", "l2t_rows", "m2m_rows", "l2l_rows", "x_rows",
                 "x_items", "v_src_slots", "v_trg_slots", "v_pairs", "u_rows", "u_items", "u_pairs", "far_items",
                 "far_pairs", "s2m_pairs", "u_src_points_streamed", "device_bytes", "v_tiles", "v_tile_partners",
                 "v_tile_max_partners", "v_tile_staged"]
        out = np.zeros(len(names), dtype=np.int64)
        n = ffi.lib().pvfmm_amd_tree_stats(self._h, out.ctypes.data_as(ffi.c_int64_p), len(names))
        if n < 0:
            ffi.check(-n)
        return dict(zip(names, out[:n].tolist()))


class PtFMMBatch:
    """Set-up trees of independent point clouds evaluated together: the first evaluation
    records every tree's evaluation into one HIP graph (one branch per tree), later calls with
    the same tensors replay it with one graph launch.  A tree set up again (with another PtFMM
    or n_vec_max) or retuned makes the next call record the graph again.  Replays are ordered
    against eager evaluations of the same trees (they share the trees' scratch buffers), also
    with sync=False.  The trees must stay alive and set up."""

    def __init__(self, trees):
        self.trees = list(trees)
        if not self.trees:
            raise ValueError("a batch needs at least one tree")
        arr = (ctypes.c_void_p * len(self.trees))(*[t._h for t in self.trees])
        self._h = ffi.lib().pvfmm_amd_batch_create(arr, len(self.trees))
        if not self._h:
            ffi.check()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                ffi.lib().pvfmm_amd_batch_destroy(self._h)
                self._h = None
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def evaluate_device(self, dens, outs=None, sync: bool = True):
        """dens: one contiguous float64 GPU tensor per tree (k0*n_src); returns the outputs."""
        import torch

        if len(dens) != len(self.trees):
            raise ValueError("one density tensor per tree")
        if outs is None:
            outs = [torch.empty(t.dims()[1] * t.n_trg, dtype=torch.float64, device=d.device)
                    for t, d in zip(self.trees, dens)]
        for t, d, o in zip(self.trees, dens, outs):
            k0, k1 = t.dims()
            if not (d.is_cuda and d.dtype == torch.float64 and d.is_contiguous() and d.numel() == k0 * t.n_src):
                raise ValueError("each density must be a contiguous float64 GPU tensor of k0*n_src values")
            if not (o.is_cuda and o.dtype == torch.float64 and o.is_contiguous() and o.numel() == k1 * t.n_trg):
                raise ValueError("each output must be a contiguous float64 GPU tensor of k1*n_trg values")
        n = len(self.trees)
        dp = (ctypes.c_void_p * n)(*[d.data_ptr() for d in dens])
        op = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        torch.cuda.current_stream(dens[0].device).synchronize()
        ffi.check(ffi.lib().pvfmm_amd_batch_evaluate_device(self._h, dp, op, 0 if sync else 1))
        return outs

    def stream_handle(self) -> int:
        return ffi.lib().pvfmm_amd_batch_stream(self._h) or 0

    def records(self) -> int:
        return ffi.lib().pvfmm_amd_batch_records(self._h)


def device_count() -> int:
    return ffi.lib().pvfmm_amd_device_count()


def direct(kernel, src, den, trg) -> np.ndarray:
    """Kernel<T>::ker_poten (include/kernel.hpp:84): FP64 all-pairs sum on the GPU.

    src [n_src, 3], den [n_src * k0], trg [n_trg, 3] -> [n_trg * k1] (AoS, input order)."""
    kid = kernel_id(kernel)
    k0, k1 = kernel_dims(kernel)
    s = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(trg, dtype=np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(den, dtype=np.float64).ravel()
    if d.size != k0 * len(s):
        raise ValueError("den must hold k0 values per source")
    out = np.zeros(k1 * len(t))
    ffi.check(ffi.lib().pvfmm_amd_direct(kid, ffi.dptr(s), len(s), ffi.dptr(d), ffi.dptr(t), len(t), ffi.dptr(out)))
    return out
