"""Differentiable Laplace FMM (adjoint FMM behind torch.autograd.Function).

For sources = targets = x (n points) and charges q the FMM gives the potential phi = A q, an
approximation of phi_i = sum_{j != i} q_j / (4 pi |x_i - x_j|) (laplace_poten,
include/kernel.txx:1276-1287), and the field u_i ~ sum_{j != i} q_j (x_i - x_j) /
(4 pi |x_i - x_j|^3) = -grad phi(x_i) (laplace_grad, include/kernel.txx:1316-1330).  A is the
FMM's own linear map; its transpose A^T is evaluated by the adjoint FMM
(PtFMMTree.evaluate_adjoint: the phases in reverse order with the transposed translations), so
the gradients with respect to the charges are those of the discrete map, exactly:

* potential phi(x, q) = A q with upstream gradient g:
    dL/dq = A^T g                      (one adjoint FMM)
    dL/dx_i = -g_i u[q]_i - q_i u[g]_i (one gradient FMM with the two densities q and g)
* energy E(x, q) = 1/2 q . A q:
    dE/dq = (A q + A^T q) / 2          (A q kept from the forward pass, one adjoint FMM)
    dE/dx_i = -q_i u[q]_i              (one gradient FMM, run beside the adjoint)

A approximates the symmetric kernel matrix, so A^T g and A g agree to the FMM's accuracy; the
adjoint makes the charge gradients consistent with the computed values to rounding.  The
positions enter through the tree: the coordinate gradients come from the gradient-kernel FMM
(the derivative of the exact kernel sum), not from differentiating the tree.

The tree is built once for the point set given to LaplaceFMM; x passed to potential()/energy()
must be those positions and only receives the gradient.  All FMM work runs in libpvfmm.so on
the GPU.  With a periodic boundary (bndry "px", "pxy" or "pxyz") the adjoint runs through the
periodic images and the lattice operator as well; the position gradients are those of a
neutral charge set there (the gradient kernel carries no uniform-background term).
"""
from __future__ import annotations

import numpy as np

from .wrapper import PtFMM, PtFMMTree


class LaplaceFMM:
    """Potential and energy of a fixed point set with FMM forward and adjoint-FMM backward.

    precision: "fp64" or "mixed" (FP32 bulk, FP64 operators and merge); bndry: "free" or a
    periodic boundary type of PtFMMTree."""

    def __init__(self, points, m: int = 10, max_pts: int = 256, precision: str = "mixed", bndry: str = "free"):
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        self.n = len(pts)
        self.points = pts
        self._tp = PtFMMTree(pts, pts, max_pts, bndry=bndry).setup(PtFMM("laplace_poten", m, precision), n_vec_max=2)
        self._tg = PtFMMTree(pts, pts, max_pts, bndry=bndry).setup(PtFMM("laplace_grad", m, precision), n_vec_max=2)

    # -- raw evaluations on float64 GPU tensors
    def _phi(self, q):
        return self._tp.evaluate_device(q.contiguous())

    def _phi_adjoint(self, g):
        return self._tp.evaluate_adjoint_device(g.contiguous())

    def _field(self, dens):
        """dens (k, n) -> (k, n, 3) fields -grad phi of each density."""
        out = self._tg.evaluate_device(dens.contiguous())
        return out.reshape(dens.shape[0], self.n, 3)

    def _adjoint_and_field(self, g, dens):
        """A^T g and the fields of dens, queued on the two trees' own streams so that the two
        FMMs may overlap on the GPU; returns after both complete."""
        import torch

        import ctypes

        from . import ffi

        g = g.contiguous()
        dens = dens.contiguous()
        adj = torch.empty_like(g)
        fld = torch.empty((dens.shape[0], 3 * self.n), dtype=torch.float64, device=dens.device)
        torch.cuda.current_stream(g.device).synchronize()
        ffi.check(ffi.lib().pvfmm_amd_tree_evaluate_adjoint_device(self._tp._h, ctypes.c_void_p(g.data_ptr()),
                                                                   ctypes.c_void_p(adj.data_ptr()), 1, 1))
        ffi.check(ffi.lib().pvfmm_amd_tree_evaluate_device(self._tg._h, ctypes.c_void_p(dens.data_ptr()),
                                                           ctypes.c_void_p(fld.data_ptr()), dens.shape[0], 1))
        torch.cuda.ExternalStream(self._tp.stream_handle()).synchronize()
        torch.cuda.ExternalStream(self._tg.stream_handle()).synchronize()
        return adj, fld.reshape(dens.shape[0], self.n, 3)

    def potential(self, x, q):
        return _Potential.apply(x, q, self)

    def energy(self, x, q):
        return _Energy.apply(x, q, self)


def _check(fmm: LaplaceFMM, x, q):
    import torch

    if x.shape != (fmm.n, 3) or q.shape != (fmm.n,):
        raise ValueError("x must be (n, 3) and q (n,) for the point set of this LaplaceFMM")
    if not (q.is_cuda and q.dtype == torch.float64):
        raise ValueError("q must be a float64 GPU tensor")


def _make_functions():
    import torch

    class Potential(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, q, fmm):
            _check(fmm, x, q)
            ctx.fmm = fmm
            ctx.save_for_backward(q)
            return fmm._phi(q.detach())

        @staticmethod
        def backward(ctx, g):
            (q,) = ctx.saved_tensors
            fmm = ctx.fmm
            g = g.contiguous()
            dq = dx = None
            if ctx.needs_input_grad[0]:
                if ctx.needs_input_grad[1]:
                    dq, u = fmm._adjoint_and_field(g, torch.stack([q.detach(), g]))
                else:
                    u = fmm._field(torch.stack([q.detach(), g]))
                dx = -(g[:, None] * u[0] + q.detach()[:, None] * u[1])
            elif ctx.needs_input_grad[1]:
                dq = fmm._phi_adjoint(g)
            return dx, dq, None

    class Energy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, q, fmm):
            _check(fmm, x, q)
            phi = fmm._phi(q.detach())
            ctx.fmm = fmm
            ctx.save_for_backward(q, phi)
            return 0.5 * torch.dot(q.detach(), phi)

        @staticmethod
        def backward(ctx, gE):
            q, phi = ctx.saved_tensors
            fmm = ctx.fmm
            qd = q.detach()
            dq = dx = None
            if ctx.needs_input_grad[0]:
                if ctx.needs_input_grad[1]:
                    aq, u = fmm._adjoint_and_field(qd, qd[None, :])
                    dq = gE * 0.5 * (phi + aq)
                else:
                    u = fmm._field(qd[None, :])
                dx = -gE * qd[:, None] * u[0]
            elif ctx.needs_input_grad[1]:
                dq = gE * 0.5 * (phi + fmm._phi_adjoint(qd))
            return dx, dq, None

    return Potential, Energy


class _Lazy:
    """Defers the torch import (and Function class creation) to first use."""

    def __init__(self, idx):
        self.idx = idx
        self.fn = None

    def apply(self, *args):
        if self.fn is None:
            self.fn = _make_functions()[self.idx]
        return self.fn.apply(*args)


_Potential = _Lazy(0)
_Energy = _Lazy(1)
