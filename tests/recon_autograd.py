"""Test-side differentiable wrappers over the C ABI (torch.autograd.Function).

* JosephProject:      y = A x   (mr_joseph_forward), backward g -> A^T g
* JosephBackproject:  x = A^T y (mr_joseph_adjoint), backward g -> A g
* FilteredBackprojection: vol = BP(F p) with MuhRec's ProjectionFilterSingle
  and MultiProjBPparallel, backward g -> F^T BP^T g (mr_filter_rows_adjoint,
  mr_bp_parallel_adjoint)
* tv():               smoothed isotropic total variation (torch autograd)
* gradient_descent(): min ||A x - y||^2 + lam * TV(x), recomputing A x every
  iteration instead of storing intermediates.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from imagingsuite_amd import _lib


JosephGeom = _lib.JosephGeom


def _lib_j():
    return _lib.load()


class Geometry:
    def __init__(self, nx: int, nz: int, n_det: int, angles, center: float, direction: int = 0):
        self.g = JosephGeom(nx, nz, n_det, center, direction)
        self.angles = np.ascontiguousarray(angles, np.float32)
        self.nx, self.nz, self.n_det = nx, nz, n_det

    def forward(self, vol: torch.Tensor) -> torch.Tensor:
        vol = vol.contiguous()
        out = torch.empty((self.angles.size, self.n_det, self.nz), dtype=torch.float32, device=vol.device)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.check(_lib_j().mr_joseph_forward(ctypes.byref(self.g), ctypes.c_void_p(vol.data_ptr()),
                                              self.angles.ctypes.data_as(fp), self.angles.size,
                                              ctypes.c_void_p(out.data_ptr()),
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out

    def adjoint(self, sino: torch.Tensor) -> torch.Tensor:
        sino = sino.contiguous()
        out = torch.empty((self.nx, self.nx, self.nz), dtype=torch.float32, device=sino.device)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.check(_lib_j().mr_joseph_adjoint(ctypes.byref(self.g), ctypes.c_void_p(sino.data_ptr()),
                                              self.angles.ctypes.data_as(fp), self.angles.size,
                                              ctypes.c_void_p(out.data_ptr()),
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out


class JosephProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, geom: Geometry):
        ctx.geom = geom
        return geom.forward(vol)

    @staticmethod
    def backward(ctx, g):
        return ctx.geom.adjoint(g), None


class JosephBackproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, geom: Geometry):
        ctx.geom = geom
        return geom.adjoint(sino)

    @staticmethod
    def backward(ctx, g):
        return ctx.geom.forward(g), None


def tv(x: torch.Tensor, eps: float = 1e-6) -> torch.Tensor:
    """Smoothed isotropic TV of a ZXY volume [y, x, z] (forward differences)."""
    dy = torch.diff(x, dim=0, append=x[-1:])
    dx = torch.diff(x, dim=1, append=x[:, -1:])
    dz = torch.diff(x, dim=2, append=x[:, :, -1:])
    return torch.sqrt(dx * dx + dy * dy + dz * dz + eps * eps).sum()


def objective(x, y, geom: Geometry, lam: float, eps: float = 1e-2):
    r = JosephProject.apply(x, geom) - y
    return (r * r).sum() + lam * tv(x, eps)


def gradient_descent(y, geom: Geometry, iters: int = 50, lam: float = 1e-3, step=None, x0=None,
                     eps: float = 1e-2):
    """Plain gradient descent with a fixed step 1/L, L = 2 ||A^T A|| (power
    iteration) + 12 lam / eps (TV); returns (x, losses, step)."""
    dev = y.device
    if step is None:
        v = torch.ones((geom.nx, geom.nx, geom.nz), device=dev)
        for _ in range(8):
            v = geom.adjoint(geom.forward(v))
            nrm = v.norm()
            v = v / nrm
        step = 1.0 / (2.0 * float(nrm) + 12.0 * lam / eps)
    x = torch.zeros((geom.nx, geom.nx, geom.nz), device=dev) if x0 is None else x0.clone()
    losses = []
    for _ in range(iters):
        x.requires_grad_(True)
        loss = objective(x, y, geom, lam, eps)
        (g,) = torch.autograd.grad(loss, x)
        losses.append(float(loss))
        with torch.no_grad():
            x = x - step * g
    return x.detach(), losses, step


class FBPOperator:
    """ProjectionFilterSingle + MultiProjBPparallel on one block, with its
    adjoint; the projections are [n_proj, slices, width]."""

    def __init__(self, width: int, slices: int, center: float, angles, weights, filtertype="hamming",
                 buffer_size: int = 16, fused: bool = False):
        from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle
        self.filt = ProjectionFilterSingle(width, filtertype=filtertype, cutoff=0.5, usebias=True,
                                           biasweight=0.1, paddingdoubler=2)
        self.bp = BackProjector("MultiProjBPparallel", width, slices, center, direction=0,
                                buffer_size=buffer_size, fused=fused)
        self.angles = np.ascontiguousarray(angles, np.float32)
        self.weights = np.ascontiguousarray(weights, np.float32)

    def forward(self, proj: torch.Tensor) -> torch.Tensor:
        work = proj.contiguous().clone()
        self.filt.process(work)
        return self.bp.process(work, self.angles, self.weights)

    def adjoint(self, grad_vol: torch.Tensor) -> torch.Tensor:
        g = self.bp.adjoint(grad_vol.contiguous(), self.angles, self.weights)
        return self.filt.adjoint(g)


class FilteredBackprojection(torch.autograd.Function):
    """vol = FBP(proj); the backward recomputes nothing and stores nothing:
    the adjoint needs only the geometry."""

    @staticmethod
    def forward(ctx, proj, op: FBPOperator):
        ctx.op = op
        return op.forward(proj)

    @staticmethod
    def backward(ctx, g):
        return ctx.op.adjoint(g), None
