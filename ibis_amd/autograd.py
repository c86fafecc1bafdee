"""The residual and the boundary force as differentiable torch operations.

Forward passes run the MI355X residual; backward passes run its hand-written
discrete adjoint (ibis_compute_dudt_vjp), which recomputes reconstructed
states instead of storing them.  The limiter is differentiated ("active") or
held fixed ("frozen").

``fv`` is either ibis_amd.finite_volume.FiniteVolume (ctypes on the C ABI)
or ibis_amd.pyapi.PyFiniteVolume (the pybind11 module python_api); both
give the same bits.
"""
from __future__ import annotations

import torch

from .finite_volume import FiniteVolume


class ResidualFunction(torch.autograd.Function):
    """dU/dt = R(prim) with prim a (7, n_total) float64 CUDA tensor; only
    the valid rows of prim receive gradient (ghost rows are set by the
    boundary actions)."""

    @staticmethod
    def forward(ctx, prim, fv: FiniteVolume, limiter: str = "active"):
        work = prim.detach().clone()
        dudt = torch.zeros((fv.ncons, fv.grid.Nt), dtype=torch.float64, device=prim.device)
        fv.compute_dudt(work, fv.grid, dudt)
        ctx.save_for_backward(prim.detach())
        ctx.fv, ctx.limiter = fv, limiter
        return dudt

    @staticmethod
    def backward(ctx, dudt_bar):
        (prim,) = ctx.saved_tensors
        fv = ctx.fv
        bar = torch.zeros_like(prim)
        db = dudt_bar.detach().contiguous().clone()
        db[:, fv.grid.N:] = 0.0
        fv.compute_dudt_vjp(prim.clone(), bar, dudt_bar=db, limiter=ctx.limiter)
        return bar, None, None


class WallForceFunction(torch.autograd.Function):
    """J = force_dir . (pressure + viscous force on the boundaries in tags)."""

    @staticmethod
    def forward(ctx, prim, fv: FiniteVolume, tags: int, force_dir, limiter: str = "active"):
        f = fv.wall_force(prim.detach().clone(), tags)
        ctx.save_for_backward(prim.detach())
        ctx.fv, ctx.tags, ctx.dir, ctx.limiter = fv, tags, tuple(force_dir), limiter
        return torch.tensor(sum(a * b for a, b in zip(f, ctx.dir)), dtype=torch.float64,
                            device=prim.device)

    @staticmethod
    def backward(ctx, g):
        (prim,) = ctx.saved_tensors
        bar = torch.zeros_like(prim)
        ctx.fv.compute_dudt_vjp(prim.clone(), bar, force_tags=ctx.tags, force_dir=ctx.dir,
                                limiter=ctx.limiter)
        return bar * g, None, None, None, None


def residual(prim, fv: FiniteVolume, limiter: str = "active"):
    return ResidualFunction.apply(prim, fv, limiter)


def wall_force(prim, fv: FiniteVolume, tags, force_dir=(1.0, 0.0, 0.0), limiter: str = "active"):
    mask = fv.tag_mask(tags) if not isinstance(tags, int) else tags
    return WallForceFunction.apply(prim, fv, mask, force_dir, limiter)
