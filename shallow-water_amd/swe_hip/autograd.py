"""torch.autograd.Function over the HIP forward solver and its discrete adjoint.

    hN, huN, hvN = swe_rollout(h0, hu0, hv0, B, grid, nsteps, dt=..., nsnap=0)

h0, hu0, hv0, B: float64 CUDA(ROCm) tensors of the grid shape (nx+4, ny+4) in the
reference layout; ghost entries are ignored (the solver mirrors the interior) and get zero
gradient.  The forward runs `nsteps` SSP-RK3 steps on the GPU with binomial checkpoints in
device memory; backward calls the hand-written adjoint kernels with dt frozen from the
forward pass (fixed `dt`, given `dts`, or recorded from `cfl`).  The library runs on its
own stream; the wrapper synchronizes the device before handing torch's buffers over, and
every library call used here returns only after its stream has drained.
"""
from __future__ import annotations

import torch

from . import GridDesc, HipAdjoint, HipSolver


class SweRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h0, hu0, hv0, B, grid: dict, nsteps: int, dt=None, cfl=None, dts=None,
                nsnap: int = 0):
        for t in (h0, hu0, hv0, B):
            if t.dtype != torch.float64 or not t.is_cuda:
                raise TypeError("swe_rollout needs float64 device tensors")
        dev = h0.device.index or 0
        solver = HipSolver(grid["nx"], grid["ny"], grid["dx"], grid["dy"], grid["g"],
                           grid.get("bc", 0), 2, dev)
        h0, hu0, hv0, B = (t.contiguous() for t in (h0, hu0, hv0, B))
        torch.cuda.synchronize(h0.device)
        solver.upload_device(h0, hu0, hv0, B)
        adj = HipAdjoint(solver, nsteps, nsnap)
        used = adj.forward(dt=dt, cfl=cfl, dts=dts)
        outs = [torch.empty_like(h0) for _ in range(3)]
        solver.download_device(*outs)
        ctx.solver, ctx.adj = solver, adj
        ctx.dts = used
        return tuple(outs)

    @staticmethod
    def backward(ctx, gh, ghu, ghv):
        solver, adj = ctx.solver, ctx.adj
        ref = next(t for t in (gh, ghu, ghv) if t is not None)
        lam = [(t if t is not None else torch.zeros_like(ref)).contiguous().to(torch.float64)
               for t in (gh, ghu, ghv)]
        out = [torch.empty_like(lam[0]) for _ in range(4)]
        torch.cuda.synchronize(lam[0].device)
        adj.backward_device(*lam, *out)
        adj.close()
        solver.close()
        return out[0], out[1], out[2], out[3], None, None, None, None, None, None


def _plain_rollout(h0, hu0, hv0, B, grid: dict, nsteps: int, dt, cfl, dts):
    """The same steps without an adjoint: no checkpoint store is allocated."""
    for t in (h0, hu0, hv0, B):
        if t.dtype != torch.float64 or not t.is_cuda:
            raise TypeError("swe_rollout needs float64 device tensors")
    dev = h0.device.index or 0
    h0, hu0, hv0, B = (t.detach().contiguous() for t in (h0, hu0, hv0, B))
    with HipSolver(grid["nx"], grid["ny"], grid["dx"], grid["dy"], grid["g"],
                   grid.get("bc", 0), 2, dev) as solver:
        torch.cuda.synchronize(h0.device)
        solver.upload_device(h0, hu0, hv0, B)
        if dts is not None:
            for d in dts:
                solver.step(float(d), 1)
        elif cfl is not None:
            solver.step_adaptive(float(cfl), nsteps)
        elif nsteps > 0:
            solver.step(float(dt), nsteps)
        outs = [torch.empty_like(h0) for _ in range(3)]
        solver.download_device(*outs)
    return tuple(outs)


def swe_rollout(h0, hu0, hv0, B, grid: dict, nsteps: int, dt=None, cfl=None, dts=None,
                nsnap: int = 0):
    """Differentiable N-step rollout; exactly one of dt / cfl / dts.  When no gradient can
    be asked for (grad mode off, or no input requires grad) the steps run without the
    adjoint's checkpoint store."""
    if sum(x is not None for x in (dt, cfl, dts)) != 1:
        raise ValueError("give exactly one of dt, cfl, dts")
    if dts is not None and len(dts) != int(nsteps):
        raise ValueError("dts must hold nsteps values")
    if not (torch.is_grad_enabled() and any(t.requires_grad for t in (h0, hu0, hv0, B))):
        return _plain_rollout(h0, hu0, hv0, B, grid, int(nsteps), dt, cfl, dts)
    return SweRollout.apply(h0, hu0, hv0, B, grid, int(nsteps), dt, cfl, dts, int(nsnap))


__all__ = ["SweRollout", "swe_rollout", "GridDesc"]
