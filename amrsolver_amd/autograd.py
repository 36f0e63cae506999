"""torch.autograd bindings of the C ABI's forward and reverse entry points,
for the validation harness: gradients of torch code that uses a level's
right-hand side or its time steps come from the hand-written reverse mode
(amr_vjp_rhs, amr_adjoint_run), not from torch's own differentiation.

States are (5, nz, ny, nx) float64 device tensors of the level's valid cells
(the level's full index space; cells no box covers are ignored and get a
zero gradient).  The ghost fill is part of the operator, so gradients
include the transposed fill.
"""
import torch

__all__ = ["level_rhs", "level_steps"]


def _covered(lev, like):
    mask = torch.zeros(like.shape[1:], dtype=torch.bool, device=like.device)
    for lo, hi in lev.boxes:
        mask[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return mask


class _LevelRHS(torch.autograd.Function):
    """dSdt = L(U) after the level's ghost fill (CNS::compute_rhs)."""

    @staticmethod
    def forward(ctx, u, lev, layout):
        S = lev.new_state(layout=layout)
        D = lev.new_state(layout=layout, ng=0)
        S.upload_tensor(u)
        lev.fill_patch(S)
        lev.compute_rhs(S, D)
        ctx.lev, ctx.S, ctx.layout = lev, S, layout
        return D.download_tensor()

    @staticmethod
    def backward(ctx, g):
        lev, layout = ctx.lev, ctx.layout
        Lm = lev.new_state(layout=layout, ng=0)
        G = lev.new_state(layout=layout, ng=0)
        Lm.upload_tensor(torch.where(_covered(lev, g), g, torch.zeros_like(g)).contiguous())
        lev.vjp_rhs(ctx.S, Lm, G)
        return G.download_tensor(), None, None


class _LevelSteps(torch.autograd.Function):
    """U after `nsteps` calls of amr_advance with a fixed dt; backward runs
    the checkpointed adjoint (checkpoints every `kcheck` steps) with the
    upstream gradient as the weights of a linear objective."""

    @staticmethod
    def forward(ctx, u, lev, dt, nsteps, kcheck, layout):
        S0 = lev.new_state(layout=layout)
        S0.upload_tensor(u)
        lev.fill_patch(S0)
        S = lev.new_state(layout=layout)
        S.upload_tensor(u)
        for _ in range(nsteps):
            lev.advance(S, 0.0, dt)
        ctx.lev, ctx.S0, ctx.args = lev, S0, (dt, nsteps, kcheck, layout)
        return S.download_tensor()

    @staticmethod
    def backward(ctx, g):
        lev = ctx.lev
        dt, nsteps, kcheck, layout = ctx.args
        W = lev.new_state(layout=layout, ng=0)
        grad = lev.new_state(layout=layout, ng=0)
        W.upload_tensor(torch.where(_covered(lev, g), g, torch.zeros_like(g)).contiguous())
        lev.adjoint_run(ctx.S0, dt, nsteps, kcheck, grad, objective=1, weights=W)
        return grad.download_tensor(), None, None, None, None, None


def level_rhs(u, lev, layout="level"):
    """dSdt of level `lev` at state u, differentiable in u."""
    return _LevelRHS.apply(u, lev, layout)


def level_steps(u, lev, dt, nsteps, kcheck=1, layout="level"):
    """The state after `nsteps` SSPRK steps of size dt, differentiable in u."""
    return _LevelSteps.apply(u, lev, dt, nsteps, kcheck, layout)
