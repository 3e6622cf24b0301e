"""Differentiable distortion energy and ARAP global step on the GPU.

The reference computes no gradients; this module adds them on top of the
same plan and kernels (param.py:294-338 energies, param.py:391-401 and
:457-474 global step):

* ``island_energy(plan, U, w, kind)`` -- per-island sum_f w_f e_f(U) (ARAP
  kind 0, symmetric Dirichlet kind 1) with a hand-written backward kernel for
  dE/dU and dE/dw.
* ``arap_step(plan, U)`` -- one refine global step U' = L^-1 b(U) (rotations,
  rhs, CG); its backward solves the adjoint system L lam = dLoss/dU' with the
  same SPD operator and maps lam back through the rhs and the rotations.
* ``arap_unrolled(plan, U0, steps)`` -- several steps; autograd keeps every
  step's input layout (the per-iteration UV checkpoints) for the backward.
* ``stretch_refine(plan, U, w, steps)`` -- gradient-based refinement of the
  weighted symmetric Dirichlet (stretch) energy: Laplacian-preconditioned
  gradient (one solve with the global step's operator), flip-safe step
  (0.9 of the largest fold-free step, per island) and Armijo backtracking.

Every call runs on the plan's GPU; plans must use refine conventions
(``IslandPlan(..., batch=False)``).
"""
from __future__ import annotations

import torch

from . import _lib
from .plan import CG_TOL, IslandPlan

ADJOINT_TOL = 1e-10   # relative residual of the adjoint solve (FP64)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, w, plan, kind):
        ctx.plan, ctx.kind = plan, kind
        ctx.save_for_backward(U, w)
        return plan.energy_fwd(U, w, kind)

    @staticmethod
    def backward(ctx, ge):
        U, w = ctx.saved_tensors
        want_w = w is not None and ctx.needs_input_grad[1]
        gU, gw = ctx.plan.energy_bwd(U.contiguous(), ge, w, ctx.kind, want_w=want_w)
        return (gU if ctx.needs_input_grad[0] else None), (gw if want_w else None), None, None


def island_energy(plan: IslandPlan, U: torch.Tensor, w: torch.Tensor | None = None, kind: int = 0):
    """(n_islands,) weighted energies; differentiable in U and w."""
    if w is None:
        return _Energy.apply(U.contiguous(), None, plan, int(kind))
    return _Energy.apply(U.contiguous(), w.contiguous(), plan, int(kind))


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, plan, tol, precision, adjoint_tol):
        U = U.contiguous()
        b = plan.local_rhs(U, _lib.UVA_MODE_REFINE)
        x = U.clone()
        plan.cg_solve(b, x, tol=tol, precision=precision)
        ctx.plan, ctx.adjoint_tol = plan, adjoint_tol
        ctx.save_for_backward(U)
        return x

    @staticmethod
    def backward(ctx, g):
        (U,) = ctx.saved_tensors
        lam = ctx.plan.adjoint_cg(g.contiguous(), tol=ctx.adjoint_tol)
        return ctx.plan.step_bwd(U, lam), None, None, None, None


def arap_step(plan: IslandPlan, U: torch.Tensor, tol: float = CG_TOL, precision: int = _lib.UVA_PREC_MIXED,
              adjoint_tol: float = ADJOINT_TOL):
    """One refine global step (no line search); differentiable in U."""
    return _Step.apply(U, plan, float(tol), int(precision), float(adjoint_tol))


def arap_unrolled(plan: IslandPlan, U0: torch.Tensor, steps: int, tol: float = CG_TOL,
                  precision: int = _lib.UVA_PREC_MIXED, adjoint_tol: float = ADJOINT_TOL):
    """``steps`` global steps from U0; returns the last layout."""
    U = U0
    for _ in range(int(steps)):
        U = arap_step(plan, U, tol, precision, adjoint_tol)
    return U


def _vertex_island(plan: IslandPlan):
    idx = getattr(plan, "_vertex_island", None)
    if idx is None:
        counts = torch.as_tensor(plan.vert_off[1:] - plan.vert_off[:-1], device=plan.dev)
        idx = torch.repeat_interleave(torch.arange(plan.n_islands, device=plan.dev), counts)
        plan._vertex_island = idx
    return idx


@torch.no_grad()
def stretch_refine(plan: IslandPlan, U: torch.Tensor, w: torch.Tensor | None = None, steps: int = 200,
                   kind: int = 1, precondition: bool = True, pc_tol: float = 1e-2, armijo: float = 1e-4,
                   max_halvings: int = 30):
    """Gradient descent on the weighted energy, each island with its own step.

    Per step: g = dE/dU (backward kernel); D = -L^-1 g (one solve with the
    global step's operator to ``pc_tol``; -g without preconditioning);
    t = min(1, 0.9 t_max) with t_max the largest fold-free step along D;
    halve t until E(U + t D) <= E(U) + armijo t g.D.  Islands whose energy is
    not finite or that find no decrease keep their layout.  U is updated in
    place; returns (U, per-step energies as a (steps+1, n_islands) tensor)."""
    plan._check_uv(U)
    vi = _vertex_island(plan)
    ones = torch.ones(plan.n_islands, dtype=torch.float64, device=plan.dev)
    e = plan.energy_fwd(U, w, kind)
    hist = [e.clone()]
    for _ in range(int(steps)):
        g, _ = plan.energy_bwd(U, ones, w, kind, want_w=False)
        if precondition:
            D = plan.adjoint_cg(g, tol=pc_tol, precision=_lib.UVA_PREC_MIXED)
            D.neg_()
        else:
            D = -g
        slope = torch.zeros(plan.n_islands, dtype=torch.float64, device=plan.dev)
        slope.index_add_(0, vi, (g * D).sum(1))
        tmax = plan.max_step(U, D)
        t = torch.clamp(0.9 * tmax, max=1.0)
        active = torch.isfinite(e) & (slope < 0)
        t = torch.where(active, t, torch.zeros_like(t))
        for _h in range(max_halvings + 1):
            Un = U + t[vi].unsqueeze(1) * D
            en = plan.energy_fwd(Un.contiguous(), w, kind)
            ok = torch.isfinite(en) & (en <= e + armijo * t * slope)
            bad = active & ~ok
            if not bool(bad.any()):
                break
            t = torch.where(bad, 0.5 * t, t)
        accept = active & ok
        t = torch.where(accept, t, torch.zeros_like(t))
        U.add_(t[vi].unsqueeze(1) * D)
        e = torch.where(accept, en, e)
        hist.append(e.clone())
    return U, torch.stack(hist)
