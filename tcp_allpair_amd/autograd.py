"""torch.autograd bindings of the energy derivatives (tcp_energy_adjoint).

No counterpart in the reference.  The GPU context computes, in FP64 with
per-particle charges q and regularisation length lambda,

    U = eps (Ur + Uo) + Um,  dU/dx = -F,  dU/dq,  dU/dlambda,  dF/dlambda,

and two autograd Functions expose them:

* energy(context, x, q, lam) -> U, differentiable in x, q and lam;
* forces(context, x, q, lam) -> F, differentiable in lam (dF/dlambda), which
  is what force matching needs.

fit_lambda() fits lambda by force matching over a batch of snapshots with
torch.optim.LBFGS; each snapshot's loss runs under activation checkpointing,
so only one snapshot's dF/dlambda lives at a time and the backward pass
re-runs that snapshot's forward.

Tensors may live on any device; the C ABI takes host buffers, so values pass
through host memory (the positions of one snapshot per call).
"""
from __future__ import annotations

import math

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from .gpu import Context


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


def _prepare(context: Context, x: torch.Tensor, q: torch.Tensor, lam) -> np.ndarray:
    n = context.p.n
    if tuple(x.shape) != (n, 3) or tuple(q.shape) != (n,):
        raise ValueError(f"expected x of shape ({n}, 3) and q of shape ({n},)")
    context.set_lambda(float(lam))
    context.set_state(x=_host(x))
    return _host(q)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, q, lam, context):
        U, dudx, dudq, dudl, _ = context.energy_adjoint(_prepare(context, x, q, lam),
                                                        with_dfdl=False)
        dev = x.device
        ctx.save_for_backward(torch.from_numpy(dudx).to(dev), torch.from_numpy(dudq).to(dev),
                              torch.tensor(dudl, dtype=torch.float64, device=dev))
        return torch.tensor(U, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g):
        dudx, dudq, dudl = ctx.saved_tensors
        return g * dudx, g * dudq, g * dudl, None


class _Forces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lam, x, q, context):
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            raise RuntimeError("forces() is differentiable in lambda only "
                               "(dF/dx and dF/dq need second derivatives)")
        _, dudx, _, _, dfdl = context.energy_adjoint(_prepare(context, x, q, lam), with_dfdl=True)
        dev = x.device
        ctx.save_for_backward(torch.from_numpy(dfdl).to(dev))
        return torch.from_numpy(-dudx).to(dev)

    @staticmethod
    def backward(ctx, gF):
        (dfdl,) = ctx.saved_tensors
        return (gF * dfdl).sum(), None, None, None


def energy(context: Context, x: torch.Tensor, q: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
    """Total potential energy U (eV) with gradients dU/dx, dU/dq, dU/dlambda."""
    return _Energy.apply(x, q, lam, context)


def forces(context: Context, x: torch.Tensor, q: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
    """Forces F = -dU/dx (eV/A), shape (N, 3), with the gradient dF/dlambda."""
    return _Forces.apply(lam, x, q, context)


def species_charges(p) -> torch.Tensor:
    return torch.tensor(np.where(np.arange(p.n) < p.ne, p.q[0], p.q[1]), dtype=torch.float64)


def force_matching_loss(context, snapshots, targets, q, lam, use_checkpoint=True):
    """sum_s |F_s(lam) - F*_s|^2 / sum_s |F*_s|^2 over the snapshot batch."""
    norm = sum(float((t ** 2).sum()) for t in targets)

    def one(lam_, x, t):
        return ((forces(context, x, q, lam_) - t) ** 2).sum()

    loss = lam.new_zeros(())
    for x, t in zip(snapshots, targets):
        if use_checkpoint:
            loss = loss + checkpoint(one, lam, x, t, use_reentrant=False)
        else:
            loss = loss + one(lam, x, t)
    return loss / norm


def fit_lambda(context: Context, snapshots, targets, lam0: float, q=None, max_iter: int = 30,
               use_checkpoint: bool = True):
    """Fit lambda to target forces with LBFGS over theta = log(lambda).

    Returns (lambda, history) where history lists (lambda, loss) per evaluation.
    """
    q = species_charges(context.p) if q is None else q
    theta = torch.tensor(math.log(lam0), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([theta], lr=1.0, max_iter=max_iter, tolerance_grad=1e-14,
                            tolerance_change=1e-16, line_search_fn="strong_wolfe")
    history = []

    def closure():
        opt.zero_grad()
        loss = force_matching_loss(context, snapshots, targets, q, theta.exp(), use_checkpoint)
        loss.backward()
        history.append((float(theta.detach().exp()), float(loss.detach())))
        return loss

    opt.step(closure)
    return float(theta.detach().exp()), history
