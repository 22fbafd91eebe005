"""Differentiable FVM rollouts: hand-written adjoint behind torch.autograd.

``rollout(step_fn, steps, checkpoint_every=8)`` returns a function of the
initial velocity data ``(u0, v0)`` whose output ``(u_T, v_T)`` supports
``torch.autograd``: the forward pass runs the fused GPU step and keeps the
state every ``checkpoint_every`` steps; the backward pass recomputes each
segment from its checkpoint and applies the reverse step
(jcfd_fvm_step_vjp: lambda* = P lambda, lambda <- lambda* + dt J^T lambda*)
from the last step to the first.  This is the gradient jax.grad takes through
funcutils.repeated(step_fn, steps) (jax_cfd/base/funcutils.py:82-88) with
jax.remat-style checkpointing.  ``curl_2d`` is differentiable the same way.
"""
from __future__ import annotations

from typing import Callable, Tuple

import torch

from jax_cfd_amd.base import _fvm


class _Rollout(torch.autograd.Function):

  @staticmethod
  def forward(ctx, u0, v0, plan, steps, every):
    ckpts = []
    u, v = u0.detach().contiguous(), v0.detach().contiguous()
    for s0 in range(0, steps, every):
      ckpts += [u, v]
      u, v = plan.step(u, v, min(every, steps - s0))
    ctx.plan, ctx.steps, ctx.every = plan, steps, every
    ctx.save_for_backward(*ckpts)
    return u, v

  @staticmethod
  def backward(ctx, gu, gv):
    plan, steps, every = ctx.plan, ctx.steps, ctx.every
    saved = ctx.saved_tensors
    n = plan.spec.n
    dev = plan.device
    lu = gu.contiguous() if gu is not None else torch.zeros((n, n), device=dev)
    lv = gv.contiguous() if gv is not None else torch.zeros((n, n), device=dev)
    nseg = len(saved) // 2
    for seg in reversed(range(nseg)):
      s0 = seg * every
      k = min(every, steps - s0)
      states = [(saved[2 * seg], saved[2 * seg + 1])]
      for _ in range(k - 1):
        states.append(plan.step(*states[-1], 1))
      for u, v in reversed(states):
        lu, lv = plan.step_vjp(u, v, lu, lv)
    return lu, lv, None, None, None


def rollout(step_fn, steps: int, checkpoint_every: int = 8) -> Callable:
  """Differentiable `repeated(step_fn, steps)` on velocity data (u0, v0).

  `step_fn` is a step of equations.semi_implicit_navier_stokes with
  advect_van_leer convection and the fast-diagonal projection."""
  if checkpoint_every < 1:
    raise ValueError('checkpoint_every must be >= 1')
  spec = step_fn.spec

  def run(u0: torch.Tensor, v0: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    plan = step_fn.plan(u0.device)
    return _Rollout.apply(u0, v0, plan, int(steps), int(checkpoint_every))
  run.spec = spec
  return run


class _Curl(torch.autograd.Function):

  @staticmethod
  def forward(ctx, u, v, plan):
    ctx.plan = plan
    return plan.curl(u, v)

  @staticmethod
  def backward(ctx, wb):
    ub, vb = ctx.plan.curl_vjp(wb.contiguous())
    return ub, vb, None


def curl_2d(u: torch.Tensor, v: torch.Tensor, grid) -> torch.Tensor:
  """Differentiable finite_differences.curl_2d of staggered velocity data."""
  plan = _fvm.plan_for(_fvm.spec_from_grid(grid), u.device)
  return _Curl.apply(u, v, plan)
