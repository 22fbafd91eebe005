"""IMEX time steppers for the spectral equations (jax_cfd/spectral/time_stepping.py).

``low_storage_runge_kutta_crank_nicolson`` of a NavierStokes2D -- with
``crank_nicolson_rk4``'s or ``crank_nicolson_rk3``'s coefficients or any
other tableau of 1 to 8 stages -- is fused on the GPU: one call of the
returned step advances the state (or a batch of states with a leading batch
axis) by whole steps.  ``crank_nicolson_rk2`` (Heun, not of the low-storage
form) raises NotImplementedError.
"""
from __future__ import annotations

from typing import Sequence

import torch

from jax_cfd_amd import _lib
from jax_cfd_amd.spectral import equations as spectral_equations

CNRK4_ALPHAS = (0, 0.1496590219993, 0.3704009573644, 0.6222557631345,
                0.9582821306748, 1)
CNRK4_BETAS = (0, -0.4178904745, -1.192151694643, -1.697784692471,
               -1.514183444257)
CNRK4_GAMMAS = (0.1496590219993, 0.3792103129999, 0.8229550293869,
                0.6994504559488, 0.1530572479681)
CNRK3_ALPHAS = (0, 1 / 3, 3 / 4, 1)
CNRK3_BETAS = (0, -5 / 9, -153 / 128)
CNRK3_GAMMAS = (1 / 3, 15 / 16, 8 / 15)


class ImplicitExplicitODE:
  """spectral/time_stepping.py:25-51 (interface only)."""

  def explicit_terms(self, state):
    raise NotImplementedError

  def implicit_terms(self, state):
    raise NotImplementedError

  def implicit_solve(self, state, step_size):
    raise NotImplementedError


class SpectralStepFn:
  """step_fn(w_hat) -> w_hat for low_storage_runge_kutta_crank_nicolson(
  tableau, NavierStokes2D, dt); `tableau` None is crank_nicolson_rk4."""

  def __init__(self, equation, dt: float, tableau=None):
    self.equation = equation
    self.dt = float(dt)
    self.tableau = tableau

  def plan(self, batch: int = 1, device=None):
    return self.equation.plan(self.dt, batch, device, self.tableau)

  def _run(self, w, nsteps):
    if not isinstance(w, torch.Tensor):
      w = torch.as_tensor(w)
    batch = 1 if w.ndim == 2 else int(w.shape[0])
    dev = w.device if w.is_cuda else None
    return self.plan(batch, dev).step(w, nsteps)

  def __call__(self, w):
    return self._run(w, 1)

  def multistep(self, w, steps: int):
    return self._run(w, int(steps))

  def rollout_frames(self, w, inner: int, count: int,
                     start_with_input: bool = False):
    """funcutils.trajectory(repeated(self, inner), count) on the device:
    frames in one (count,) + w.shape tensor, each written in place by the
    library call that produces it.  Returns (final state, frames)."""
    if not isinstance(w, torch.Tensor):
      w = torch.as_tensor(w)
    inner, count = int(inner), int(count)
    if inner < 0 or count < 0:
      raise ValueError('step and frame counts must be >= 0')
    batch = 1 if w.ndim == 2 else int(w.shape[0])
    plan = self.plan(batch, w.device if w.is_cuda else None)
    w = plan._field(w)
    frames = torch.empty((count,) + tuple(w.shape), dtype=w.dtype, device=w.device)
    if count == 0:
      return w.clone(), frames
    if start_with_input:
      frames[0].copy_(w)
      for t in range(1, count):
        plan.step(frames[t - 1], inner, out=frames[t])
      return plan.step(frames[count - 1], inner), frames
    prev = w
    for t in range(count):
      plan.step(prev, inner, out=frames[t])
      prev = frames[t]
    return frames[count - 1].clone(), frames


def low_storage_runge_kutta_crank_nicolson(alphas: Sequence[float], betas: Sequence[float],
                                           gammas: Sequence[float], equation,
                                           time_step: float):
  """spectral/time_stepping.py:117-170 for a NavierStokes2D, fused on the GPU:
  any tableau of 1 to 8 stages (k + 1 alphas, k betas, k gammas)."""
  if len(alphas) - 1 != len(betas) != len(gammas):
    raise ValueError('number of RK coefficients does not match')
  if len(alphas) - 1 != len(betas) or len(betas) != len(gammas):
    raise NotImplementedError('the GPU steppers need k + 1 alphas, k betas and '
                              'k gammas')
  if not 1 <= len(betas) <= _lib.MAX_RK_STAGES:
    raise NotImplementedError(
        f'tableaux of 1 to {_lib.MAX_RK_STAGES} stages are built')
  if not isinstance(equation, spectral_equations.NavierStokes2D):
    raise NotImplementedError('only NavierStokes2D is fused on the GPU')
  tableau = tuple(tuple(float(x) for x in c) for c in (alphas, betas, gammas))
  if tableau == (CNRK4_ALPHAS, CNRK4_BETAS, CNRK4_GAMMAS):
    tableau = None  # the plan's built-in crank_nicolson_rk4 coefficients
  return SpectralStepFn(equation, time_step, tableau)


def crank_nicolson_rk4(equation, time_step: float) -> SpectralStepFn:
  """spectral/time_stepping.py:186-197."""
  return low_storage_runge_kutta_crank_nicolson(
      CNRK4_ALPHAS, CNRK4_BETAS, CNRK4_GAMMAS, equation, time_step)


def crank_nicolson_rk3(equation, time_step: float) -> SpectralStepFn:
  """spectral/time_stepping.py:173-183 (Williamson's coefficients)."""
  return low_storage_runge_kutta_crank_nicolson(
      CNRK3_ALPHAS, CNRK3_BETAS, CNRK3_GAMMAS, equation, time_step)


def crank_nicolson_rk2(equation, time_step):
  raise NotImplementedError('crank_nicolson_rk2 is not built')
