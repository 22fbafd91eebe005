"""Pressure projection (jax_cfd/base/pressure.py:67-198).

``projection`` runs the fast-diagonalisation (rfft) or the CG solve on the
GPU.  ``solve_fast_diag`` called directly returns q; ``solve_cg`` is a marker
read by ``projection`` and the step factory.
"""
from __future__ import annotations

from typing import Callable, Optional

from jax_cfd_amd import _lib
from jax_cfd_amd.base import _fvm, boundaries, grids


def _check_velocity(v):
  if len(v) != 2:
    raise NotImplementedError('the projection kernels are built for 2-D velocity')
  if not boundaries.has_all_periodic_boundary_conditions(*v):
    raise NotImplementedError('non-periodic boundary conditions')
  return grids.consistent_grid(*[u.array for u in v])


def solve_fast_diag(v, q0=None, pressure_bc=None, implementation: Optional[str] = None):
  """pressure.py:115-157 (periodic): q = irfftn(Lambda^+ rfftn(div v))."""
  del q0, pressure_bc
  if implementation not in (None, 'rfft', 'fft'):
    raise NotImplementedError(f'implementation={implementation!r}')
  grid = _check_velocity(v)
  plan = _fvm.plan_for(_fvm.spec_from_grid(grid), v[0].data.device)
  # q = v - P(v) componentwise is a gradient; recover q through the filter
  # with the pseudo-inverse multiplier on div v.
  div = plan.divergence(v[0].data, v[1].data)
  lam = _poisson_multiplier(grid, div.device)
  q = plan.spectral_filter(div, lam)
  return grids.GridArray(q, grid.cell_center, grid)


def _poisson_multiplier(grid, device):
  import numpy as np
  import torch
  n = grid.shape[0]
  dx, dy = grid.step
  p = np.arange(n)
  c = np.arange(n // 2 + 1)
  lx = (-2.0 / dx ** 2) + 2.0 * np.cos(2 * np.pi * p / n) / dx ** 2
  ly = (-2.0 / dy ** 2) + 2.0 * np.cos(2 * np.pi * c / n) / dy ** 2
  s = lx[:, None] + ly[None, :]
  cutoff = float(10 * np.finfo(np.float32).eps)
  with np.errstate(divide='ignore'):
    lam = np.where(np.abs(s) > cutoff, 1.0 / s, 0.0)
  return torch.as_tensor(lam.astype(np.float32), device=device)


def solve_cg(v, q0=None, pressure_bc=None, rtol: float = 1e-6, atol: float = 1e-6,
             maxiter: Optional[int] = None):
  """pressure.py:67-112: unpreconditioned CG (x0 = 0) on the periodic
  Laplacian.  A marker: projection(v, solve_cg) and the step factory run it
  on the GPU (the solution q itself is not returned on its own)."""
  del v, q0, pressure_bc, rtol, atol, maxiter
  raise NotImplementedError('use projection(v, solve_cg) or pass solve_cg to '
                            'semi_implicit_navier_stokes')


def cg_options(solve, rtol=1e-6, atol=1e-6, maxiter=None) -> dict:
  """Plan fields for a solve_cg callable (possibly functools.partial)."""
  import functools
  if isinstance(solve, functools.partial):
    kw = dict(solve.keywords)
    rtol = kw.get('rtol', rtol)
    atol = kw.get('atol', atol)
    maxiter = kw.get('maxiter', maxiter)
  return dict(pressure=_lib.PRESSURE_CG, cg_rtol=float(rtol), cg_atol=float(atol),
              cg_maxiter=int(maxiter or 0))


def is_solve_cg(solve) -> bool:
  import functools
  return solve is solve_cg or (isinstance(solve, functools.partial)
                               and solve.func is solve_cg)


def projection(v: grids.GridVariableVector,
               solve: Callable = solve_fast_diag) -> grids.GridVariableVector:
  """pressure.py:181-198: v - grad q with q from `solve` (fast-diagonal or
  CG, both on the GPU)."""
  grid = _check_velocity(v)
  if solve is solve_fast_diag:
    spec = _fvm.spec_from_grid(grid)
  elif is_solve_cg(solve):
    spec = _fvm.spec_from_grid(grid, **cg_options(solve))
  else:
    raise NotImplementedError('projection on the GPU uses solve_fast_diag or solve_cg')
  plan = _fvm.plan_for(spec, v[0].data.device)
  uo, vo = plan.projection(v[0].data, v[1].data)
  return tuple(grids.GridVariable(grids.GridArray(d, u.offset, grid), u.bc)
               for d, u in zip((uo, vo), v))
