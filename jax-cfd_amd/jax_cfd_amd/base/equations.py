"""Navier-Stokes step factories (jax_cfd/base/equations.py).

``semi_implicit_navier_stokes`` and ``implicit_diffusion_navier_stokes``
keep the reference signatures (equations.py:120-151, 154-195) and returns a ``step_fn(v) -> v`` for the state
``(GridVariable(u), GridVariable(v))``.  The operator callables are read
once on probe velocities (see jax_cfd_amd._symbolic) and mapped onto one
fused GPU plan; combinations the kernels do not implement raise
NotImplementedError when the factory is called.
"""
from __future__ import annotations

import functools
from typing import Callable, Optional

import numpy as np
import torch

from jax_cfd_amd import _lib, _symbolic
from jax_cfd_amd.base import (_fvm, advection, boundaries, diffusion, grids,
                              pressure, time_stepping)


def sum_fields(*args):
  """Component-wise sum of tuples of fields (equations.py:41-42)."""
  return tuple(sum(a) for a in zip(*args))


def stable_time_step(max_velocity: float, max_courant_number: float,
                     viscosity: float, grid: grids.Grid,
                     implicit_diffusion: bool = False) -> float:
  """equations.py:45-59."""
  dt = advection.stable_time_step(max_velocity, max_courant_number, grid)
  if not implicit_diffusion:
    diffusion_dt = diffusion.stable_time_step(viscosity, grid)
    if diffusion_dt < dt:
      raise ValueError(f'stable time step for diffusion is smaller than '
                       f'the chosen timestep: {diffusion_dt} vs {dt}')
  return dt


def dynamic_time_step(v, max_courant_number: float, viscosity: float,
                      grid: grids.Grid, implicit_diffusion: bool = False) -> float:
  """equations.py:62-70 (the max is taken on the device, one scalar read)."""
  v_max = float(sum(u.data.float() ** 2 for u in v).max().sqrt())
  return stable_time_step(v_max, max_courant_number, viscosity, grid,
                          implicit_diffusion)


def _read_convection(convect, grid, dt):
  probe = _symbolic.probe_velocity(grid)
  if convect is None:
    convect = lambda v: tuple(  # noqa: E731  (the reference default, equations.py:89-92)
        advection.advect_van_leer_using_limiters(u, v, dt) for u in v)
  out = convect(probe)
  kinds = []
  dts = []
  for i, term in enumerate(out):
    t = getattr(term, 'data', None)
    if not isinstance(t, _symbolic.Term) or t.info.get('comp') != i:
      raise NotImplementedError(
          'convect must advect each velocity component by the velocity with '
          'advect_van_leer; got ' + repr(t))
    kinds.append(t.kind)
    dts.append(t.info['dt'])
  if len(out) != 2 or len(set(kinds)) != 1 or len(set(dts)) != 1:
    raise NotImplementedError('convect must treat both components alike')
  kinds_built = {'advect_van_leer': _lib.CONVECT_VAN_LEER,
                 'advect_van_leer_using_limiters': _lib.CONVECT_VAN_LEER_LIMITERS,
                 'advect_linear': _lib.CONVECT_LINEAR,
                 'advect_upwind': _lib.CONVECT_UPWIND}
  if kinds[0] not in kinds_built:
    raise NotImplementedError(f'convection {kinds[0]} is not built')
  # linear and upwind advection take no dt
  if kinds[0] not in ('advect_linear', 'advect_upwind') and dts[0] != dt:
    raise NotImplementedError('advection dt differing from the step dt')
  return kinds_built[kinds[0]]


def _read_forcing(forcing, grid):
  """Returns (linear_coefficient, forcing_x) or (None, None) for no forcing."""
  if forcing is None:
    return None, None
  probe = _symbolic.probe_velocity(grid)
  out = forcing(probe)
  if len(out) != 2:
    raise NotImplementedError('forcing must return two components')
  coefs, consts = [], []
  for i, f in enumerate(out):
    d = getattr(f, 'data', f)
    if isinstance(d, _symbolic.Affine):
      if d.coef != 0 and d.comp != i:
        raise NotImplementedError('forcing couples velocity components')
      coefs.append(d.coef)
      consts.append(d.const)
    elif isinstance(d, _symbolic.Term):
      raise NotImplementedError(f'forcing term {d.kind} is not built')
    else:  # constant field, no linear part
      coefs.append(0.0)
      consts.append(_symbolic.Affine._as_np(d))
    if tuple(getattr(f, 'offset', grid.cell_faces[i])) != grid.cell_faces[i]:
      raise NotImplementedError('forcing must live on the velocity offsets')
  if coefs[0] != coefs[1]:
    raise NotImplementedError('different linear coefficients per component')
  n = grid.shape[0]
  fx = np.zeros(n, np.float32)
  if consts[0] is not None:
    c0 = np.broadcast_to(np.asarray(consts[0], np.float32), grid.shape)
    if not np.array_equal(c0, np.broadcast_to(c0[:1, :], grid.shape)):
      raise NotImplementedError('constant x-forcing must depend on y only '
                                '(Kolmogorov form)')
    fx = np.ascontiguousarray(c0[0, :])
  if consts[1] is not None and np.any(np.asarray(consts[1]) != 0):
    raise NotImplementedError('constant y-forcing is not built')
  return coefs[0], tuple(float(x) for x in fx)


class FvmStepFn:
  """step_fn(v) -> v for one FVM instance; backed by a jcfd_fvm_plan."""

  def __init__(self, grid: grids.Grid, spec: _fvm.FvmSpec):
    self.grid = grid
    self.spec = spec
    self._plans = {}

  def plan(self, device=None) -> _fvm.FvmPlan:
    if device is None:
      device = torch.device('cuda', torch.cuda.current_device())
    key = device.index if device.index is not None else 0
    if key not in self._plans:
      self._plans[key] = _fvm.FvmPlan(self.spec, torch.device('cuda', key))
    return self._plans[key]

  def _checked(self, state):
    if len(state) != 2:
      raise ValueError('FVM state must be a tuple of two GridVariables')
    u, v = state
    for comp, off in zip(state, self.grid.cell_faces):
      if tuple(comp.offset) != tuple(off):
        raise ValueError(f'velocity offsets must be {self.grid.cell_faces}')
    if not boundaries.has_all_periodic_boundary_conditions(*state):
      raise NotImplementedError('non-periodic boundary conditions')
    data = u.data
    dev = data.device if getattr(data, 'is_cuda', False) else None
    return u, v, self.plan(dev)

  def _wrap(self, state, uo, vo):
    u, v = state
    return (grids.GridVariable(grids.GridArray(uo, u.offset, self.grid), u.bc),
            grids.GridVariable(grids.GridArray(vo, v.offset, self.grid), v.bc))

  def _run(self, state, nsteps: int):
    u, v, plan = self._checked(state)
    uo, vo = plan.step(u.data, v.data, nsteps)
    return self._wrap(state, uo, vo)

  def __call__(self, state):
    return self._run(state, 1)

  def multistep(self, state, steps: int):
    """funcutils.repeated(self, steps)(state) as one library call."""
    return self._run(state, int(steps))

  def rollout_frames(self, state, inner: int, count: int,
                     start_with_input: bool = False):
    """funcutils.trajectory(repeated(self, inner), count) on the device.

    Frames live in one (count, n, n) tensor per component; each library
    call reads the previous frame and its last kernel writes the next frame
    in place, so frames cost no copies.  Returns (final state, frames).
    """
    u, v, plan = self._checked(state)
    inner, count = int(inner), int(count)
    if inner < 0 or count < 0:
      raise ValueError('step and frame counts must be >= 0')
    uc, vc = plan._field(u.data), plan._field(v.data)
    fu = torch.empty((count,) + tuple(uc.shape), dtype=uc.dtype, device=uc.device)
    fv = torch.empty_like(fu)
    if count == 0:
      final = (uc.clone(), vc.clone())
    elif start_with_input:
      fu[0].copy_(uc)
      fv[0].copy_(vc)
      for t in range(1, count):
        plan.step(fu[t - 1], fv[t - 1], inner, out=(fu[t], fv[t]))
      final = plan.step(fu[count - 1], fv[count - 1], inner)
    else:
      prev = (uc, vc)
      for t in range(count):
        plan.step(prev[0], prev[1], inner, out=(fu[t], fv[t]))
        prev = (fu[t], fv[t])
      final = (fu[count - 1].clone(), fv[count - 1].clone())
    return self._wrap(state, *final), self._wrap(state, fu, fv)


def semi_implicit_navier_stokes(
    density: float,
    viscosity: Optional[float],
    dt: float,
    grid: grids.Grid,
    convect: Optional[Callable] = None,
    diffuse: Callable = diffusion.diffuse,
    pressure_solve: Callable = pressure.solve_fast_diag,
    forcing: Optional[Callable] = None,
    time_stepper: Callable = time_stepping.forward_euler,
) -> FvmStepFn:
  """Returns step_fn for explicit-diffusion Navier-Stokes (equations.py:120-151)."""
  tableau = time_stepping.tableau_of(time_stepper)
  if tableau is None:
    raise NotImplementedError(
        'time_stepper must be forward_euler, midpoint_rk2, heun_rk2, classic_rk4 or '
        'functools.partial(navier_stokes_rk, tableau)')
  rk = time_stepping.plan_fields(tableau)
  if diffuse is not diffusion.diffuse:
    raise NotImplementedError('only diffusion.diffuse (explicit) is built')
  if pressure_solve is pressure.solve_fast_diag:
    solve_kw = {}
  elif pressure.is_solve_cg(pressure_solve):
    solve_kw = pressure.cg_options(pressure_solve)
  else:
    raise NotImplementedError(f'pressure_solve {pressure_solve!r} is not built')
  conv = _read_convection(convect, grid, dt)
  lin, fx = _read_forcing(forcing, grid)
  if rk and solve_kw:
    raise NotImplementedError('Runge-Kutta tableaux are built with solve_fast_diag')
  spec = _fvm.spec_from_grid(
      grid, dt=float(dt), density=float(density),
      viscosity=None if viscosity is None else float(viscosity),
      convect=conv, linear_coefficient=lin, forcing_x=fx, **solve_kw, **rk)
  return FvmStepFn(grid, spec)


def implicit_diffusion_navier_stokes(
    density: float,
    viscosity: float,
    dt: float,
    grid: grids.Grid,
    convect: Optional[Callable] = None,
    diffusion_solve: Callable = diffusion.solve_fast_diag,
    pressure_solve: Callable = pressure.solve_fast_diag,
    forcing: Optional[Callable] = None,
) -> FvmStepFn:
  """Returns step_fn of Navier-Stokes with implicit diffusion
  (equations.py:154-195): convection (+ forcing / density), forward Euler,
  projection, then diffusion_solve(v, viscosity, dt)."""
  if diffusion.is_solve_fast_diag(diffusion_solve):
    diff_kw = dict(diffusion=_lib.DIFFUSION_IMPLICIT_FAST_DIAG)
  elif diffusion.is_solve_cg(diffusion_solve):
    diff_kw = dict(diffusion=_lib.DIFFUSION_IMPLICIT_CG, **diffusion.cg_spec_fields(
        **getattr(diffusion_solve, 'keywords', {})))
  else:
    raise NotImplementedError(f'diffusion_solve {diffusion_solve!r} is not built '
                              '(diffusion.solve_fast_diag and solve_cg are)')
  if pressure_solve is pressure.solve_fast_diag:
    solve_kw = {}
  elif pressure.is_solve_cg(pressure_solve):
    solve_kw = pressure.cg_options(pressure_solve)
  else:
    raise NotImplementedError(f'pressure_solve {pressure_solve!r} is not built')
  conv = _read_convection(convect, grid, dt)
  lin, fx = _read_forcing(forcing, grid)
  spec = _fvm.spec_from_grid(
      grid, dt=float(dt), density=float(density), viscosity=float(viscosity),
      convect=conv, linear_coefficient=lin, forcing_x=fx, **diff_kw, **solve_kw)
  return FvmStepFn(grid, spec)


def navier_stokes_explicit_terms(density, viscosity, dt, grid, convect=None,
                                 diffuse=diffusion.diffuse, forcing=None):
  """equations.py:77-116: returns v -> dv/dt (the explicit-terms seam)."""
  if diffuse is not diffusion.diffuse:
    raise NotImplementedError('only diffusion.diffuse is built')
  conv = _read_convection(convect, grid, dt)
  lin, fx = _read_forcing(forcing, grid)
  spec = _fvm.spec_from_grid(
      grid, dt=float(dt), density=float(density),
      viscosity=None if viscosity is None else float(viscosity),
      convect=conv, linear_coefficient=lin, forcing_x=fx)

  def explicit_terms(v):
    plan = _fvm.plan_for(spec, v[0].data.device)
    du, dv = plan.explicit_terms(v[0].data, v[1].data)
    return tuple(grids.GridVariable(grids.GridArray(d, u.offset, grid), u.bc)
                 for d, u in zip((du, dv), v))
  return explicit_terms
