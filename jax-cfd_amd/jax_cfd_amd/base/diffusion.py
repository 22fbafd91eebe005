"""Diffusion (jax_cfd/base/diffusion.py:35-57, 137-212)."""
from __future__ import annotations

from jax_cfd_amd import _symbolic
from jax_cfd_amd.base import grids


def diffuse(c: grids.GridVariable, nu: float) -> grids.GridArray:
  """nu * laplacian(c) (diffusion.py:35-37).  Fused into the FVM step kernel;
  read symbolically by the step factory."""
  if _symbolic.is_probe(c):
    return grids.GridArray(_symbolic.Term('diffuse', comp=c.data.comp, nu=nu),
                           c.offset, c.grid)
  raise NotImplementedError(
      'diffuse runs fused inside the FVM step kernel; it has no stand-alone '
      'GPU kernel')


def stable_time_step(viscosity: float, grid: grids.Grid) -> float:
  """diffusion.py:40-57."""
  if viscosity == 0:
    return float('inf')
  dx = min(grid.step)
  return dx ** 2 / (viscosity * 2 ** grid.ndim)


def solve_fast_diag(v, nu: float, dt: float, implementation=None):
  """diffusion.py:170-212 (periodic): (1 - nu dt L)^-1 applied to each
  velocity component by fast diagonalisation, on the GPU.  Also the marker
  implicit_diffusion_navier_stokes reads as its `diffusion_solve`."""
  from jax_cfd_amd.base import _fvm, boundaries
  if implementation not in (None, 'rfft', 'fft'):
    raise NotImplementedError(f'implementation={implementation!r}')
  if len(v) != 2 or not boundaries.has_all_periodic_boundary_conditions(*v):
    raise NotImplementedError('the diffusion kernels are built for periodic 2-D velocity')
  grid = grids.consistent_grid(*[u.array for u in v])
  from jax_cfd_amd import _lib
  spec = _fvm.spec_from_grid(grid, dt=float(dt), viscosity=float(nu),
                             diffusion=_lib.DIFFUSION_IMPLICIT_FAST_DIAG)
  plan = _fvm.plan_for(spec, v[0].data.device)
  uo, vo = plan.diffuse(v[0].data, v[1].data)
  return tuple(grids.GridVariable(grids.GridArray(d, u.offset, grid), u.bc)
               for d, u in zip((uo, vo), v))


def solve_cg(v, nu: float, dt: float, rtol: float = 1e-6, atol: float = 1e-6,
             maxiter=None):
  """diffusion.py:137-167: each velocity component u becomes the CG solution
  of (1 - nu dt L) x = u from x0 = u, with jax's cg tolerances (stop once
  <r, r> <= max(rtol^2 <u, u>, atol^2) or after maxiter iterations, default
  10 n^2), on the GPU.  Also a marker implicit_diffusion_navier_stokes reads
  as its `diffusion_solve` (with functools.partial keywords)."""
  from jax_cfd_amd import _lib
  from jax_cfd_amd.base import _fvm, boundaries
  if not boundaries.has_all_periodic_boundary_conditions(*v):
    raise ValueError('solve_cg() expects periodic BC')
  if len(v) != 2:
    raise NotImplementedError('the diffusion kernels are built for 2-D velocity')
  grid = grids.consistent_grid(*[u.array for u in v])
  spec = _fvm.spec_from_grid(grid, dt=float(dt), viscosity=float(nu),
                             diffusion=_lib.DIFFUSION_IMPLICIT_CG,
                             **cg_spec_fields(rtol=rtol, atol=atol, maxiter=maxiter))
  plan = _fvm.plan_for(spec, v[0].data.device)
  uo, vo = plan.diffuse(v[0].data, v[1].data)
  return tuple(grids.GridVariable(grids.GridArray(d, u.offset, grid), u.bc)
               for d, u in zip((uo, vo), v))


def cg_spec_fields(rtol: float = 1e-6, atol: float = 1e-6, maxiter=None) -> dict:
  """solve_cg's keywords as FvmSpec fields (maxiter None: 0, read as 10 n^2)."""
  return dict(diff_cg_rtol=float(rtol), diff_cg_atol=float(atol),
              diff_cg_maxiter=0 if maxiter is None else int(maxiter))


def is_solve_cg(solve) -> bool:
  import functools
  if isinstance(solve, functools.partial):
    return solve.func is solve_cg and not solve.args and set(solve.keywords) <= {
        'rtol', 'atol', 'maxiter'}
  return solve is solve_cg


def is_solve_fast_diag(solve) -> bool:
  import functools
  if isinstance(solve, functools.partial):
    return solve.func is solve_fast_diag and solve.keywords.get(
        'implementation') in (None, 'rfft', 'fft') and not solve.args
  return solve is solve_fast_diag
