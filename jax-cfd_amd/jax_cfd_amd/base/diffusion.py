"""Diffusion (jax_cfd/base/diffusion.py:35-57, 170-212)."""
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


def is_solve_fast_diag(solve) -> bool:
  import functools
  if isinstance(solve, functools.partial):
    return solve.func is solve_fast_diag and solve.keywords.get(
        'implementation') in (None, 'rfft', 'fft') and not solve.args
  return solve is solve_fast_diag
