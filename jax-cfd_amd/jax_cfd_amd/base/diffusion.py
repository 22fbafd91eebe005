"""Explicit diffusion (jax_cfd/base/diffusion.py:35-57)."""
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
