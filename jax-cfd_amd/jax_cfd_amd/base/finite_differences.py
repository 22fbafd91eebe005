"""Finite differences on the GPU (jax_cfd/base/finite_differences.py).

Only the operators this path evaluates on device data are provided:
``divergence`` (136-143) and ``curl_2d`` (186-193).
"""
from __future__ import annotations

from jax_cfd_amd.base import _fvm, grids


def _plan(v):
  if len(v) != 2:
    raise NotImplementedError('built for 2-D velocity fields')
  grid = grids.consistent_grid(*[getattr(u, 'array', u) for u in v])
  return grid, _fvm.plan_for(_fvm.spec_from_grid(grid), v[0].data.device)


def divergence(v) -> grids.GridArray:
  """Backward-difference divergence of a staggered velocity."""
  grid, plan = _plan(v)
  out = plan.divergence(v[0].data, v[1].data)
  return grids.GridArray(out, grid.cell_center, grid)


def curl_2d(v) -> grids.GridArray:
  """Forward-difference vorticity dv/dx - du/dy, offset (1, 1)."""
  grid, plan = _plan(v)
  return grids.GridArray(plan.curl(v[0].data, v[1].data), (1.0, 1.0), grid)
