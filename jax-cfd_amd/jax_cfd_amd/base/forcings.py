"""Forcing functions (jax_cfd/base/forcings.py:63-178).

The constant parts are built on the host in fp32 at construction time, as
the reference builds them at trace time; the step kernels receive them as
the per-row-constant column f_x(y_j) plus a linear coefficient.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from jax_cfd_amd.base import grids


def _as_data(arr: np.ndarray, v):
  """Constant forcing data placed next to the velocity it will be added to."""
  data = getattr(v[0], 'data', None)
  if isinstance(data, torch.Tensor):
    return torch.as_tensor(arr, device=data.device)
  return arr


def kolmogorov_forcing(grid: grids.Grid, scale: float = 1, k: int = 2,
                       swap_xy: bool = False,
                       offsets: Optional[Tuple[Tuple[float, ...], ...]] = None):
  """forcings.py:63-104 (2-D)."""
  if grid.ndim != 2:
    raise NotImplementedError('kolmogorov_forcing is built for 2-D grids')
  if offsets is None:
    offsets = grid.cell_faces
  if swap_xy:
    x = grid.mesh(offsets[1])[0]
    fv = (scale * np.sin(k * x)).astype(np.float32)
    fu = np.zeros_like(fv)
    f_offsets = ((1, 1 / 2), offsets[1])
  else:
    y = grid.mesh(offsets[0])[1]
    fu = (scale * np.sin(k * y)).astype(np.float32)
    fv = np.zeros_like(fu)
    f_offsets = (offsets[0], (1 / 2, 1))

  def forcing(v):
    return (grids.GridArray(_as_data(fu, v), f_offsets[0], grid),
            grids.GridArray(_as_data(fv, v), f_offsets[1], grid))
  forcing.constant_fields = (fu, fv)
  forcing.offsets = f_offsets
  return forcing


def linear_forcing(grid, coefficient: float):
  """forcings.py:107-113."""
  del grid

  def forcing(v):
    return tuple(coefficient * u.array for u in v)
  return forcing


def no_forcing(grid):
  """forcings.py:116-122."""
  del grid

  def forcing(v):
    return tuple(0 * u.array for u in v)
  return forcing


def sum_forcings(*forcings):
  """forcings.py:125-129."""
  from jax_cfd_amd.base import equations

  def forcing(v):
    return equations.sum_fields(*[f(v) for f in forcings])
  return forcing


FORCING_FUNCTIONS = dict(kolmogorov=kolmogorov_forcing)


def simple_turbulence_forcing(grid: grids.Grid, constant_magnitude: float = 0,
                              constant_wavenumber: int = 2,
                              linear_coefficient: float = 0,
                              forcing_type: str = 'kolmogorov'):
  """forcings.py:136-178."""
  linear_force = linear_forcing(grid, linear_coefficient)
  constant_force_fn = FORCING_FUNCTIONS.get(forcing_type)
  if constant_force_fn is None:
    raise NotImplementedError(f'forcing_type {forcing_type!r} has no kernel here')
  constant_force = constant_force_fn(grid, constant_magnitude,
                                     constant_wavenumber)
  return sum_forcings(linear_force, constant_force)
