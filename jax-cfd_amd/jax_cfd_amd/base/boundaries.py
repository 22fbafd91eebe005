"""Periodic boundary conditions (jax_cfd/base/boundaries.py:44-559).

Only the periodic subset is on this path: every configuration the kernels
serve is periodic.  Other boundary types are accepted as descriptions but
any operator asked to use them raises NotImplementedError.
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Sequence, Tuple

import torch

from jax_cfd_amd.base import grids


class BCType:
  PERIODIC = 'periodic'
  DIRICHLET = 'dirichlet'
  NEUMANN = 'neumann'


class Padding:
  MIRROR = 'mirror'
  EXTEND = 'extend'


@dataclasses.dataclass(init=False, frozen=True)
class ConstantBoundaryConditions:
  """Boundary types and values per axis side (boundaries.py:44-80)."""
  types: Tuple[Tuple[str, str], ...]
  bc_values: Tuple[Tuple[Optional[float], Optional[float]], ...]

  def __init__(self, types: Sequence[Tuple[str, str]],
               values: Sequence[Tuple[Optional[float], Optional[float]]]):
    object.__setattr__(self, 'types', tuple(tuple(t) for t in types))
    object.__setattr__(self, 'bc_values', tuple(tuple(v) for v in values))

  def _require_periodic(self, axis: int):
    if self.types[axis] != (BCType.PERIODIC, BCType.PERIODIC):
      raise NotImplementedError(
          f'only periodic boundaries are supported, got {self.types[axis]}')

  def shift(self, u: grids.GridArray, offset: int, axis: int,
            mode: Optional[str] = Padding.EXTEND) -> grids.GridArray:
    """Periodic shift: out[i] = u[i + offset] (boundaries.py:81-104)."""
    del mode
    self._require_periodic(axis)
    new_offset = tuple(o + offset if a == axis else o
                       for a, o in enumerate(u.offset))
    return grids.GridArray(torch.roll(u.data, -offset, dims=axis), new_offset,
                           u.grid)

  def impose_bc(self, u: grids.GridArray) -> grids.GridVariable:
    return grids.GridVariable(u, self)

  def trim_boundary(self, u: grids.GridArray) -> grids.GridArray:
    return u


BoundaryConditions = ConstantBoundaryConditions


def periodic_boundary_conditions(ndim: int) -> ConstantBoundaryConditions:
  """boundaries.py:556-559."""
  return ConstantBoundaryConditions(
      ((BCType.PERIODIC, BCType.PERIODIC),) * ndim, ((None, None),) * ndim)


def has_all_periodic_boundary_conditions(*arrays) -> bool:
  for array in arrays:
    for lower, upper in array.bc.types:
      if lower != BCType.PERIODIC or upper != BCType.PERIODIC:
        return False
  return True


def get_pressure_bc_from_velocity(v) -> ConstantBoundaryConditions:
  """boundaries.py:724-735 (periodic case)."""
  if not has_all_periodic_boundary_conditions(*v):
    raise NotImplementedError('non-periodic velocity boundary conditions')
  return periodic_boundary_conditions(len(v))
