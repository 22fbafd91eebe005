"""Symbolic stand-ins used to read user-supplied operator callables.

The reference composes a step from Python callables (convect, forcing,
pressure_solve, time_stepper; jax_cfd/base/equations.py:77-151) and lets
JAX trace them.  This package maps those callables onto fused kernels, so a
factory calls them once on *probe* velocities: the operators of this package
recognise probes and return descriptions of themselves instead of computing.
Anything that does not reduce to a form the kernels implement raises
NotImplementedError -- there is no silent fallback.
"""
from __future__ import annotations

import numbers

import numpy as np
import torch


class Affine:
  """coef * velocity[comp] + const: the form of the forcings on this path."""

  __array_ufunc__ = None  # make NumPy defer to the reflected operators

  def __init__(self, comp: int, coef: float = 1.0, const=None):
    self.comp = comp
    self.coef = float(coef)
    self.const = const  # None or an array (numpy or torch, host or device)

  @staticmethod
  def _as_np(x):
    if isinstance(x, torch.Tensor):
      return x.detach().cpu().numpy()
    return np.asarray(x)

  def _with(self, coef, const):
    return Affine(self.comp, coef, const)

  def __add__(self, other):
    if isinstance(other, Affine):
      if other.comp != self.comp and other.coef != 0 and self.coef != 0:
        raise NotImplementedError('forcing couples velocity components')
      comp = self.comp if self.coef != 0 else other.comp
      const = _add_const(self.const, other.const)
      return Affine(comp, self.coef + other.coef, const)
    if isinstance(other, numbers.Number):
      if other == 0:
        return self
      return self._with(self.coef, _add_const(self.const, float(other)))
    if isinstance(other, (np.ndarray, torch.Tensor)):
      return self._with(self.coef, _add_const(self.const, self._as_np(other)))
    return NotImplemented

  __radd__ = __add__

  def __mul__(self, other):
    if isinstance(other, numbers.Number):
      c = None if self.const is None else self.const * float(other)
      return self._with(self.coef * float(other), c)
    return NotImplemented

  __rmul__ = __mul__

  def __truediv__(self, other):
    if isinstance(other, numbers.Number):
      return self * (1.0 / float(other))
    return NotImplemented

  def __neg__(self):
    return self * -1.0

  def __sub__(self, other):
    return self + (-other if not isinstance(other, (np.ndarray, torch.Tensor))
                   else -self._as_np(other))

  def __rsub__(self, other):
    return (-self) + other


def _add_const(a, b):
  if a is None:
    return b if not isinstance(b, torch.Tensor) else b.detach().cpu().numpy()
  if b is None:
    return a
  if isinstance(b, torch.Tensor):
    b = b.detach().cpu().numpy()
  return np.asarray(a) + b


class Term:
  """Result of an operator of this package applied to probes."""

  def __init__(self, kind: str, **info):
    self.kind = kind
    self.info = info

  def __repr__(self):
    return f'Term({self.kind}, {self.info})'


def is_probe(x) -> bool:
  data = getattr(x, 'data', x)
  return isinstance(data, (Affine, Term))


def probe_velocity(grid):
  """Tuple of GridVariables whose data are probes of components 0..ndim-1."""
  from jax_cfd_amd.base import boundaries, grids
  bc = boundaries.periodic_boundary_conditions(grid.ndim)
  return tuple(grids.GridVariable(grids.GridArray(Affine(i), o, grid), bc)
               for i, o in enumerate(grid.cell_faces))
