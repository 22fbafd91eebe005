"""Explicit Navier-Stokes time steppers (jax_cfd/base/time_stepping.py).

The step factory maps ``forward_euler`` onto the fused FVM kernels; the other
Butcher tableaux are recognised and reported as not built.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Sequence, TypeVar

PyTreeState = TypeVar('PyTreeState')
TimeStepFn = Callable[[PyTreeState], PyTreeState]


class ExplicitNavierStokesODE:
  """time_stepping.py:26-45."""

  def __init__(self, explicit_terms, pressure_projection):
    self.explicit_terms = explicit_terms
    self.pressure_projection = pressure_projection


@dataclasses.dataclass
class ButcherTableau:
  """time_stepping.py:48-56."""
  a: Sequence[Sequence[float]]
  b: Sequence[float]

  def __post_init__(self):
    if len(self.a) + 1 != len(self.b):
      raise ValueError('inconsistent Butcher tableau')


def navier_stokes_rk(tableau: ButcherTableau, equation, time_step: float):
  """time_stepping.py:59-106 -- only the b = [1] tableau (forward Euler) is
  fused on the GPU; build steps through equations.semi_implicit_navier_stokes."""
  raise NotImplementedError(
      'generic Runge-Kutta tableaux are not built; use forward_euler through '
      'semi_implicit_navier_stokes')


def forward_euler(equation, time_step: float):
  """time_stepping.py:109-118.  Marker for the step factory."""
  raise NotImplementedError(
      'call semi_implicit_navier_stokes(..., time_stepper=forward_euler); the '
      'fused kernels implement the whole composition')


def midpoint_rk2(equation, time_step):
  raise NotImplementedError('midpoint_rk2 is not built')


def heun_rk2(equation, time_step):
  raise NotImplementedError('heun_rk2 is not built')


def classic_rk4(equation, time_step):
  raise NotImplementedError('classic_rk4 is not built')
