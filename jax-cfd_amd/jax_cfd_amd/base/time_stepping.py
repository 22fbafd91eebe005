"""Explicit Navier-Stokes time steppers (jax_cfd/base/time_stepping.py).

``forward_euler``, ``midpoint_rk2``, ``heun_rk2``, ``classic_rk4`` and
``functools.partial(navier_stokes_rk, tableau)`` are markers the step factory
(equations.semi_implicit_navier_stokes) maps onto GPU plans: forward Euler
onto the fused step kernels, other tableaux (up to 4 stages) onto the
Runge-Kutta stage sequence (explicit terms, combination + projection).
"""
from __future__ import annotations

import dataclasses
import functools
from typing import Callable, Optional, Sequence, Tuple, TypeVar

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
  """time_stepping.py:59-106.  Runs on the GPU when handed to
  semi_implicit_navier_stokes as functools.partial(navier_stokes_rk, tableau)."""
  del tableau, equation, time_step
  raise NotImplementedError(
      'pass functools.partial(navier_stokes_rk, tableau) as time_stepper to '
      'semi_implicit_navier_stokes; the GPU plan implements the whole step')


def _marker(name):
  def stepper(equation, time_step):
    raise NotImplementedError(
        f'pass {name} as time_stepper to semi_implicit_navier_stokes; the GPU '
        'plan implements the whole step')
  stepper.__name__ = name
  stepper.__doc__ = f'time_stepping.py:109-158 ({name}).  Marker for the step factory.'
  return stepper


forward_euler = _marker('forward_euler')
midpoint_rk2 = _marker('midpoint_rk2')
heun_rk2 = _marker('heun_rk2')
classic_rk4 = _marker('classic_rk4')

# (a, b) of the named steppers (time_stepping.py:109-158).
TABLEAUS = {
    forward_euler: ButcherTableau(a=[], b=[1]),
    midpoint_rk2: ButcherTableau(a=[[1 / 2]], b=[0, 1]),
    heun_rk2: ButcherTableau(a=[[1]], b=[1 / 2, 1 / 2]),
    classic_rk4: ButcherTableau(a=[[1 / 2], [0, 1 / 2], [0, 0, 1]],
                                b=[1 / 6, 1 / 3, 1 / 3, 1 / 6]),
}


def tableau_of(time_stepper) -> Optional[ButcherTableau]:
  """The tableau a time_stepper argument stands for, or None."""
  if time_stepper in TABLEAUS:
    return TABLEAUS[time_stepper]
  if (isinstance(time_stepper, functools.partial) and time_stepper.func is navier_stokes_rk
      and len(time_stepper.args) == 1 and not time_stepper.keywords):
    return time_stepper.args[0]
  return None


def plan_fields(tableau: ButcherTableau) -> dict:
  """FvmSpec fields for a tableau (forward Euler -> the fused kernels)."""
  stages = len(tableau.b)
  if stages < 1 or stages > 4:
    raise NotImplementedError('Runge-Kutta tableaux of 1 to 4 stages are built')
  for i, row in enumerate(tableau.a):
    if len(row) > i + 1 or any(row[j] for j in range(len(row)) if j > i):
      raise NotImplementedError('only explicit (lower-triangular) tableaux are built')
  if stages == 1 and tableau.b[0] == 1:
    return {}
  a: Tuple[float, ...] = tuple(float(row[j]) if j < len(row) else 0.0
                               for i, row in enumerate(tableau.a) for j in range(i + 1))
  return dict(rk_stages=stages, rk_a=a, rk_b=tuple(float(x) for x in tableau.b))
