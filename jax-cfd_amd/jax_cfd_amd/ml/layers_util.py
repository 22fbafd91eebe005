"""Stencil constraint constants (jax_cfd/ml/layers_util.py).

Host-side construction of the polynomial-accuracy constraints and the
standard low-order coefficients that seed PolynomialConstraint; the
per-cell work that uses them runs in the GPU kernels.
"""
from __future__ import annotations

import enum
import functools
import itertools
import math
from typing import Iterator, Optional, Sequence, Tuple

import numpy as np


class Method(enum.Enum):
  """Discretization method (layers_util.py:18-21)."""
  FINITE_DIFFERENCE = 1
  FINITE_VOLUME = 2


def _exponents(degree: int, ndim: int) -> Iterator[Tuple[int, ...]]:
  if ndim == 0:
    yield ()
    return
  for d in range(degree + 1):
    for rest in _exponents(degree - d, ndim - 1):
      yield (d,) + rest


def _cell_moment(stencil: np.ndarray, m: int, method: Method, h: Optional[float]) -> np.ndarray:
  """x^m at the stencil points (finite differences) or its mean over the
  cell of width h centred there (finite volumes)."""
  if method is Method.FINITE_DIFFERENCE:
    return stencil ** m
  if method is not Method.FINITE_VOLUME:
    raise ValueError(f'unexpected method: {method}')
  if h is None:
    raise ValueError('grid_step is required for finite volumes')
  return ((stencil + h / 2) ** (m + 1) - (stencil - h / 2) ** (m + 1)) / (h * (m + 1))


def polynomial_accuracy_constraints(stencils: Sequence[np.ndarray], method: Method,
                                    derivative_orders: Sequence[int], accuracy_order: int,
                                    grid_step: Optional[float] = None):
  """(A, b) with A c = b for stencil coefficients c (row-major over the
  product stencil) exact on every monomial of total degree below
  accuracy_order + sum(derivative_orders) (layers_util.py:56-146); rows in
  sorted order of A's rows, duplicates merged."""
  if len(stencils) != len(derivative_orders):
    raise ValueError('mismatched lengths for stencils and derivative_orders')
  if accuracy_order < 1:
    raise ValueError(f'cannot compute constraints with non-positive accuracy_order: '
                     f'{accuracy_order}')
  rows = {}
  top = accuracy_order + sum(derivative_orders) - 1
  for exps in _exponents(top, len(stencils)):
    lhs = functools.reduce(np.kron, [_cell_moment(np.asarray(s, float), m, method, grid_step)
                                     for m, s in zip(exps, stencils)])
    rhs = math.prod(math.factorial(m) if m == d else 0
                    for m, d in zip(exps, derivative_orders))
    key = tuple(lhs)
    if key in rows and rows[key] != rhs:
      raise ValueError('conflicting constraints')
    rows[key] = rhs
  keys = sorted(rows)
  return np.array(keys), np.array([rows[k] for k in keys])


def _solve_1d(stencil: np.ndarray, method: Method, order: int, grid_step) -> np.ndarray:
  a, b = polynomial_accuracy_constraints([stencil], method, [order], stencil.size - order,
                                         grid_step)
  return np.linalg.solve(a, b)


def polynomial_accuracy_coefficients(stencils: Sequence[np.ndarray], method: Method,
                                     derivative_orders: Sequence[int],
                                     accuracy_order: Optional[int] = None,
                                     grid_step: Optional[float] = None) -> np.ndarray:
  """Outer product of per-axis highest-order coefficients on the centred
  sub-stencil the accuracy order needs (layers_util.py:176-227)."""
  out = np.zeros(tuple(len(s) for s in stencils))
  index, parts = [], []
  for s, order in zip(stencils, derivative_orders):
    s = np.asarray(s, float)
    trim = 0 if accuracy_order is None else (s.size - order - accuracy_order) // 2
    sl = slice(trim, s.size - trim)
    index.append(sl)
    parts.append(_solve_1d(s[sl], method, order, grid_step))
  out[tuple(index)] = functools.reduce(np.multiply.outer, parts)
  return out


def get_roll_and_shift(input_offset, target_offset):
  """Integer roll and fractional shift in [0, 1) of target - input
  (layers_util.py:230-237)."""
  delta = [t - i for t, i in zip(target_offset, input_offset)]
  roll = tuple(-math.floor(d) for d in delta)
  return roll, tuple(d + r for d, r in zip(delta, roll))


def get_stencils(stencil_sizes, offset, steps):
  """Cell-centre positions relative to the target point, (size - 1) // 2
  cells on the low side (layers_util.py:240-269)."""
  out = []
  for size, o, h in zip(stencil_sizes, offset, steps):
    lo = -((size - 1) // 2)
    out.append((np.arange(lo, lo + size) - o) * h)
  return tuple(out)


def patch_offsets(patch_shape) -> list:
  """Index offsets of extract_patches' channels (layers_util.py:324-341),
  row-major: per axis -size // 2 + 1 .. size // 2."""
  return list(itertools.product(*[range(-s // 2 + 1, s // 2 + 1) for s in patch_shape]))
