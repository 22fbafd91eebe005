"""Learned interpolation (FusedLearnedInterpolation) restated in NumPy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows the reference's ml/ package:
  * layers_util.py:56-146  polynomial_accuracy_constraints
  * layers_util.py:149-227 _high_order_coefficients_1d, polynomial_accuracy_coefficients
  * layers_util.py:230-269 get_roll_and_shift, get_stencils
  * layers_util.py:324-446 extract_patches (the 'roll' method), apply_coefficients
  * layers.py:451-544      PolynomialConstraint (bias + SVD nullspace)
  * layers.py:586-642      SpatialDerivativeFromLogits
  * layers.py:24-116 + haiku Conv2D: PeriodicConv2D (wrap padding, VALID
                           cross-correlation, weights (kh, kw, cin, cout), bias)
  * towers.py:144-262      forward_tower_factory / forward_flex_tower_factory
                           (conv, relu, ..., conv; identity scale functions)
  * interpolations.py:31-121 FusedLearnedInterpolation (one tower over the
                           stacked velocity components; logits split per
                           (input offset, target offset, tag) in insertion order)
  * advections.py:48-85, base/advection.py:34-116 modular_self_advection /
                           self_advection / advect_general / _advect_aligned
  * ml/equations.py:92-146 modular_navier_stokes_model (the convection module
                           is built from the step's input velocity)

Pinning (tests/test_oracle_ml.py): the reference's own known-answer tests
(layers_util_test.py: constraint matrices, finite-difference and
finite-volume coefficients, extract_patches; layers_test.py: constrained
outputs satisfy the constraints, periodic-convolution equivariance), a
brute-force per-pixel periodic cross-correlation, delta kernels that reduce
the convolution to a roll, zero logits reducing every interpolation to the
pinned multilinear interpolation and the whole advection to advect_linear
exactly, and classical stencils (4-point finite-volume face value
(-1, 7, 7, -1)/12) recovered from the constraint space.
"""
from __future__ import annotations

import collections
import dataclasses
import enum
import functools
import itertools
import math
from typing import Dict, Iterator, Optional, Sequence, Tuple

import numpy as np

from oracle import fvm as ofvm
from oracle import grid_ops as go
from oracle.grid_ops import A, Grid


class Method(enum.Enum):
  FINITE_DIFFERENCE = 1
  FINITE_VOLUME = 2


def _kron(arrays):
  return functools.reduce(np.kron, arrays)


def exponents_up_to_degree(degree: int, ndim: int) -> Iterator[Tuple[int, ...]]:
  """layers_util.py:29-53: exponent tuples of total degree <= degree, the
  first axis' exponent outermost."""
  if ndim == 0:
    yield tuple()
  else:
    for d in range(degree + 1):
      for rest in exponents_up_to_degree(degree - d, ndim - 1):
        yield (d,) + rest


def polynomial_accuracy_constraints(stencils, method, derivative_orders, accuracy_order,
                                    grid_step=None):
  """layers_util.py:56-146: A c = b for coefficients c (row-major over the
  stencil product) exact on monomials of degree < accuracy_order +
  sum(derivative_orders); rows sorted by their left-hand side."""
  if len(stencils) != len(derivative_orders):
    raise ValueError('mismatched lengths for stencils and derivative_orders')
  if accuracy_order < 1:
    raise ValueError('accuracy_order must be positive')
  constraints = {}
  max_degree = accuracy_order + sum(derivative_orders) - 1
  for exponents in exponents_up_to_degree(max_degree, len(stencils)):
    lhs_terms, rhs_terms = [], []
    for m, stencil, order in zip(exponents, stencils, derivative_orders):
      if method is Method.FINITE_VOLUME:
        if grid_step is None:
          raise ValueError('grid_step is required for finite volumes')
        lhs_terms.append(1 / grid_step * ((stencil + grid_step / 2) ** (m + 1)
                                          - (stencil - grid_step / 2) ** (m + 1)) / (m + 1))
      elif method is Method.FINITE_DIFFERENCE:
        lhs_terms.append(stencil ** m)
      else:
        raise ValueError(method)
      rhs_terms.append(math.factorial(m) if m == order else 0)
    lhs = tuple(_kron(lhs_terms))
    rhs = np.prod(rhs_terms)
    if lhs in constraints and constraints[lhs] != rhs:
      raise ValueError('conflicting constraints')
    constraints[lhs] = rhs
  rows, rhs = zip(*sorted(constraints.items()))
  return np.array(rows), np.array(rhs)


def _high_order_coefficients_1d(stencil, method, derivative_order, grid_step=None):
  """layers_util.py:149-173: the unique solution of the square system."""
  a, b = polynomial_accuracy_constraints([stencil], method, [derivative_order],
                                         stencil.size - derivative_order, grid_step)
  return np.linalg.solve(a, b)


def polynomial_accuracy_coefficients(stencils, method, derivative_orders, accuracy_order=None,
                                     grid_step=None):
  """layers_util.py:176-227: per axis the highest-order coefficients on the
  centred sub-stencil that the accuracy order needs, outer product."""
  slices, sizes, coefs = [], [], []
  for stencil, order in zip(stencils, derivative_orders):
    excess = 0 if accuracy_order is None else stencil.size - order - accuracy_order
    sl = slice(excess // 2, stencil.size - excess // 2)
    c = _high_order_coefficients_1d(stencil[sl], method, order, grid_step)
    slices.append(sl)
    sizes.append(stencil[sl].size)
    coefs.append(c)
  result = np.zeros(tuple(s.size for s in stencils))
  result[tuple(slices)] = _kron(coefs).reshape(sizes)
  return result


def get_roll_and_shift(input_offset, target_offset):
  """layers_util.py:230-237."""
  delta = [t - i for t, i in zip(target_offset, input_offset)]
  roll = tuple(-math.floor(d) for d in delta)
  shift = tuple(d + r for d, r in zip(delta, roll))
  return roll, shift


def get_stencils(stencil_sizes, offset, steps):
  """layers_util.py:240-269: cell centres relative to the target point."""
  out = []
  for size, o, step in zip(stencil_sizes, offset, steps):
    left = -((size - 1) // 2)
    out.append(np.array([(-o + s) * step for s in range(left, left + size)]))
  return tuple(out)


def patch_shifts(patch_shape):
  """Index offsets of extract_patches' channels (layers_util.py:324-341):
  per axis range(-size // 2 + 1, size // 2 + 1), row-major product."""
  per_axis = [range(-size // 2 + 1, size // 2 + 1) for size in patch_shape]
  return list(itertools.product(*per_axis))


def extract_patches(x, patch_shape):
  """layers_util.py:324-341 ('roll'): x (..., 1) -> (..., prod(patch_shape)),
  channel k = x rolled by -shift_k (value at index + shift_k)."""
  x = np.squeeze(x, -1)
  chans = [np.roll(x, tuple(-s for s in sh), axis=tuple(range(x.ndim)))
           for sh in patch_shifts(patch_shape)]
  return np.stack(chans, axis=-1)


def apply_coefficients(coefficients, stencil_values):
  """layers_util.py:445-447."""
  return np.sum(coefficients * stencil_values, axis=-1, keepdims=True)


class PolynomialConstraint:
  """layers.py:451-544: coefficients = bias + inputs . nullspace, where the
  nullspace rows are the last right-singular vectors of the constraint
  matrix (np.linalg.svd), divided by prod(step**derivative_orders)."""

  def __init__(self, stencils, derivative_orders, method, steps, accuracy_order=1,
               bias_accuracy_order=1, bias=None):
    grid_steps = {*steps}
    if len(grid_steps) != 1:
      raise ValueError('nonuniform steps not supported by PolynomialConstraint')
    (grid_step,) = grid_steps
    a, b = polynomial_accuracy_constraints(stencils, method, derivative_orders,
                                           accuracy_order, grid_step)
    if bias is None:
      bias = polynomial_accuracy_coefficients(stencils, method, derivative_orders,
                                              bias_accuracy_order, grid_step).ravel()
    self.bias = bias
    if np.linalg.norm(a @ bias - b) > 1e-8:
      raise ValueError('invalid bias, not in nullspace')
    _, _, vt = np.linalg.svd(a)
    size = a.shape[1] - a.shape[0]
    if not size:
      raise ValueError('there is only one valid solution accurate to this order')
    self.subspace_size = size
    self.nullspace = vt[-size:] / (grid_step ** np.array(derivative_orders)).prod()

  def __call__(self, inputs):
    """bias + inputs . nullspace with both constants rounded to the inputs'
    dtype first (jnp.asarray of the NumPy constants with x64 disabled)."""
    dtype = np.asarray(inputs).dtype
    return (self.bias.astype(dtype)
            + np.tensordot(inputs, self.nullspace.astype(dtype), axes=[-1, 0]))


class SpatialDerivativeFromLogits:
  """layers.py:586-642 (extract_patch_method 'roll', no tiling)."""

  def __init__(self, stencil_shape, input_offset, target_offset, derivative_orders, steps,
               method=Method.FINITE_VOLUME):
    self.stencil_shape = tuple(stencil_shape)
    self.roll, shift = get_roll_and_shift(input_offset, target_offset)
    stencils = get_stencils(self.stencil_shape, shift, steps)
    self.constraint = PolynomialConstraint(stencils, derivative_orders, method, steps)

  @property
  def subspace_size(self):
    return self.constraint.subspace_size

  def __call__(self, inputs, logits):
    """inputs (..., 1), logits (..., subspace_size) -> (..., 1)."""
    if logits.shape[-1] != self.subspace_size:
      raise ValueError('last dimension of logits does not match the subspace size')
    coefficients = self.constraint(logits)
    rolled = np.roll(inputs, self.roll, axis=tuple(range(len(self.roll))))
    return apply_coefficients(coefficients, extract_patches(rolled, self.stencil_shape))


def periodic_conv2d(x, w, b):
  """PeriodicConv2D (layers.py:24-116) with haiku's Conv2D: x (n0, n1, cin),
  w (kh, kw, cin, cout), b (cout,); wrap-pad kernel//2 before (the rest
  after), VALID cross-correlation, + b.  Sums in x's dtype."""
  kh, kw = w.shape[:2]
  pads = [(k // 2, k - k // 2 - 1) for k in (kh, kw)]
  xp = np.pad(x, pads + [(0, 0)], mode='wrap')
  n0, n1 = x.shape[:2]
  out = np.zeros((n0, n1, w.shape[3]), dtype=x.dtype)
  for a in range(kh):
    for c in range(kw):
      out = out + np.tensordot(xp[a:a + n0, c:c + n1, :], w[a, c], axes=[-1, 0]).astype(x.dtype)
  return (out + b).astype(x.dtype)


def forward_tower(params, inputs, nonlinearity=lambda x: np.maximum(x, 0)):
  """towers.py:201-262 with identity input/output scaling: conv, relu, ...,
  conv.  params: a list of (w, b) per convolution."""
  x = inputs
  for k, (w, b) in enumerate(params):
    x = periodic_conv2d(x, w, b)
    if k + 1 < len(params):
      x = nonlinearity(x)
  return x


def tower_shapes(num_output_channels, num_inputs=2, num_hidden_channels=16,
                 num_hidden_layers=2, kernel_size=3):
  """(kh, kw, cin, cout) of forward_tower_factory's convolutions."""
  shapes, cin = [], num_inputs
  for _ in range(num_hidden_layers):
    shapes.append((kernel_size, kernel_size, cin, num_hidden_channels))
    cin = num_hidden_channels
  shapes.append((kernel_size, kernel_size, cin, num_output_channels))
  return shapes


class FusedLearnedInterpolation:
  """interpolations.py:31-121 (tags, stencil_size int, 'roll' patches,
  unfused constraints).  `params` are the tower's (w, b) pairs."""

  def __init__(self, grid: Grid, v: Sequence[A], params, tags=(None,), stencil_size=4):
    self.derivatives = collections.OrderedDict()
    for u in v:
      for target in go.control_volume_offsets(u):
        for tag in tags:
          key = (tuple(u.offset), tuple(target), tag)
          self.derivatives[key] = SpatialDerivativeFromLogits(
              (stencil_size,) * grid.ndim, u.offset, target, (0,) * grid.ndim, grid.step)
    sizes = [d.subspace_size for d in self.derivatives.values()]
    inputs = np.stack([u.data for u in v], axis=-1)
    self.logits = forward_tower(params, inputs)
    if self.logits.shape[-1] != sum(sizes):
      raise ValueError('tower output channels do not match the interpolations')
    split = np.split(self.logits, np.cumsum(sizes)[:-1], axis=-1)
    self.split_logits = dict(zip(self.derivatives, split))

  def __call__(self, c: A, offset, v=None, dt=None, tag=None) -> A:
    if tuple(c.offset) == tuple(offset):
      return c
    key = (tuple(c.offset), tuple(offset), tag)
    if key not in self.derivatives:
      raise KeyError(f'No interpolator for key {key}')
    out = self.derivatives[key](c.data[..., None], self.split_logits[key])
    return A(out[..., 0], tuple(offset))


def advect_general(c: A, v: Sequence[A], u_interp, c_interp, grid: Grid) -> A:
  """advection.py:81-116 with _advect_aligned (34-78), periodic."""
  targets = go.control_volume_offsets(c)
  aligned_v = tuple(u_interp(u, t, v) for u, t in zip(v, targets))
  aligned_c = tuple(c_interp(c, t, aligned_v) for t in targets)
  flux = tuple(A(cc.data * u.data, u.offset) for cc, u in zip(aligned_c, aligned_v))
  adv = go.divergence(flux, grid)
  return A(-adv.data, adv.offset)


def learned_self_convection(grid: Grid, v: Sequence[A], params, tags=('u', 'c'),
                            stencil_size=4):
  """self_advection(modular_self_advection(FusedLearnedInterpolation)):
  advections.py:48-85 -- one interpolation module built from v, the
  velocity interpolated with tag 'u' and the advected component with 'c'."""
  interp = FusedLearnedInterpolation(grid, v, params, tags, stencil_size)
  u_fn = lambda c, o, vv=None: interp(c, o, tag='u')  # noqa: E731
  c_fn = lambda c, o, vv=None: interp(c, o, tag='c')  # noqa: E731
  return tuple(advect_general(u, v, u_fn, c_fn, grid).data for u in v)


@dataclasses.dataclass
class LearnedInterpolationStep:
  """modular_navier_stokes_model (ml/equations.py:92-146) with the
  learned self-advection: equation 'semi_implicit' (explicit diffusion,
  forward Euler: equations.py:120-151) or 'implicit_diffusion'
  (equations.py:154-195), fast-diagonal pressure."""
  grid: Grid
  dt: float
  params: list
  density: float = 1.0
  viscosity: float = 1e-3
  forcing: Optional[go.Forcing] = None
  equation: str = 'semi_implicit'
  tags: Tuple = ('u', 'c')
  stencil_size: int = 4

  def convection(self, v):
    return learned_self_convection(self.grid, v, self.params, self.tags, self.stencil_size)

  def explicit_terms(self, v):
    dv = list(self.convection(v))
    nu = self.viscosity / self.density
    dv = [d + go.diffuse(u, nu, self.grid).data for d, u in zip(dv, v)]
    if self.forcing is not None:
      f = self.forcing(v, self.grid)
      dv = [d + fi / self.density for d, fi in zip(dv, f)]
    return tuple(A(d, u.offset) for d, u in zip(dv, v))

  def __call__(self, v):
    if self.equation == 'semi_implicit':
      k = self.explicit_terms(v)
      vs = tuple(A(u.data + self.dt * (0 + 1 * kk.data), u.offset) for u, kk in zip(v, k))
      return ofvm.projection(vs, self.grid, ofvm.solve_fast_diag)
    if self.equation != 'implicit_diffusion':
      raise ValueError(self.equation)
    dv = [0 + c for c in self.convection(v)]
    if self.forcing is not None:
      f = self.forcing(v, self.grid)
      dv = [d + fi / self.density for d, fi in zip(dv, f)]
    vs = tuple(A(u.data + d * self.dt, u.offset) for u, d in zip(v, dv))
    vs = ofvm.projection(vs, self.grid, ofvm.solve_fast_diag)
    return ofvm.solve_fast_diag_diffusion(vs, self.viscosity, self.dt, self.grid)

  def rollout(self, v, steps):
    for _ in range(steps):
      v = self(v)
    return v


def init_tower_params(shapes, seed=0, dtype=np.float32, bias_scale=0.0, last_scale=1.0):
  """Seeded weights with haiku Conv2D's default initialiser: truncated
  normal (2 standard deviations) of stddev sqrt(1 / fan_in) / 0.87962566,
  fan_in = kh kw cin; biases zero unless bias_scale > 0 (then uniform in
  +-bias_scale).  last_scale multiplies the output convolution's weights."""
  rng = np.random.default_rng(seed)
  params = []
  for k, (kh, kw, cin, cout) in enumerate(shapes):
    std = np.sqrt(1.0 / (kh * kw * cin)) / 0.87962566103423978
    w = rng.standard_normal((kh, kw, cin, cout))
    bad = np.abs(w) > 2
    while bad.any():
      w[bad] = rng.standard_normal(int(bad.sum()))
      bad = np.abs(w) > 2
    w = w * std
    if k == len(shapes) - 1:
      w = w * last_scale
    b = (rng.uniform(-bias_scale, bias_scale, cout) if bias_scale > 0 else np.zeros(cout))
    params.append((w.astype(dtype), b.astype(dtype)))
  return params
