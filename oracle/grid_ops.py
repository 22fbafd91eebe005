"""Offset-tracking arrays and the discrete operators of the FVM path (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Everything here is periodic.  An ``A`` pairs a NumPy array with its staggered
offset, as jax-cfd's ``GridArray`` does (jax_cfd/base/grids.py:40-108), and
the functions mirror the reference functions named in their docstrings,
including operation order, so that fp32 results round the same way up to the
FFT/`sin` library differences.  Python-float scalars act on fp32 arrays as
weak types (NumPy 2 / NEP 50), the same promotion rule JAX applies.

Pinning (tests/test_oracle_fvm.py): shift/roll KATs, finite-difference KATs,
linear vs scipy interpn, upwind / Lax-Wendroff KATs, Van Leer vs
Van-Leer-using-limiters on cell-centred scalars, TVD and mass conservation,
Laplacian of constants/quadratics.  The velocity-component form of
``advect_van_leer`` (the stale-loop-variable quirk) is pinned by a brute-force
per-cell evaluation of the hand-derived stencil in the tests.
``map_coordinates_wrap_linear`` is pinned against SciPy's independent
``ndimage.map_coordinates(order=1, mode='grid-wrap')``, and
``advect_step_semilagrangian`` by exact periodic rolls (whole-cell
displacements), the two-point average at Courant 0.5 and the reference's
Gaussian translation case (advection_test.py:140-170).
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Optional, Sequence, Tuple

import numpy as np


@dataclasses.dataclass(frozen=True)
class Grid:
  """Periodic grid geometry; mirrors jax_cfd/base/grids.py:499-555."""
  shape: Tuple[int, ...]
  step: Tuple[float, ...]
  domain: Tuple[Tuple[float, float], ...]

  @classmethod
  def from_domain(cls, shape, domain):
    shape = tuple(int(s) for s in shape)
    if isinstance(domain, (int, float)):
      domain = ((0.0, float(domain)),) * len(shape)
    domain = tuple((float(lo), float(hi)) for lo, hi in domain)
    step = tuple((hi - lo) / n for (lo, hi), n in zip(domain, shape))
    return cls(shape, step, domain)

  @classmethod
  def from_step(cls, shape, step):
    shape = tuple(int(s) for s in shape)
    if isinstance(step, (int, float)):
      step = (float(step),) * len(shape)
    domain = tuple((0.0, float(s * n)) for s, n in zip(step, shape))
    return cls.from_domain(shape, domain)

  @property
  def ndim(self):
    return len(self.shape)

  @property
  def cell_center(self):
    return (0.5,) * self.ndim

  @property
  def cell_faces(self):
    """grids.py:567-572: offsets (1, .5, ...), (.5, 1, ...), ..."""
    d = self.ndim
    return tuple(tuple(1.0 if i == j else 0.5 for i in range(d))
                 for j in range(d))

  def axes(self, offset, dtype=np.float32):
    """grids.py:584-602: lower + (arange + offset) * step, in `dtype`."""
    out = []
    for (lo, _), o, n, h in zip(self.domain, offset, self.shape, self.step):
      a = np.arange(n, dtype=np.int32)
      x = (a + dtype(o)).astype(dtype) * dtype(h)
      out.append((dtype(lo) + x).astype(dtype))
    return tuple(out)

  def mesh(self, offset, dtype=np.float32):
    return tuple(np.meshgrid(*self.axes(offset, dtype), indexing='ij'))

  def fft_freqs(self, dtype=np.float32):
    """grids.py:604-658 `fftfreq(n, d=step)` / `rfftfreq` as computed in fp32.

    JAX forms the integer ordinals and divides by ``d*n`` in the working dtype.
    """
    out = []
    for axis, (n, h) in enumerate(zip(self.shape, self.step)):
      if axis == self.ndim - 1:
        k = np.arange(n // 2 + 1)
      else:
        k = np.concatenate([np.arange((n + 1) // 2), -np.arange(n // 2, 0, -1)])
      out.append((k.astype(dtype) / dtype(h * n)).astype(dtype))
    return tuple(out)

  def rfft_mesh(self, dtype=np.float32):
    return tuple(np.meshgrid(*self.fft_freqs(dtype), indexing='ij'))


@dataclasses.dataclass
class A:
  """Array with a staggered offset (the oracle's GridArray)."""
  data: np.ndarray
  offset: Tuple[float, ...]

  @property
  def dtype(self):
    return self.data.dtype


def shift(c: A, k: int, axis: int) -> A:
  """Periodic shift: out[i] = c[i + k]; offset += k.

  Mirrors ConstantBoundaryConditions.shift for periodic BCs
  (jax_cfd/base/boundaries.py:81-104, 193-199: pad(wrap) + trim == roll).
  """
  offset = tuple(o + k if a == axis else o for a, o in enumerate(c.offset))
  return A(np.roll(c.data, -k, axis=axis), offset)


def _linear_along_axis(c: A, offset: float, axis: int) -> A:
  """jax_cfd/base/interpolation.py:36-64."""
  delta = offset - c.offset[axis]
  if delta == 0:
    return c
  new_offset = tuple(offset if j == axis else o for j, o in enumerate(c.offset))
  if int(delta) == delta:
    return A(shift(c, int(delta), axis).data, new_offset)
  floor = int(np.floor(delta))
  ceil = int(np.ceil(delta))
  floor_weight = ceil - delta
  ceil_weight = 1.0 - floor_weight
  data = (floor_weight * shift(c, floor, axis).data
          + ceil_weight * shift(c, ceil, axis).data)
  return A(data, new_offset)


def linear(c: A, offset: Sequence[float]) -> A:
  """Multilinear interpolation to `offset` (interpolation.py:67-93)."""
  out = c
  for a, o in enumerate(offset):
    out = _linear_along_axis(out, o, a)
  return out


def control_volume_offsets(c: A):
  """grids.py:448-454."""
  return tuple(tuple(o + 0.5 if i == j else o for i, o in enumerate(c.offset))
               for j in range(len(c.offset)))


# ---------------------------------------------------------------- differences
def backward_difference(u: A, axis: int, step: float) -> A:
  """(u[i] - u[i-1]) / h  (finite_differences.py:95-102, 44-54)."""
  s = shift(u, -1, axis)
  offset = tuple(0.5 * (a + b) for a, b in zip(u.offset, s.offset))
  return A((u.data + (-s.data)) / step, offset)


def forward_difference(u: A, axis: int, step: float) -> A:
  """(u[i+1] - u[i]) / h  (finite_differences.py:117-124)."""
  s = shift(u, +1, axis)
  offset = tuple(0.5 * (a + b) for a, b in zip(u.offset, s.offset))
  return A((s.data + (-u.data)) / step, offset)


def divergence(v: Sequence[A], grid: Grid) -> A:
  """Sum of backward differences, starting from 0 (fd.py:136-143)."""
  diffs = [backward_difference(u, ax, grid.step[ax]) for ax, u in enumerate(v)]
  data = 0
  for d in diffs:
    data = data + d.data
  return A(data, diffs[0].offset)


def laplacian(u: A, grid: Grid) -> A:
  """fd.py:127-133: -2u*sum(scales) + sum_axis (u[-1]+u[+1])*scale."""
  dtype = u.data.dtype.type
  scales = np.square(1 / np.array(grid.step, dtype=dtype))
  result = -2 * u.data * np.sum(scales)
  for axis in range(grid.ndim):
    result = result + (shift(u, -1, axis).data
                       + shift(u, +1, axis).data) * scales[axis]
  return A(result, u.offset)


def curl_2d(v: Sequence[A], grid: Grid) -> A:
  """fd.py:186-193: d v/dx (forward) - d u/dy (forward)."""
  a = forward_difference(v[1], 0, grid.step[0])
  b = forward_difference(v[0], 1, grid.step[1])
  return A(a.data - b.data, a.offset)


# ------------------------------------------------------------------ advection
def advect_van_leer(c: A, v: Sequence[A], dt: float, grid: Grid) -> A:
  """Van Leer flux-limited advection; mirrors advection.py:210-342 line by line.

  Periodic boundaries only, so the flux-BC and mode arguments of the
  reference are identities here.  The loop variable ``ax`` left over from the
  first loop (advection.py:259-267) is what selects ``c_center_arr`` at
  advection.py:291, exactly as in the reference.
  """
  ndim = grid.ndim
  c_left_var = c
  ax = 0
  for ax in range(ndim):
    c_left_var = shift(c_left_var, int(c.offset[ax] % 1 - c.offset[ax]), ax)
  offsets = control_volume_offsets(c_left_var)
  aligned_v = tuple(linear(u, off) for u, off in zip(v, offsets))
  flux = []
  for axis, u in enumerate(aligned_v):
    c_center = c_left_var.data
    c_right = shift(c_left_var, +1, axis).data
    flux.append(A(np.where(u.data > 0, u.data * c_center, u.data * c_right),
                  u.offset))
  for axis, (u, h) in enumerate(zip(aligned_v, grid.step)):
    u = shift(u, int(u.offset[axis] % 1 - u.offset[axis]), axis)
    c_center_arr = shift(c, int(1 - c.offset[ax]), ax)  # stale `ax` (ref:291)
    c_center = c_center_arr.data
    c_right = shift(c_center_arr, +1, axis).data
    c_left = shift(c_center_arr, -1, axis).data
    c_left_left = shift(c_center_arr, -2, axis).data
    numerator_positive = c_left - c_left_left
    numerator_negative = c_right - c_center
    numerator = np.where(u.data > 0, numerator_positive, numerator_negative)
    denominator = c_center - c_left
    denom_avoid_nans = np.where(abs(denominator) > 0,
                                abs(denominator) + abs(numerator),
                                denominator.dtype.type(1.0))
    denom_inv = denominator / denom_avoid_nans
    phi = numerator * (np.sign(denominator) + np.sign(numerator)) * denom_inv
    abs_velocity = abs(u.data)
    courant = (dt / h) * abs_velocity
    pre_factor = 0.5 * (1 - courant) * abs_velocity
    correction = A(pre_factor * phi, u.offset)
    correction = shift(correction, int(offsets[axis][axis] - u.offset[axis]),
                       axis)
    flux[axis] = A(flux[axis].data + correction.data, flux[axis].offset)
  adv = divergence(flux, grid)
  adv = A(-adv.data, adv.offset)
  for ax in range(ndim):
    adv = shift(adv, -int(c.offset[ax] % 1 - c.offset[ax]), ax)
  return adv


def upwind(c: A, offset, v: Sequence[A]) -> A:
  """interpolation.py:96-155."""
  if tuple(c.offset) == tuple(offset):
    return c
  axes = [a for a, (p, q) in enumerate(zip(c.offset, offset)) if p != q]
  axis, = axes
  u = v[axis]
  delta = u.offset[axis] - c.offset[axis]
  if int(delta) == delta:
    return A(shift(c, int(delta), axis).data, tuple(offset))
  floor = int(np.floor(delta))
  ceil = int(np.ceil(delta))
  return A(np.where(u.data > 0, shift(c, floor, axis).data,
                    shift(c, ceil, axis).data), tuple(offset))


def lax_wendroff(c: A, offset, v: Sequence[A], dt: float, grid: Grid) -> A:
  """interpolation.py:158-221."""
  if tuple(c.offset) == tuple(offset):
    return c
  axes = [a for a, (p, q) in enumerate(zip(c.offset, offset)) if p != q]
  axis, = axes
  u = v[axis]
  delta = u.offset[axis] - c.offset[axis]
  floor = int(np.floor(delta))
  ceil = int(np.ceil(delta))
  courant = (dt / grid.step[axis]) * u.data
  cf = shift(c, floor, axis).data
  cc = shift(c, ceil, axis).data
  pos = cf + 0.5 * (1 - courant) * (cc - cf)
  neg = cc - 0.5 * (1 + courant) * (cc - cf)
  return A(np.where(u.data > 0, pos, neg), tuple(offset))


def safe_div(x, y, default_numerator=1):
  """interpolation.py:224-226."""
  return x / np.where(y != 0, y, y.dtype.type(default_numerator))


def van_leer_limiter(r):
  """interpolation.py:229-231."""
  with np.errstate(divide='ignore', invalid='ignore'):
    return np.where(r > 0, safe_div(2 * r, 1 + r), r.dtype.type(0.0))


def tvd_van_leer_lax_wendroff(c: A, offset, v: Sequence[A], dt: float,
                              grid: Grid) -> A:
  """apply_tvd_limiter(lax_wendroff, van_leer_limiter) (interpolation.py:234-303)."""
  for axis, axis_offset in enumerate(offset):
    interp_offset = tuple(co if i != axis else axis_offset
                          for i, co in enumerate(c.offset))
    if interp_offset != tuple(c.offset):
      if interp_offset[axis] - c.offset[axis] != 0.5:
        raise NotImplementedError('forward interpolation to faces only')
      c_low = upwind(c, offset, v)
      c_high = lax_wendroff(c, offset, v, dt, grid)
      c_left = shift(c, -1, axis).data
      c_right = shift(c, 1, axis).data
      c_next_right = shift(c, 2, axis).data
      with np.errstate(divide='ignore', invalid='ignore'):
        pos_r = safe_div(c.data - c_left, c_right - c.data)
        neg_r = safe_div(c_next_right - c_right, c_right - c.data)
      phi = np.where(v[axis].data > 0, van_leer_limiter(pos_r),
                     van_leer_limiter(neg_r))
      data = c_low.data - (c_low.data - c_high.data) * phi
      c = A(data, interp_offset)
  return c


def advect_linear(c: A, v: Sequence[A], grid: Grid) -> A:
  """advect_general(c, v, linear, linear) (advection.py:34-123): both c and
  the velocity interpolated linearly to the control-volume faces, flux
  c_face * u_face, minus its divergence."""
  targets = control_volume_offsets(c)
  aligned_v = tuple(linear(u, t) for u, t in zip(v, targets))
  aligned_c = tuple(linear(c, t) for t in targets)
  flux = tuple(A(cc.data * u.data, u.offset)
               for cc, u in zip(aligned_c, aligned_v))
  adv = divergence(flux, grid)
  return A(-adv.data, adv.offset)


def advect_upwind(c: A, v: Sequence[A], grid: Grid) -> A:
  """advect_general(c, v, linear, upwind) (advection.py:34-116, 126-130)."""
  targets = control_volume_offsets(c)
  aligned_v = tuple(linear(u, t) for u, t in zip(v, targets))
  aligned_c = tuple(upwind(c, t, aligned_v) for t in targets)
  flux = tuple(A(cc.data * u.data, u.offset)
               for cc, u in zip(aligned_c, aligned_v))
  adv = divergence(flux, grid)
  return A(-adv.data, adv.offset)


def advect_van_leer_using_limiters(c: A, v: Sequence[A], dt: float,
                                   grid: Grid) -> A:
  """advect_general(c, v, linear, tvd(lax_wendroff)) (advection.py:34-116, 387-395)."""
  targets = control_volume_offsets(c)
  aligned_v = tuple(linear(u, t) for u, t in zip(v, targets))
  aligned_c = tuple(tvd_van_leer_lax_wendroff(c, t, aligned_v, dt, grid)
                    for t in targets)
  flux = tuple(A(cc.data * u.data, u.offset)
               for cc, u in zip(aligned_c, aligned_v))
  adv = divergence(flux, grid)
  return A(-adv.data, adv.offset)


def map_coordinates_wrap_linear(a: np.ndarray, indices: Sequence[np.ndarray]) -> np.ndarray:
  """jax.scipy.ndimage.map_coordinates(a, indices, order=1, mode='wrap')
  (JAX's published implementation): per axis lower = floor(x), weights
  (1 - (x - lower), x - lower) at indices (lower, lower + 1) mod size; the
  contributions in itertools.product order over the axes, each the product
  of its weights (left to right) times the value, summed left to right."""
  import itertools
  dtype = a.dtype.type
  nodes = []
  for x, size in zip(indices, a.shape):
    x = np.asarray(x, dtype)
    lower = np.floor(x)
    w_up = (x - lower).astype(dtype)
    w_lo = (dtype(1) - w_up).astype(dtype)
    k = lower.astype(np.int64)
    nodes.append([(np.mod(k, size), w_lo), (np.mod(k + 1, size), w_up)])
  out = None
  for items in itertools.product(*nodes):
    w = items[0][1]
    for t in items[1:]:
      w = w * t[1]
    term = w * a[tuple(t[0] for t in items)]
    out = term if out is None else out + term
  return out


def advect_step_semilagrangian(c: A, v: Sequence[A], dt: float, grid: Grid) -> A:
  """advection.py:345-382: departure points x - dt * linear(u, c.offset) on
  c's mesh (grids.py:584-602, float32), fractional indices x / step - offset,
  then map_coordinates_wrap_linear.  Returns the advected field."""
  if not all(lo == 0 for lo, _ in grid.domain):
    raise ValueError(f'Grid domains currently must start at zero. Found {grid.domain}')
  dtype = c.data.dtype.type
  indices = []
  for axis, u in enumerate(v):
    n, h, o = grid.shape[axis], grid.step[axis], c.offset[axis]
    x1 = (np.arange(n, dtype=np.int32) + dtype(o)).astype(dtype)
    x1 = (dtype(grid.domain[axis][0]) + x1 * dtype(h)).astype(dtype)
    x = np.meshgrid(*[x1 if a == axis else np.zeros(grid.shape[a], dtype)
                      for a in range(grid.ndim)], indexing='ij')[axis]
    coord = x - dtype(dt) * linear(u, c.offset).data.astype(dtype)
    indices.append((coord / dtype(h) - dtype(o)).astype(dtype))
  return A(map_coordinates_wrap_linear(c.data, indices), c.offset)


def stable_time_step_advection(max_velocity: float, max_courant_number: float,
                               grid: Grid) -> float:
  """advection.py:398-416."""
  return max_courant_number * min(grid.step) / max_velocity


# ----------------------------------------------------------------- diffusion
def diffuse(c: A, nu: float, grid: Grid) -> A:
  """diffusion.py:35-37: nu * laplacian(c)."""
  lap = laplacian(c, grid)
  return A(nu * lap.data, lap.offset)


def stable_time_step_diffusion(viscosity: float, grid: Grid) -> float:
  """diffusion.py:40-57."""
  if viscosity == 0:
    return float('inf')
  return min(grid.step) ** 2 / (viscosity * 2 ** grid.ndim)


# ------------------------------------------------------------------- forcing
@dataclasses.dataclass(frozen=True)
class Forcing:
  """Kolmogorov (sin(k y) on the x component) plus linear drag.

  Mirrors simple_turbulence_forcing(..., forcing_type='kolmogorov')
  (forcings.py:136-178) = sum_forcings(linear_forcing, kolmogorov_forcing)
  (forcings.py:63-129); ``offsets=None`` means the velocity's own
  cell-face offsets, as kolmogorov_forcing's default.
  """
  scale: float = 0.0
  wavenumber: int = 2
  linear_coefficient: float = 0.0
  offsets: Optional[Tuple[Tuple[float, ...], ...]] = None

  def constant_terms(self, grid: Grid, dtype=np.float32):
    offsets = self.offsets if self.offsets is not None else grid.cell_faces
    y = grid.mesh(offsets[0], dtype)[1]
    fx = (self.scale * np.sin(self.wavenumber * y)).astype(dtype)
    return fx, np.zeros_like(fx)

  def __call__(self, v: Sequence[A], grid: Grid):
    dtype = v[0].data.dtype.type
    fx, fy = self.constant_terms(grid, dtype)
    out = []
    for u, f in zip(v, (fx, fy)):
      lin = self.linear_coefficient * u.data
      out.append((0 + lin) + f)
    return tuple(out)
