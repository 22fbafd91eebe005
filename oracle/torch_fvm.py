"""Differentiable torch restatement of the FVM step (oracle, CPU).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The jax.grad stand-in for the adjoint: the same operations as oracle.fvm /
oracle.grid_ops (advect_van_leer with its reference structure,
advection.py:210-342; diffusion; Kolmogorov + linear forcing; forward Euler;
the fast-diagonal projection through torch.fft), differentiated by
torch.autograd.  The differentiation rules match jax's for every operation
used: `where` sends the cotangent to the selected branch, sign has zero
derivative, division follows the quotient rule, and |x| is written as
where(x >= 0, x, -x) so its derivative at 0 is +1, as jax's JVP rule for abs
on reals (select(x >= 0, g, -g)) gives.  That rule at 0 is parity unpinned:
no reference test reaches |U| at U = 0 with a nonzero limiter value.  Pinned
by the reference's gradient tests (advection_test.py:298-351, restated in
tests/test_oracle_torch.py), central differences, and equality with the
NumPy oracle's forward step.  Runs in the dtype of its inputs (fp64 for the
gradient reference).
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Tuple

import numpy as np
import torch

from oracle import fvm as ofvm
from oracle.grid_ops import Grid


def jabs(x):
  return torch.where(x >= 0, x, -x)


@dataclasses.dataclass
class T:
  data: torch.Tensor
  offset: Tuple[float, ...]


def shift(c: T, k: int, axis: int) -> T:
  off = tuple(o + k if a == axis else o for a, o in enumerate(c.offset))
  return T(torch.roll(c.data, -k, dims=axis), off)


def linear(c: T, offset) -> T:
  out = c
  for axis, o in enumerate(offset):
    delta = o - out.offset[axis]
    if delta == 0:
      continue
    new_off = tuple(o if j == axis else x for j, x in enumerate(out.offset))
    if int(delta) == delta:
      out = T(shift(out, int(delta), axis).data, new_off)
      continue
    floor, ceil = int(np.floor(delta)), int(np.ceil(delta))
    fw = ceil - delta
    cw = 1.0 - fw
    out = T(fw * shift(out, floor, axis).data + cw * shift(out, ceil, axis).data, new_off)
  return out


def cv_offsets(c: T):
  return tuple(tuple(o + .5 if i == j else o for i, o in enumerate(c.offset))
               for j in range(len(c.offset)))


def divergence(v, grid: Grid):
  data = 0
  for ax, u in enumerate(v):
    data = data + (u.data + (-shift(u, -1, ax).data)) / grid.step[ax]
  return data


def advect_van_leer(c: T, v, dt: float, grid: Grid) -> torch.Tensor:
  """advection.py:210-342 (periodic), same structure as oracle.grid_ops."""
  ndim = grid.ndim
  c_left = c
  ax = 0
  for ax in range(ndim):
    c_left = shift(c_left, int(c.offset[ax] % 1 - c.offset[ax]), ax)
  offsets = cv_offsets(c_left)
  aligned = tuple(linear(u, o) for u, o in zip(v, offsets))
  flux = []
  for axis, u in enumerate(aligned):
    cc = c_left.data
    cr = shift(c_left, +1, axis).data
    flux.append(T(torch.where(u.data > 0, u.data * cc, u.data * cr), u.offset))
  for axis, (u, h) in enumerate(zip(aligned, grid.step)):
    u = shift(u, int(u.offset[axis] % 1 - u.offset[axis]), axis)
    cca = shift(c, int(1 - c.offset[ax]), ax)  # stale `ax`, as the reference
    c_center = cca.data
    c_right = shift(cca, +1, axis).data
    c_lft = shift(cca, -1, axis).data
    c_ll = shift(cca, -2, axis).data
    num = torch.where(u.data > 0, c_lft - c_ll, c_right - c_center)
    den = c_center - c_lft
    dnn = torch.where(jabs(den) > 0, jabs(den) + jabs(num), torch.ones_like(den))
    inv = den / dnn
    phi = num * (torch.sign(den) + torch.sign(num)) * inv
    av = jabs(u.data)
    pre = 0.5 * (1 - (dt / h) * av) * av
    corr = shift(T(pre * phi, u.offset), int(offsets[axis][axis] - u.offset[axis]), axis)
    flux[axis] = T(flux[axis].data + corr.data, flux[axis].offset)
  adv = T(-divergence(flux, grid), c.offset)  # offset bookkeeping only
  for ax2 in range(ndim):
    adv = shift(adv, -int(c.offset[ax2] % 1 - c.offset[ax2]), ax2)
  return adv.data


def laplacian(x: torch.Tensor, grid: Grid) -> torch.Tensor:
  dtype = np.float32 if x.dtype == torch.float32 else np.float64
  scales = np.square(1 / np.array(grid.step, dtype=dtype))
  out = -2 * x * float(np.sum(scales))
  for axis in range(grid.ndim):
    out = out + (torch.roll(x, 1, axis) + torch.roll(x, -1, axis)) * float(scales[axis])
  return out


@dataclasses.dataclass
class TorchFVMStep:
  """Forward-Euler step of semi_implicit_navier_stokes with advect_van_leer."""
  grid: Grid
  dt: float
  viscosity: float = 1e-3
  density: float = 1.0
  linear_coefficient: Optional[float] = None
  forcing_x: Optional[np.ndarray] = None  # f_x(y_j)

  def __post_init__(self):
    self._diag = {}

  def _pinv(self, dtype):
    if dtype not in self._diag:
      d = ofvm.poisson_pinv_diagonal(self.grid, np.float64 if dtype == torch.float64
                                     else np.float32)
      self._diag[dtype] = torch.as_tensor(d)
    return self._diag[dtype]

  def projection(self, u, v):
    faces = self.grid.cell_faces
    rhs = divergence((T(u, faces[0]), T(v, faces[1])), self.grid)
    q = torch.fft.irfftn(self._pinv(u.dtype) * torch.fft.rfftn(rhs), s=rhs.shape)
    gx = (torch.roll(q, -1, 0) + (-q)) / self.grid.step[0]
    gy = (torch.roll(q, -1, 1) + (-q)) / self.grid.step[1]
    return u - gx, v - gy

  def __call__(self, u, v):
    faces = self.grid.cell_faces
    vel = (T(u, faces[0]), T(v, faces[1]))
    k = [advect_van_leer(c, vel, self.dt, self.grid) for c in vel]
    nu = self.viscosity / self.density
    k = [kk + nu * laplacian(c.data, self.grid) for kk, c in zip(k, vel)]
    if self.linear_coefficient is not None:
      fx = torch.as_tensor(self.forcing_x if self.forcing_x is not None
                           else np.zeros(self.grid.shape[1]), dtype=u.dtype)
      f = [0 + self.linear_coefficient * u + fx[None, :],
           0 + self.linear_coefficient * v + 0.0]
      k = [kk + ff / self.density for kk, ff in zip(k, f)]
    us = u + self.dt * (0 + 1 * k[0])
    vs = v + self.dt * (0 + 1 * k[1])
    return self.projection(us, vs)


def curl_2d(u, v, grid: Grid):
  """finite_differences.py:186-193."""
  return ((torch.roll(v, -1, 0) + (-v)) / grid.step[0]
          - (torch.roll(u, -1, 1) + (-u)) / grid.step[1])
