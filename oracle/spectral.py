"""Pseudospectral vorticity step of jax-cfd restated on the CPU (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``NavierStokes2D`` mirrors jax_cfd/spectral/equations.py:129-190 (and
``ForcedNavierStokes2D`` 194-220); ``crank_nicolson_rk4`` and
``crank_nicolson_rk3`` mirror spectral/time_stepping.py:117-170, 173-197; ``brick_wall_filter_2d`` and
``vorticity_to_velocity`` mirror spectral/utils.py:128-134, 154-189.  The
state is the complex64 rfft2 of the vorticity, shape (n, n//2 + 1).

Pinning (tests/test_oracle_spectral.py): CN-RK4 and CN-RK3 against
closed-form linear ODE solutions with the reference's per-method tolerances
(spectral/time_stepping_test.py:42-187); the
Taylor-Green vortex decay in closed form (spectral/equations_test.py:263-303
pattern); the forcing curl closed form (-2 n^2 at [0, 4] for sin(4y));
the 2/3 filter index sets counted by hand for n in {64, 128, 256, 512};
vorticity_to_velocity against the FVM curl round trip
(spectral/utils_test.py:61-85 pattern); rfft2 against the O(n^2) DFT.
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Tuple

import numpy as np

from oracle import fftlib
from oracle.grid_ops import Forcing, Grid

# Carpenter-Kennedy coefficients, spectral/time_stepping.py:192-194.
CNRK4_ALPHAS = (0, 0.1496590219993, 0.3704009573644, 0.6222557631345,
                0.9582821306748, 1)
CNRK4_BETAS = (0, -0.4178904745, -1.192151694643, -1.697784692471,
               -1.514183444257)
CNRK4_GAMMAS = (0.1496590219993, 0.3792103129999, 0.8229550293869,
                0.6994504559488, 0.1530572479681)
# Williamson coefficients, spectral/time_stepping.py:178-180.
CNRK3_ALPHAS = (0, 1 / 3, 3 / 4, 1)
CNRK3_BETAS = (0, -5 / 9, -153 / 128)
CNRK3_GAMMAS = (1 / 3, 15 / 16, 8 / 15)


def brick_wall_filter_2d(grid: Grid, dtype=np.float32) -> np.ndarray:
  """2/3 rule mask (spectral/utils.py:128-134), Python int/floor semantics."""
  n, m = grid.shape
  f = np.zeros((n, m // 2 + 1), dtype)
  cols = int(2 / 3 * (m // 2 + 1))
  f[:int(2 / 3 * n) // 2, :cols] = 1
  f[-int(2 / 3 * n) // 2:, :cols] = 1
  return f


def _ctype(dtype):
  return np.complex64 if np.dtype(dtype) == np.float32 else np.complex128


def vorticity_to_velocity(grid: Grid, dtype=np.float32):
  """spectral/utils.py:154-189."""
  kx, ky = grid.rfft_mesh(dtype)
  two_pi_i = 2 * np.pi * 1j
  laplace = (two_pi_i ** 2 * (abs(kx) ** 2 + abs(ky) ** 2)).astype(_ctype(dtype))
  laplace[0, 0] = 1

  def ret(vorticity_hat):
    psi_hat = -1 / laplace * vorticity_hat
    vxhat = two_pi_i * ky * psi_hat
    vyhat = -two_pi_i * kx * psi_hat
    return vxhat.astype(vorticity_hat.dtype), vyhat.astype(vorticity_hat.dtype)
  return ret


@dataclasses.dataclass
class NavierStokes2D:
  """spectral/equations.py:129-190.  forcing: offsets must be set explicitly."""
  viscosity: float
  grid: Grid
  drag: float = 0.0
  smooth: Optional[bool] = True
  forcing: Optional[Forcing] = None
  dtype: type = np.float32

  def __post_init__(self):
    ct = _ctype(self.dtype)
    self.kx, self.ky = self.grid.rfft_mesh(self.dtype)
    self.laplace = ((np.pi * 2j) ** 2 * (self.kx ** 2 + self.ky ** 2)).astype(ct)
    self.filter_ = brick_wall_filter_2d(self.grid, self.dtype)
    self.linear_term = (self.viscosity * self.laplace - self.drag).astype(ct)
    self._v2v = vorticity_to_velocity(self.grid, self.dtype)
    self.forcing_curl = None
    if self.forcing is not None:
      fx, fy = self.forcing.constant_terms(self.grid, self.dtype)
      fx_hat, fy_hat = np.fft.rfft2(fx), np.fft.rfft2(fy)
      self.forcing_curl = (2j * np.pi * (fy_hat * self.kx - fx_hat * self.ky)
                           ).astype(ct)

  def explicit_terms(self, vorticity_hat):
    """spectral/equations.py:160-184."""
    n = self.grid.shape
    vxhat, vyhat = self._v2v(vorticity_hat)
    vx, vy = fftlib.irfftn(vxhat, s=n, axes=(0, 1)), fftlib.irfftn(vyhat, s=n, axes=(0, 1))
    grad_x_hat = 2j * np.pi * self.kx * vorticity_hat
    grad_y_hat = 2j * np.pi * self.ky * vorticity_hat
    grad_x = fftlib.irfftn(grad_x_hat, s=n, axes=(0, 1))
    grad_y = fftlib.irfftn(grad_y_hat, s=n, axes=(0, 1))
    advection = -(grad_x * vx + grad_y * vy)
    advection_hat = fftlib.rfftn(advection)
    if self.smooth is not None:
      advection_hat = advection_hat * self.filter_
    terms = advection_hat
    if self.forcing_curl is not None:
      # The reference adds curl(rfft2(f)) of the linear-forcing part too; the
      # forcings used on this path have no linear part, so the term is
      # constant.  (Kolmogorov forcing ignores its velocity argument.)
      if self.forcing.linear_coefficient != 0:
        raise NotImplementedError('linear forcing in the spectral equation')
      terms = terms + self.forcing_curl
    return terms.astype(vorticity_hat.dtype)

  def implicit_terms(self, vorticity_hat):
    return self.linear_term * vorticity_hat

  def implicit_solve(self, vorticity_hat, time_step):
    return 1 / (1 - time_step * self.linear_term) * vorticity_hat


def ForcedNavierStokes2D(viscosity, grid, smooth, dtype=np.float32):
  """spectral/equations.py:194-220: Kolmogorov k=4 at offsets (0,0), drag 0.1."""
  forcing = Forcing(scale=1, wavenumber=4, linear_coefficient=0.0,
                    offsets=((0, 0), (0, 0)))
  return NavierStokes2D(viscosity, grid, drag=0.1, smooth=smooth,
                        forcing=forcing, dtype=dtype)


def low_storage_rk_crank_nicolson(alphas, betas, gammas, equation, time_step):
  """spectral/time_stepping.py:117-170."""
  dt = time_step
  F, G, G_inv = (equation.explicit_terms, equation.implicit_terms,
                 equation.implicit_solve)

  def step_fn(u):
    h = 0
    for k in range(len(betas)):
      h = F(u) + betas[k] * h
      mu = 0.5 * dt * (alphas[k + 1] - alphas[k])
      u = G_inv(u + gammas[k] * dt * h + mu * G(u), mu)
    return u
  return step_fn


def crank_nicolson_rk4(equation, time_step):
  """spectral/time_stepping.py:186-197."""
  return low_storage_rk_crank_nicolson(CNRK4_ALPHAS, CNRK4_BETAS, CNRK4_GAMMAS,
                                       equation, time_step)


def crank_nicolson_rk3(equation, time_step):
  """spectral/time_stepping.py:173-183 (Williamson's coefficients)."""
  return low_storage_rk_crank_nicolson(CNRK3_ALPHAS, CNRK3_BETAS, CNRK3_GAMMAS,
                                       equation, time_step)


def rollout(step_fn, state, steps: int):
  for _ in range(steps):
    state = step_fn(state)
  return state
