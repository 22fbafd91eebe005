"""Pseudospectral Navier-Stokes in vorticity form (jax_cfd/spectral/equations.py).

``NavierStokes2D`` keeps the reference fields and methods (129-190).  Its
``explicit_terms`` runs the fused HIP kernels; ``crank_nicolson_rk4`` of it
(jax_cfd_amd.spectral.time_stepping) advances whole steps on the GPU.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Optional

import numpy as np
import torch

from jax_cfd_amd import _lib, _symbolic
from jax_cfd_amd.base import boundaries, forcings, grids
from jax_cfd_amd.spectral import _ns2d
from jax_cfd_amd.spectral import utils as spectral_utils


def _forcing_curl_entries(forcing, grid):
  """Sparse constant curl of a constant forcing: 2 pi i (fy^ kx - fx^ ky)
  with fx^, fy^ = rfft2 of the forcing fields in complex64
  (spectral/equations.py:175-182).  Entries below 1e-5 of the largest are
  fp32 FFT leakage of the exact (sparse) transform and are dropped."""
  if forcing is None:
    return ()
  bc = boundaries.periodic_boundary_conditions(2)
  probe = tuple(grids.GridVariable(grids.GridArray(_symbolic.Affine(i),
                                                   (0.5, 0.5), grid), bc)
                for i in range(2))
  fx, fy = forcing(probe)
  consts = []
  for f in (fx, fy):
    d = getattr(f, 'data', f)
    if isinstance(d, _symbolic.Affine):
      if d.coef != 0:
        raise NotImplementedError('velocity-dependent forcing in the spectral '
                                  'equation is not built')
      d = d.const if d.const is not None else np.zeros(grid.shape, np.float32)
    consts.append(np.broadcast_to(np.asarray(_symbolic.Affine._as_np(d), np.float32),
                                  grid.shape))
  kx, ky = grid.rfft_mesh()
  fxh = np.fft.rfft2(consts[0])
  fyh = np.fft.rfft2(consts[1])
  curl = (2j * np.pi * (fyh * kx - fxh * ky)).astype(np.complex64)
  mag = np.abs(curl)
  if mag.max() == 0:
    return ()
  idx = np.argwhere(mag > 1e-5 * mag.max())
  if len(idx) > 256:
    raise NotImplementedError('forcing curl with more than 256 modes')
  return tuple((int(p), int(c), float(curl[p, c].real), float(curl[p, c].imag))
               for p, c in idx)


@dataclasses.dataclass
class NavierStokes2D:
  """Implicit (linear) / explicit (nonlinear) split of 2-D Navier-Stokes.

  Attributes as in spectral/equations.py:129-147: viscosity, grid, drag,
  smooth (2/3 rule applied whenever not None), forcing_fn(grid) -> forcing.
  """
  viscosity: float
  grid: grids.Grid
  drag: float = 0.
  smooth: Optional[bool] = True
  forcing_fn: Optional[Callable] = None
  _forcing_fn_with_grid = None

  def __post_init__(self):
    if self.grid.ndim != 2 or self.grid.shape[0] != self.grid.shape[1]:
      raise NotImplementedError('square 2-D grids only')
    self.kx, self.ky = self.grid.rfft_mesh()
    lap = ((np.pi * 2j) ** 2 * (self.kx ** 2 + self.ky ** 2)).astype(np.complex64)
    self.laplace = lap
    self.filter_ = spectral_utils.brick_wall_filter_2d(self.grid)
    self.linear_term = (self.viscosity * lap - self.drag).astype(np.complex64)
    if self.forcing_fn is not None:
      self._forcing_fn_with_grid = self.forcing_fn(self.grid)
    self._forcing_entries = _forcing_curl_entries(self._forcing_fn_with_grid,
                                                  self.grid)
    self._plans = {}

  def spec(self, dt: float, tableau=None) -> _ns2d.Ns2dSpec:
    n = self.grid.shape[0]
    return _ns2d.Ns2dSpec(n=n, dx=self.grid.step[0], dy=self.grid.step[1],
                          viscosity=float(self.viscosity), drag=float(self.drag),
                          smooth=self.smooth is not None, dt=float(dt),
                          forcing=self._forcing_entries, tableau=tableau)

  def plan(self, dt: float, batch: int = 1, device=None, tableau=None) -> _ns2d.Ns2dPlan:
    """The GPU plan of `batch` instances stepped with the low-storage CN-RK
    `tableau` ((alphas, betas, gammas); None: crank_nicolson_rk4)."""
    if device is None:
      device = torch.device('cuda', torch.cuda.current_device())
    key = (float(dt), int(batch), device.index or 0, tableau)
    if key not in self._plans:
      self._plans[key] = _ns2d.Ns2dPlan(self.spec(dt, tableau), batch, device)
    return self._plans[key]

  def explicit_terms(self, vorticity_hat):
    """spectral/equations.py:160-184 on the GPU (any leading batch dim)."""
    w = vorticity_hat
    batch = 1 if w.ndim == 2 else int(w.shape[0])
    dev = w.device if isinstance(w, torch.Tensor) and w.is_cuda else None
    return self.plan(0.0, batch, dev).explicit_terms(w)

  def implicit_terms(self, vorticity_hat):
    """L * w_hat (spectral/equations.py:186-187); element-wise."""
    lt = torch.as_tensor(self.linear_term, device=vorticity_hat.device)
    return lt * vorticity_hat

  def implicit_solve(self, vorticity_hat, time_step):
    """w_hat / (1 - dt L) (spectral/equations.py:189-190); element-wise."""
    lt = torch.as_tensor(self.linear_term, device=vorticity_hat.device)
    return 1 / (1 - time_step * lt) * vorticity_hat


def ForcedNavierStokes2D(viscosity, grid, smooth):  # pylint: disable=invalid-name
  """spectral/equations.py:194-220: Kolmogorov k=4 at offsets (0, 0), drag 0.1."""
  wave_number = 4
  offsets = ((0, 0), (0, 0))
  forcing_fn = lambda grid: forcings.kolmogorov_forcing(  # noqa: E731
      grid, k=wave_number, offsets=offsets)
  return NavierStokes2D(viscosity, grid, drag=0.1, smooth=smooth,
                        forcing_fn=forcing_fn)
