"""Spectral helpers (jax_cfd/spectral/utils.py:128-203).

Host-side constants (the 2/3 mask) and element-wise helpers; the step itself
runs in the fused kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from jax_cfd_amd.base import grids


def brick_wall_filter_2d(grid: grids.Grid) -> np.ndarray:
  """2/3 rule mask, fp32 (n, m//2 + 1) (spectral/utils.py:128-134)."""
  n, m = grid.shape
  f = np.zeros((n, m // 2 + 1), np.float32)
  f[:int(2 / 3 * n) // 2, :int(2 / 3 * (m // 2 + 1))] = 1
  f[-int(2 / 3 * n) // 2:, :int(2 / 3 * (m // 2 + 1))] = 1
  return f


def vorticity_to_velocity(grid: grids.Grid):
  """spectral/utils.py:154-189: w_hat -> (vx_hat, vy_hat), element-wise."""
  kx, ky = grid.rfft_mesh()
  two_pi_i = 2 * np.pi * 1j
  laplace = (two_pi_i ** 2 * (abs(kx) ** 2 + abs(ky) ** 2)).astype(np.complex64)
  laplace[0, 0] = 1

  def ret(vorticity_hat):
    dev = vorticity_hat.device
    lap = torch.as_tensor(laplace, device=dev)
    psi_hat = -1 / lap * vorticity_hat
    vxhat = two_pi_i * torch.as_tensor(ky, device=dev) * psi_hat
    vyhat = -two_pi_i * torch.as_tensor(kx, device=dev) * psi_hat
    return vxhat, vyhat
  return ret


def spectral_curl_2d(mesh, velocity_hat):
  """spectral/utils.py:199-203."""
  kx, ky = mesh
  uhat, vhat = velocity_hat
  dev = uhat.device
  return 2j * np.pi * (vhat * torch.as_tensor(kx, device=dev)
                       - uhat * torch.as_tensor(ky, device=dev))
