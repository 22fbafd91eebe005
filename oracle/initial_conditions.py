"""filtered_velocity_field restated on the CPU with this project's seeded noise.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

jax_cfd/base/initial_conditions.py:71-121 draws N(0, 1) noise per component
with jax.random (threefry), filters it with a log-normal spectral density
(filter_utils.py:25-42, initial_conditions.py:60-64, 95-96), then applies
(projection, rescale to max speed) three times.  JAX is not available, so the
noise comes from ``noise(seed, shape)`` -- NumPy's PCG64 generator, standard
normal, fp32 -- which the product path uses too (it is input data, generated
on the host).  Everything after the noise follows the reference.

Pinning (tests/test_oracle_ic.py): divergence after the 3 projections and the
exact maximum speed (initial_conditions_test.py:39-60); the spectral filter
keeps the energy at |k| near the peak wavenumber.  Bitwise equality with
jax.random streams is impossible without JAX ("parity unpinned" for the noise
itself; both sides of every comparison use the same noise).
"""
from __future__ import annotations

import numpy as np

from oracle import fvm
from oracle.grid_ops import A, Grid


def noise(seed: int, shape, components: int = 2) -> np.ndarray:
  """Seeded standard-normal noise, fp32, one array per velocity component."""
  rng = np.random.Generator(np.random.PCG64(seed))
  return np.stack([rng.standard_normal(shape, dtype=np.float32)
                   for _ in range(components)])


def log_normal_pdf(x, mode, variance=.25):
  """initial_conditions.py:60-64."""
  mean = float(np.log(mode)) + variance
  logx = np.log(x)
  return np.exp(-(mean - logx) ** 2 / 2 / variance - logx)


def angular_frequency_magnitude(grid: Grid, dtype=np.float32):
  """filter_utils.py:25-29 (fp32 frequencies, 2*pi*fftfreq)."""
  freqs = []
  for n, h in zip(grid.shape, grid.step):
    k = np.concatenate([np.arange((n + 1) // 2), -np.arange(n // 2, 0, -1)])
    freqs.append((2 * np.pi * (k.astype(dtype) / dtype(h * n))).astype(dtype))
  mesh = np.meshgrid(*freqs, indexing='ij')
  return np.sqrt(sum(m.astype(dtype) ** 2 for m in mesh)).astype(dtype)


def spectral_filter(grid: Grid, peak_wavenumber: float, dtype=np.float32):
  """where(k > 0, lognormal(k, peak) / k^(ndim-1), 0) on the full fftn grid."""
  k = angular_frequency_magnitude(grid, dtype)
  with np.errstate(divide='ignore', invalid='ignore'):
    dens = log_normal_pdf(k, peak_wavenumber) / k ** (grid.ndim - 1)
  return np.where(k > 0, dens, 0.0).astype(dtype)


def filter_noise(noise_c: np.ndarray, filt: np.ndarray) -> np.ndarray:
  """filter_utils.py:32-42: ifftn(fftn(array) * filters).real."""
  return np.fft.ifftn(np.fft.fftn(noise_c) * filt).real.astype(noise_c.dtype)


def max_speed(v):
  """initial_conditions.py:67-68."""
  return np.sqrt(sum(u.data ** 2 for u in v)).max()


def filtered_velocity_field(seed: int, grid: Grid, maximum_velocity=1.0,
                            peak_wavenumber=3.0, iterations=3,
                            dtype=np.float32):
  """initial_conditions.py:71-121 with ``noise(seed)`` in place of jax.random."""
  nz = noise(seed, grid.shape, grid.ndim).astype(dtype)
  filt = spectral_filter(grid, peak_wavenumber, dtype)
  comps = [filter_noise(nz[c], filt) for c in range(grid.ndim)]
  v = tuple(A(c, o) for c, o in zip(comps, grid.cell_faces))
  for _ in range(iterations):
    v = fvm.projection(v, grid)
    vmax = max_speed(v)
    v = tuple(A(maximum_velocity * u.data / vmax, u.offset) for u in v)
  return v


def vorticity_hat_from_velocity(v, grid: Grid):
  """rfftn(curl_2d(v)) -- how the spectral tests and notebooks start
  (spectral/equations_test.py:216-219)."""
  from oracle import grid_ops as go
  w = go.curl_2d(v, grid).data
  return np.fft.rfftn(w)
