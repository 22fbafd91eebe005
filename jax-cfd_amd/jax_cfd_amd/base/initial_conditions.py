"""Initial velocity fields (jax_cfd/base/initial_conditions.py:40-166).

``filtered_velocity_field`` follows the reference: N(0, 1) noise per
component, spectral filter with a log-normal density, then three rounds of
(projection, rescale to the maximum speed).  The filter and the projections
run on the GPU.  The noise cannot be jax.random's threefry stream (JAX is not
a dependency); it is NumPy's PCG64 standard normal in fp32, seeded by the
integer ``rng_key`` (an int, or a sequence whose last entry is the seed).
"""
from __future__ import annotations

import concurrent.futures
import threading
from typing import Callable, Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from jax_cfd_amd.base import _fvm, boundaries, grids, pressure


def _seed(rng_key) -> int:
  if isinstance(rng_key, (int, np.integer)):
    return int(rng_key)
  arr = np.asarray(rng_key).ravel()
  return int(arr[-1])


def noise(seed: int, shape, components: int) -> np.ndarray:
  """Seeded fp32 standard-normal noise, one array per component."""
  rng = np.random.Generator(np.random.PCG64(seed))
  return np.stack([rng.standard_normal(shape, dtype=np.float32)
                   for _ in range(components)])


# Noise drawn ahead of use on worker threads (NumPy releases the GIL while it
# fills an array), straight into pinned host memory so the copy to the
# device does not block the caller.  Keyed by (seed, shape, components);
# the same values as noise().
_prefetched: Dict[tuple, concurrent.futures.Future] = {}
_prefetch_lock = threading.Lock()
_executor: Optional[concurrent.futures.ThreadPoolExecutor] = None


def _pinned_noise(seed: int, shape, components: int) -> torch.Tensor:
  out = torch.empty((components,) + tuple(shape), dtype=torch.float32, pin_memory=True)
  out.numpy()[...] = noise(seed, shape, components)
  return out


def prefetch_noise(seed: int, shape, components: int) -> None:
  """Start drawing noise(seed, shape, components) on a worker thread; the
  next filtered_velocity_field with that seed and shape uses it."""
  global _executor
  key = (int(seed), tuple(shape), int(components))
  with _prefetch_lock:
    if key in _prefetched:
      return
    if _executor is None:
      _executor = concurrent.futures.ThreadPoolExecutor(max_workers=4)
    _prefetched[key] = _executor.submit(_pinned_noise, *key)


def _noise_on_device(seed: int, shape, components: int, device) -> torch.Tensor:
  key = (int(seed), tuple(shape), int(components))
  with _prefetch_lock:
    fut = _prefetched.pop(key, None)
  host = fut.result() if fut is not None else _pinned_noise(*key)
  return host.to(device, non_blocking=True)


_multipliers: Dict[tuple, torch.Tensor] = {}


def _multiplier_on_device(grid: grids.Grid, peak_wavenumber: float, device) -> torch.Tensor:
  """spectral_density_multiplier on the device, built once per grid."""
  key = (grid.shape, grid.step, float(peak_wavenumber), str(device))
  if key not in _multipliers:
    _multipliers[key] = torch.as_tensor(spectral_density_multiplier(grid, peak_wavenumber),
                                        device=device)
  return _multipliers[key]


def _log_normal_pdf(x, mode, variance=.25):
  """initial_conditions.py:60-64."""
  mean = float(np.log(mode)) + variance
  logx = np.log(x)
  return np.exp(-(mean - logx) ** 2 / 2 / variance - logx)


def spectral_density_multiplier(grid: grids.Grid, peak_wavenumber: float) -> np.ndarray:
  """The filter of filter_utils.filter (filter_utils.py:25-42) on the rfft
  half plane: where(|k| > 0, lognormal(|k|) / |k|^(ndim-1), 0), fp32."""
  freqs = []
  for axis, (n, h) in enumerate(zip(grid.shape, grid.step)):
    if axis == grid.ndim - 1:
      k = np.arange(n // 2 + 1)
    else:
      k = np.concatenate([np.arange((n + 1) // 2), -np.arange(n // 2, 0, -1)])
    freqs.append((2 * np.pi * (k.astype(np.float32) / np.float32(h * n))).astype(np.float32))
  mesh = np.meshgrid(*freqs, indexing='ij')
  kmag = np.sqrt(sum(m ** 2 for m in mesh)).astype(np.float32)
  with np.errstate(divide='ignore', invalid='ignore'):
    dens = _log_normal_pdf(kmag, peak_wavenumber) / kmag ** (grid.ndim - 1)
  return np.where(kmag > 0, dens, 0.0).astype(np.float32)


def wrap_variables(var: Sequence, grid: grids.Grid, bcs: Sequence,
                   offsets: Optional[Sequence[Tuple[float, ...]]] = None):
  """initial_conditions.py:40-57 (no batch dimension)."""
  if offsets is None:
    offsets = grid.cell_faces
  return tuple(bc.impose_bc(grids.GridArray(u, o, grid))
               for u, o, bc in zip(var, offsets, bcs))


def _max_speed(v) -> torch.Tensor:
  return torch.sqrt(sum(u.data * u.data for u in v)).max()


def filtered_velocity_field(rng_key, grid: grids.Grid, maximum_velocity: float = 1,
                            peak_wavenumber: float = 3, iterations: int = 3,
                            device: Optional[torch.device] = None):
  """initial_conditions.py:71-121 with this project's seeded noise."""
  if device is None:
    device = torch.device('cuda', torch.cuda.current_device())
  plan = _fvm.plan_for(_fvm.spec_from_grid(grid), device)
  nz = _noise_on_device(_seed(rng_key), grid.shape, grid.ndim, device)
  mult = _multiplier_on_device(grid, peak_wavenumber, device)
  comps = [plan.spectral_filter(nz[c], mult) for c in range(grid.ndim)]
  bcs = [boundaries.periodic_boundary_conditions(grid.ndim)] * grid.ndim
  v = wrap_variables(comps, grid, bcs)
  for _ in range(iterations):
    v = pressure.projection(v)
    vmax = _max_speed(v)
    v = tuple(grids.GridVariable(maximum_velocity * u.array / vmax, u.bc) for u in v)
  return v


def initial_velocity_field(velocity_fns: Sequence[Callable], grid: grids.Grid,
                           velocity_bc: Optional[Sequence] = None,
                           pressure_solve: Callable = pressure.solve_fast_diag,
                           iterations: Optional[int] = None,
                           device: Optional[torch.device] = None):
  """initial_conditions.py:124-166: each component evaluated on its face
  mesh (fp32, as grid.eval_on_mesh), placed on the GPU, then projected
  `iterations` times (if given) with `pressure_solve` on the GPU."""
  if device is None:
    device = torch.device('cuda', torch.cuda.current_device())
  if velocity_bc is None:
    velocity_bc = (boundaries.periodic_boundary_conditions(grid.ndim),) * grid.ndim
  v = []
  for fn, offset, bc in zip(velocity_fns, grid.cell_faces, velocity_bc):
    arr = grid.eval_on_mesh(fn, offset)
    data = torch.as_tensor(np.asarray(arr.data), dtype=torch.float32).to(device)
    v.append(grids.GridVariable(grids.GridArray(data, offset, grid), bc))
  v = tuple(v)
  for _ in range(iterations or 0):
    v = pressure.projection(v, pressure_solve)
  return v
