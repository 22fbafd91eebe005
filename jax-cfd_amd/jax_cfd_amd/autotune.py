"""Tile-shape and FFT-radix autotuner for the FVM step kernels.

The step kernels are built for a menu of tile shapes (rows per workgroup of
the stencil and projection kernels, spectrum columns per workgroup of the
Poisson column kernel) and two FFT radix schedules (largest radix 8 or 4).
Which combination is fastest depends on the grid size and on how many
instances a launch carries: one 256^2 field wants many small tiles to fill
the chip, 64 of them want larger tiles for fewer, longer workgroups.  The
tuner times every menu entry of a plan on the device with HIP events
(jcfd_fvm_time_kernels on the plan's own stream), keeps the fastest tile of
each kernel class per radix, and picks the radix with the fastest step.

Tile shapes never change results (they regroup the same per-cell and
per-column arithmetic).  The radix schedule changes the FFTs' rounding, so
`radices=(8,)` (the default) keeps results bit-identical to untuned plans;
pass `radices=(8, 4)` to let the tuner choose the schedule too.

Choices are cached per (device, solver, n, batch, radices) in a JSON file
(JCFD_TUNE_CACHE, default ~/.cache/jax_cfd_amd/tiles.json).
"""
from __future__ import annotations

import json
import os
import threading
from typing import Dict, Optional, Sequence

import torch

FVM_MENU = {
    'rows_forward': (1, 2, 4, 8),
    'poisson_cols': (1, 2, 4, 8, 16),
    'rows_project': (1, 2, 4, 8),
}
_CLASSES = tuple(FVM_MENU)
_lock = threading.Lock()


def cache_path() -> str:
  return os.environ.get('JCFD_TUNE_CACHE',
                        os.path.join(os.path.expanduser('~'), '.cache', 'jax_cfd_amd',
                                     'tiles.json'))


def _load_cache(path: str) -> Dict[str, dict]:
  try:
    with open(path) as f:
      data = json.load(f)
    return data if isinstance(data, dict) else {}
  except (OSError, ValueError):
    return {}


def _store_cache(path: str, key: str, value: dict) -> None:
  with _lock:
    data = _load_cache(path)
    data[key] = value
    os.makedirs(os.path.dirname(path) or '.', exist_ok=True)
    tmp = path + '.tmp'
    with open(tmp, 'w') as f:
      json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _device_name(device: torch.device) -> str:
  try:
    return torch.cuda.get_device_name(device)
  except Exception:  # no device query available
    return 'unknown'


def fvm_key(plan, radices: Sequence[int]) -> str:
  s = plan.spec
  return (f'{_device_name(plan.device)}|fvm|n={s.n}|batch={plan.batch}|'
          f'convect={s.convect}|radices={",".join(map(str, radices))}')


def select(timings: Dict[int, Dict[str, Dict[int, float]]]) -> dict:
  """Pick tiles from timings[radix][kernel_class][tile] = ms per launch:
  the fastest tile of each class per radix, then the radix whose summed
  best times are smallest (ties keep the earlier radix)."""
  best = None
  for radix, per_class in timings.items():
    choice, total = {}, 0.0
    for cls in _CLASSES:
      options = per_class.get(cls, {})
      if not options:
        raise ValueError(f'no timing for {cls} at radix {radix}')
      tile = min(options, key=lambda t: (options[t], t))
      choice[cls] = int(tile)
      total += options[tile]
    if best is None or total < best[0]:
      best = (total, dict(choice, radix=int(radix)))
  if best is None:
    raise ValueError('no timings')
  return dict(best[1], ms_per_step=best[0])


def measure_fvm(plan, u: torch.Tensor, v: torch.Tensor, radices: Sequence[int] = (8,),
                reps: int = 20, trials: int = 2) -> Dict[int, Dict[str, Dict[int, float]]]:
  """timings[radix][class][tile] = ms per launch for every menu entry the
  plan's size has.  Configuration t sets tile t on every class whose menu
  has it for this size (the plan's current tile elsewhere) and times the
  three classes together (best of `trials` runs of `reps` launches)."""
  default = plan.tiles
  out: Dict[int, Dict[str, Dict[int, float]]] = {}
  values = sorted(set(t for menu in FVM_MENU.values() for t in menu))
  try:
    for radix in radices:
      per_class: Dict[str, Dict[int, float]] = {c: {} for c in _CLASSES}
      for t in values:
        cfg = {c: t if _valid(plan, default, c, t, radix) else default[c] for c in _CLASSES}
        plan.set_tiles(cfg['rows_forward'], cfg['poisson_cols'], cfg['rows_project'], radix)
        ms = None
        for _ in range(trials):
          m = plan.time_kernels(u, v, reps)
          ms = m if ms is None else {k: min(ms[k], m[k]) for k in ms}
        for c, kname in zip(_CLASSES, plan.KERNELS):
          per_class[c][cfg[c]] = min(per_class[c].get(cfg[c], float('inf')), ms[kname])
      out[int(radix)] = per_class
  finally:
    plan.set_tiles(default['rows_forward'], default['poisson_cols'], default['rows_project'],
                   default['radix'])
  return out


def _valid(plan, default, cls, tile, radix) -> bool:
  """Whether the plan's size has `tile` on the menu of kernel class `cls`."""
  if tile not in FVM_MENU[cls]:
    return False
  cfg = dict(default)
  cfg[cls] = tile
  try:
    plan.set_tiles(cfg['rows_forward'], cfg['poisson_cols'], cfg['rows_project'], radix)
    return True
  except NotImplementedError:
    return False


def tune_fvm(plan, u: Optional[torch.Tensor] = None, v: Optional[torch.Tensor] = None,
             radices: Sequence[int] = (8,), reps: int = 20, use_cache: bool = True) -> dict:
  """Tune `plan` in place and return the chosen tiles (with the timings).
  u, v: fields of the plan's shape to time on (random if omitted)."""
  key = fvm_key(plan, radices)
  path = cache_path()
  if use_cache:
    hit = _load_cache(path).get(key)
    if hit is not None:
      plan.set_tiles(hit['rows_forward'], hit['poisson_cols'], hit['rows_project'], hit['radix'])
      return dict(hit, cached=True)
  if u is None or v is None:
    gen = torch.Generator(device=plan.device).manual_seed(0)
    u = torch.randn(plan.shape, device=plan.device, generator=gen)
    v = torch.randn(plan.shape, device=plan.device, generator=gen)
  timings = measure_fvm(plan, u, v, radices, reps)
  choice = select(timings)
  plan.set_tiles(choice['rows_forward'], choice['poisson_cols'], choice['rows_project'],
                 choice['radix'])
  result = dict(choice, timings={str(r): {c: {str(t): ms for t, ms in d.items()}
                                          for c, d in pc.items()}
                                 for r, pc in timings.items()})
  if use_cache:
    _store_cache(path, key, {k: choice[k] for k in ('rows_forward', 'poisson_cols',
                                                   'rows_project', 'radix', 'ms_per_step')})
  return dict(result, cached=False)
