"""Tile-shape and FFT-radix autotuner for the FVM and spectral step kernels.

The step kernels are built for a menu of tile shapes and two FFT radix
schedules (largest radix 8 or 4).  FVM plans tune the rows per workgroup of
the stencil kernel (16 and 32 select its marching form, n >= 256) and of the
projection kernel and the spectrum columns per workgroup of
the Poisson column kernel; spectral plans tune the packed-spectrum columns
per workgroup of their three column kernels and the rows per workgroup of
the nonlinear-product row kernel.  Which combination is fastest depends on
the grid size and on how many instances a launch carries: one 256^2 field
wants many small tiles to fill the chip, 64 of them want larger tiles for
fewer, longer workgroups.  The tuner times every menu entry of a plan on the
device with HIP events (the plan's time_kernels, on its own stream), keeps
the fastest value of each tunable per radix (kernel times weighted by
launches per step), and picks the radix with the fastest step.

Tile shapes never change results (they regroup the same per-cell, per-row
and per-column arithmetic).  The radix schedule changes the FFTs' rounding,
so `radices=(8,)` (the default) keeps results bit-identical to untuned
plans; pass `radices=(8, 4)` to let the tuner choose the schedule too.

Choices are cached per (device, solver, n, batch, radices) in a JSON file
(JCFD_TUNE_CACHE, default ~/.cache/jax_cfd_amd/tiles.json).
"""
from __future__ import annotations

import dataclasses
import json
import os
import threading
from typing import Callable, Dict, Optional, Sequence, Tuple

import torch

_lock = threading.Lock()


@dataclasses.dataclass(frozen=True)
class Tunables:
  """menu[param] = candidate values; weights[param] = {kernel: launches per
  step} (the kernels whose time the parameter sets)."""
  solver: str
  menu: Dict[str, Tuple[int, ...]]
  weights: Dict[str, Dict[str, int]]


FVM = Tunables(
    'fvm',
    menu={'rows_forward': (1, 2, 4, 8, 16, 32), 'poisson_cols': (1, 2, 4, 8, 16),
          'rows_project': (1, 2, 4, 8)},
    weights={'rows_forward': {'fvm_rows_forward': 1},
             'poisson_cols': {'fvm_poisson_cols': 1},
             'rows_project': {'fvm_rows_project': 1}})
NS2D = Tunables(
    'spectral',
    menu={'cols': (1, 2, 4, 8, 16), 'rows': (1, 2, 4)},
    weights={'cols': {'ns_cols_inverse': 1, 'ns_cols_update_inverse': 4, 'ns_cols_update': 1},
             'rows': {'ns_rows': 5}})
FVM_MENU = FVM.menu


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


def _device_name(device) -> str:
  """ISA and compute-unit count (e.g. 'gfx950x256'): read from the HSA agent,
  so the key is the same whether or not the marketing-name table is present
  (it is absent under some profilers)."""
  try:
    props = torch.cuda.get_device_properties(device)
    return f"{props.gcnArchName.split(':')[0]}x{props.multi_processor_count}"
  except Exception:  # no device query available
    return 'unknown'


def _key(plan, t: Tunables, radices: Sequence[int]) -> str:
  s = plan.spec
  extra = f'|convect={s.convect}' if t.solver == 'fvm' else ''
  return (f'{_device_name(plan.device)}|{t.solver}|n={s.n}|batch={plan.batch}{extra}|'
          f'radices={",".join(map(str, radices))}')


def select(timings: Dict[int, Dict[str, Dict[int, float]]]) -> dict:
  """Pick from timings[radix][param][value] = ms per step spent in the
  kernels `param` sets: the fastest value of each parameter per radix, then
  the radix whose summed best times are smallest (ties keep the smaller
  value and the earlier radix)."""
  best = None
  for radix, per_param in timings.items():
    choice, total = {}, 0.0
    for param, options in per_param.items():
      if not options:
        raise ValueError(f'no timing for {param} at radix {radix}')
      value = min(options, key=lambda v: (options[v], v))
      choice[param] = int(value)
      total += options[value]
    if best is None or total < best[0]:
      best = (total, dict(choice, radix=int(radix)))
  if best is None:
    raise ValueError('no timings')
  return dict(best[1], ms_per_step=best[0])


def _valid(plan, default: dict, param: str, value: int, radix: int) -> bool:
  """Whether the plan's size has `value` on the menu of `param`."""
  cfg = dict(default, radix=radix)
  cfg[param] = value
  try:
    plan.set_tiles(**cfg)
    return True
  except NotImplementedError:
    return False


def measure(plan, t: Tunables, time_fn: Callable[[], Dict[str, float]],
            radices: Sequence[int] = (8,), trials: int = 2) -> Dict[int, Dict[str, Dict[int, float]]]:
  """timings[radix][param][value] for every menu entry the plan's size has.
  Configuration v sets value v on every parameter whose menu has it for this
  size (the plan's current value elsewhere); time_fn() times all kernels at
  once (best of `trials`)."""
  default = plan.tiles
  params = tuple(t.menu)
  values = sorted(set(v for menu in t.menu.values() for v in menu))
  out: Dict[int, Dict[str, Dict[int, float]]] = {}
  try:
    for radix in radices:
      per_param: Dict[str, Dict[int, float]] = {p: {} for p in params}
      for v in values:
        cfg = {p: v if (v in t.menu[p] and _valid(plan, default, p, v, radix)) else default[p]
               for p in params}
        plan.set_tiles(**cfg, radix=radix)
        ms = None
        for _ in range(trials):
          m = time_fn()
          ms = m if ms is None else {k: min(ms[k], m[k]) for k in ms}
        for p in params:
          cost = sum(w * ms[k] for k, w in t.weights[p].items())
          per_param[p][cfg[p]] = min(per_param[p].get(cfg[p], float('inf')), cost)
      out[int(radix)] = per_param
  finally:
    plan.set_tiles(**default)
  return out


def _tune(plan, t: Tunables, time_fn, radices, use_cache) -> dict:
  key = _key(plan, t, radices)
  path = cache_path()
  if use_cache:
    hit = _load_cache(path).get(key)
    if hit is not None:
      plan.set_tiles(**{p: hit[p] for p in t.menu}, radix=hit['radix'])
      return dict(hit, cached=True)
  timings = measure(plan, t, time_fn, radices)
  choice = select(timings)
  plan.set_tiles(**{p: choice[p] for p in t.menu}, radix=choice['radix'])
  if use_cache:
    _store_cache(path, key, choice)
  return dict(choice, cached=False,
              timings={str(r): {p: {str(v): ms for v, ms in d.items()} for p, d in pp.items()}
                       for r, pp in timings.items()})


def tune_fvm(plan, u: Optional[torch.Tensor] = None, v: Optional[torch.Tensor] = None,
             radices: Sequence[int] = (8,), reps: int = 20, use_cache: bool = True) -> dict:
  """Tune an FvmPlan in place; returns the choice (and timings when measured).
  u, v: fields of the plan's shape to time on (random if omitted)."""
  if u is None or v is None:
    gen = torch.Generator(device=plan.device).manual_seed(0)
    u = torch.randn(plan.shape, device=plan.device, generator=gen)
    v = torch.randn(plan.shape, device=plan.device, generator=gen)
  return _tune(plan, FVM, lambda: plan.time_kernels(u, v, reps), radices, use_cache)


def tune_ns2d(plan, w: Optional[torch.Tensor] = None, radices: Sequence[int] = (8,),
              reps: int = 10, use_cache: bool = True) -> dict:
  """Tune an Ns2dPlan in place; w: a state of the plan's shape (small random
  vorticity spectrum if omitted)."""
  if w is None:
    gen = torch.Generator(device=plan.device).manual_seed(0)
    w = (torch.randn(plan.shape, device=plan.device, generator=gen)
         + 1j * torch.randn(plan.shape, device=plan.device, generator=gen)).to(torch.complex64)
    w = w * 1e-3
  return _tune(plan, NS2D, lambda: plan.time_kernels(w, reps), radices, use_cache)
