"""The GPU path against the committed golden fixtures (tests/golden/, the
oracle's fp32 outputs on seeded inputs; no oracle call at run time).

Tolerances: single operations and one step, max-norm relative error
<= 1e-5; several steps, relative L2 <= 1e-4.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
TWO_PI = 2 * np.pi


def _load(name):
  with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as f:
    return {k: f[k] for k in f.files}


def _maxrel(a, b):
  return float(np.abs(a - b).max() / np.abs(b).max())


def _rel_l2(a, b):
  return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _dev():
  return torch.device('cuda', 0)


def _grid(n):
  import jax_cfd_amd.base as cfd
  return cfd.grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))


def _state(u, v, n):
  import jax_cfd_amd.base as cfd
  g = _grid(n)
  bc = cfd.boundaries.periodic_boundary_conditions(2)
  return tuple(cfd.grids.GridVariable(cfd.grids.GridArray(
      torch.as_tensor(x, device=_dev()), off, g), bc) for x, off in zip((u, v), g.cell_faces))


FVM = ['fvm_van_leer_32_forced', 'fvm_van_leer_limiters_32_unforced', 'fvm_van_leer_64_forced']


@pytest.mark.parametrize('name', FVM)
def test_fvm_golden(name):
  import jax_cfd_amd.base as cfd
  d = _load(name)
  n, dt, forced = int(d['n']), float(d['dt']), bool(d['forced'])
  g = _grid(n)
  convect = None
  if '_van_leer_limiters_' not in name:
    convect = lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v)  # noqa: E731
  forcing = (cfd.forcings.simple_turbulence_forcing(g, 1.0, 4, -0.1, 'kolmogorov')
             if forced else None)
  kw = dict(convect=convect, forcing=forcing)
  print(f'golden {name}', flush=True)
  s = _state(d['u0'], d['v0'], n)
  k = cfd.equations.navier_stokes_explicit_terms(1.0, 1e-3, dt, g, **kw)(s)
  for got, key in zip(k, ('explicit_u', 'explicit_v')):
    err = _maxrel(got.data.cpu().numpy(), d[key])
    print(f'  {key}: {err:.2e}')
    assert err <= 1e-5
  pin = _state(d['u0'] + 0.1 * d['explicit_u'], d['v0'] + 0.1 * d['explicit_v'], n)
  p = cfd.pressure.projection(pin, cfd.pressure.solve_fast_diag)
  for got, key in zip(p, ('projected_u', 'projected_v')):
    err = _maxrel(got.data.cpu().numpy(), d[key])
    print(f'  {key}: {err:.2e}')
    assert err <= 1e-5
  step = cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, dt, g, **kw)
  one = step(s)
  for got, key in zip(one, ('u1', 'v1')):
    err = _maxrel(got.data.cpu().numpy(), d[key])
    print(f'  {key}: {err:.2e}')
    assert err <= 1e-5
  five = cfd.funcutils.repeated(step, 5)(s)
  for got, key in zip(five, ('u5', 'v5')):
    err = _rel_l2(got.data.cpu().numpy(), d[key])
    print(f'  {key} (rel L2): {err:.2e}')
    assert err <= 1e-4


@pytest.mark.parametrize('name', ['spectral_32_forced', 'spectral_64_drag'])
def test_spectral_golden(name):
  import jax_cfd_amd.spectral as sp
  d = _load(name)
  n, forced = int(d['n']), bool(d['forced'])
  g = _grid(n)
  eq = (sp.equations.ForcedNavierStokes2D(1e-3, g, True) if forced
        else sp.equations.NavierStokes2D(1e-3, g, drag=0.1))
  w0 = torch.as_tensor(d['w0'], device=_dev())
  print(f'golden {name}', flush=True)
  err = _maxrel(eq.explicit_terms(w0).cpu().numpy(), d['explicit'])
  print(f'  explicit: {err:.2e}')
  assert err <= 1e-5
  step = sp.time_stepping.crank_nicolson_rk4(eq, 1e-3)
  err = _maxrel(step(w0).cpu().numpy(), d['w1'])
  print(f'  w1: {err:.2e}')
  assert err <= 1e-5
  from jax_cfd_amd.base import funcutils
  err = _rel_l2(funcutils.repeated(step, 10)(w0).cpu().numpy(), d['w10'])
  print(f'  w10 (rel L2): {err:.2e}')
  assert err <= 1e-4


def test_initial_velocity_golden():
  import jax_cfd_amd.base as cfd
  d = _load('initial_velocity_32')
  v = cfd.initial_conditions.filtered_velocity_field(
      int(d['seed']), _grid(32), float(d['vmax']), int(d['peak']), device=_dev())
  for got, key in zip(v, ('u', 'v')):
    err = _maxrel(got.data.cpu().numpy(), d[key])
    print(f'  initial {key}: {err:.2e}')
    assert err <= 1e-5
