"""Selection logic and cache of the tile autotuner, on a stand-in plan whose
kernel timings are a known function of the tiles (no GPU needed)."""
import json

import pytest

from jax_cfd_amd import autotune


class FakePlan:
  """Mimics FvmPlan's tile API: poisson_cols 16 is off this size's menu;
  rows_forward is fastest at 2, poisson_cols at 4, rows_project at 1; the
  radix-4 schedule costs 10% more."""
  KERNELS = ('fvm_rows_forward', 'fvm_poisson_cols', 'fvm_rows_project')

  class spec:
    n, convect = 64, 0

  batch = 3
  device = 'cpu'
  shape = (3, 64, 64)

  def __init__(self):
    self._tiles = dict(rows_forward=4, poisson_cols=8, rows_project=4, radix=8)
    self.calls = 0

  @property
  def tiles(self):
    return dict(self._tiles)

  def set_tiles(self, rows_forward, poisson_cols, rows_project, radix=8):
    if poisson_cols == 16 or radix not in (4, 8):
      raise NotImplementedError('not on the menu')
    self._tiles = dict(rows_forward=rows_forward, poisson_cols=poisson_cols,
                       rows_project=rows_project, radix=radix)

  def time_kernels(self, u, v, reps):
    self.calls += 1
    t = self._tiles
    f = 1.1 if t['radix'] == 4 else 1.0
    return {'fvm_rows_forward': f * (1 + abs(t['rows_forward'] - 2)),
            'fvm_poisson_cols': f * (1 + abs(t['poisson_cols'] - 4)),
            'fvm_rows_project': f * (1 + abs(t['rows_project'] - 1))}


def test_select_picks_fastest_per_class_and_radix():
  timings = {8: {'rows_forward': {1: 2.0, 2: 1.0}, 'poisson_cols': {4: 1.0, 8: 3.0},
                 'rows_project': {1: 1.0}},
             4: {'rows_forward': {1: 0.5}, 'poisson_cols': {4: 0.5}, 'rows_project': {1: 0.5}}}
  c = autotune.select(timings)
  assert (c['rows_forward'], c['poisson_cols'], c['rows_project'], c['radix']) == (1, 4, 1, 4)
  assert c['ms_per_step'] == pytest.approx(1.5)
  # ties keep the smaller tile and the earlier radix
  tie = {8: {'rows_forward': {2: 1.0, 1: 1.0}, 'poisson_cols': {8: 1.0},
             'rows_project': {4: 1.0}},
         4: {'rows_forward': {1: 1.0}, 'poisson_cols': {8: 1.0}, 'rows_project': {4: 1.0}}}
  c = autotune.select(tie)
  assert c['rows_forward'] == 1 and c['radix'] == 8


def test_tune_sets_plan_and_caches(tmp_path, monkeypatch):
  cache = tmp_path / 'tiles.json'
  monkeypatch.setenv('JCFD_TUNE_CACHE', str(cache))
  plan = FakePlan()
  res = autotune.tune_fvm(plan, u=object(), v=object(), radices=(8, 4), reps=1)
  assert not res['cached']
  assert plan.tiles == dict(rows_forward=2, poisson_cols=4, rows_project=1, radix=8)
  # the menu entry off this size's menu was never timed
  assert '16' not in res['timings']['8']['poisson_cols']
  stored = json.loads(cache.read_text())
  assert len(stored) == 1
  # a fresh plan of the same key takes the cached choice without timing
  plan2 = FakePlan()
  res2 = autotune.tune_fvm(plan2, u=object(), v=object(), radices=(8, 4), reps=1)
  assert res2['cached'] and plan2.calls == 0
  assert plan2.tiles == plan.tiles


def test_measure_restores_tiles_on_error():
  plan = FakePlan()

  def boom(*a):
    raise RuntimeError('device error')
  plan.time_kernels = boom
  with pytest.raises(RuntimeError):
    autotune.measure(plan, autotune.FVM, lambda: plan.time_kernels(None, None, 1), radices=(8,))
  assert plan.tiles == dict(rows_forward=4, poisson_cols=8, rows_project=4, radix=8)


class FakeNsPlan:
  """Spectral stand-in: the column tile sets three kernels launched 1, 4 and
  1 times per step; ns_cols_inverse alone prefers 4, the 4x-launched
  ns_cols_update_inverse prefers 8, so the weighted choice is 8."""

  class spec:
    n = 128

  batch = 2
  device = 'cpu'

  def __init__(self):
    self._tiles = dict(cols=8, rows=2, radix=8)

  @property
  def tiles(self):
    return dict(self._tiles)

  def set_tiles(self, cols, rows, radix=8):
    if rows == 8 or rows == 16:
      raise NotImplementedError('row tile off the menu')
    self._tiles = dict(cols=cols, rows=rows, radix=radix)

  def time_kernels(self, w, reps):
    t = self._tiles
    return {'ns_cols_inverse': 1 + abs(t['cols'] - 4),
            'ns_rows': 1 + abs(t['rows'] - 1),
            'ns_cols_update_inverse': 1 + abs(t['cols'] - 8),
            'ns_cols_update': 1.0}


def test_spectral_choice_weights_kernels_by_launches(tmp_path, monkeypatch):
  monkeypatch.setenv('JCFD_TUNE_CACHE', str(tmp_path / 'tiles.json'))
  plan = FakeNsPlan()
  res = autotune.tune_ns2d(plan, w=object(), reps=1)
  assert plan.tiles == dict(cols=8, rows=1, radix=8)
  assert res['timings']['8']['rows'].keys() == {'1', '2', '4'}
