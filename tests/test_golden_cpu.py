"""The oracle still produces the committed golden fixtures (tests/golden/,
written by tests/golden/make_fixtures.py): a change to the oracle that moves
its results fails here.  Tolerance: max-norm relative 1e-6 (the same NumPy
arithmetic; only a different FFT code path on another host could move the
last bits)."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _load(name):
  with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as f:
    return {k: f[k] for k in f.files}


def test_fixtures_present_and_plain_arrays():
  names = sorted(f for f in os.listdir(HERE) if f.endswith('.npz'))
  assert len(names) == 6
  for n in names:
    d = _load(n[:-4])
    for a in d.values():
      assert a.dtype.kind in 'fci'


@pytest.fixture(scope='module')
def regenerated():
  import importlib.util
  spec = importlib.util.spec_from_file_location('make_fixtures', os.path.join(HERE, 'make_fixtures.py'))
  mod = importlib.util.module_from_spec(spec)
  spec.loader.exec_module(mod)
  return mod.fixtures()


def test_oracle_reproduces_fixtures(regenerated):
  for name, arrays in regenerated.items():
    saved = _load(name)
    assert set(saved) == set(arrays), name
    for key, a in arrays.items():
      b = saved[key]
      assert a.dtype == b.dtype and a.shape == b.shape, (name, key)
      scale = max(float(np.abs(b).max()), 1e-30)
      err = float(np.abs(a - b).max()) / scale
      assert err <= 1e-6, (name, key, err)
