"""Host-side pieces of the server that need no GPU."""
import numpy as np
import pytest

from jax_cfd_amd import serving


def test_poisson_requests_seeded_and_mixed():
  a = serving.poisson_requests(400, rate=200.0, seed=0)
  b = serving.poisson_requests(400, rate=200.0, seed=0)
  assert [(r.solver, r.n, r.steps, r.seed, r.arrival) for r in a] == \
         [(r.solver, r.n, r.steps, r.seed, r.arrival) for r in b]
  arr = np.array([r.arrival for r in a])
  assert arr[0] == 0 and np.all(np.diff(arr) >= 0)
  assert abs(len(a) / arr[-1] - 200.0) < 30.0  # mean rate
  assert {r.solver for r in a} == {'fvm', 'spectral'}
  assert {r.n for r in a} == {64, 128, 256, 512}
  steps = np.array([r.steps for r in a])
  assert steps.min() >= 100 and steps.max() <= 2000


def test_summarize_percentiles():
  res = [serving.RolloutResult(i, 'fvm', 64, 100, 'done', None, 0.0, 0.0, float(i + 1))
         for i in range(100)]
  res.append(serving.RolloutResult(100, 'fvm', 512, 100, 'rejected'))
  s = serving.summarize(res, 10.0)
  assert s['done'] == 100 and s['rejected'] == 1
  assert s['cell_steps'] == 100 * 64 * 64 * 100
  assert abs(s['p50_latency_s'] - 50.5) < 1e-9
  assert 99.0 <= s['p99_latency_s'] <= 100.0


def test_invalid_requests_raise_before_any_work():
  import pytest
  srv = serving.RolloutServer.__new__(serving.RolloutServer)  # no device needed
  with pytest.raises(ValueError):
    serving.RolloutServer.run(srv, [serving.RolloutRequest(0, 'fvm', 64, -1, 0)])
  with pytest.raises(ValueError):
    serving.RolloutServer.run(srv, [serving.RolloutRequest(0, 'fvm', 64, 10, 0, save_every=-2)])


def test_unknown_group_solver_raises_before_any_device_work():
  from jax_cfd_amd import serving
  with pytest.raises(ValueError):
    serving.RolloutServer(group=('lbm',))
