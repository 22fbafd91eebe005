"""Pins the differentiable torch restatement used as the jax.grad stand-in."""
import numpy as np
import torch

from oracle import fvm as ofvm
from oracle import grid_ops as go
from oracle import initial_conditions as oic
from oracle import torch_fvm as tf
from oracle.grid_ops import Grid


def _setup(n=16, forced=True):
  grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  v0 = oic.filtered_velocity_field(3, grid, 2.0, 2, dtype=np.float64)
  dt = ofvm.stable_time_step(2.0, 0.5, 1e-3, grid)
  forcing = go.Forcing(1.0, 2, -0.1) if forced else None
  ostep = ofvm.FVMStep(grid, dt=dt, viscosity=1e-3, forcing=forcing)
  fx = forcing.constant_terms(grid, np.float64)[0][0] if forced else None
  tstep = tf.TorchFVMStep(grid, dt, 1e-3, 1.0, -0.1 if forced else None, fx)
  return grid, v0, ostep, tstep


def test_torch_step_equals_numpy_oracle():
  grid, v0, ostep, tstep = _setup()
  ref = ostep.rollout(v0, 3)
  u, v = (torch.tensor(x.data) for x in v0)
  for _ in range(3):
    u, v = tstep(u, v)
  np.testing.assert_allclose(u.numpy(), ref[0].data, rtol=0, atol=1e-12)
  np.testing.assert_allclose(v.numpy(), ref[1].data, rtol=0, atol=1e-12)


def test_torch_gradient_matches_central_differences():
  """Directional derivatives of a curl loss after 3 steps (fp64)."""
  grid, v0, ostep, tstep = _setup()
  u0, v0t = (torch.tensor(x.data, requires_grad=True) for x in v0)
  rng = np.random.default_rng(0)
  target = torch.tensor(rng.standard_normal(grid.shape))

  def loss(u, v):
    for _ in range(3):
      u, v = tstep(u, v)
    return ((tf.curl_2d(u, v, grid) - target) ** 2).mean()

  L = loss(u0, v0t)
  gu, gv = torch.autograd.grad(L, (u0, v0t))
  for _ in range(3):
    du = torch.tensor(rng.standard_normal(grid.shape))
    dv = torch.tensor(rng.standard_normal(grid.shape))
    eps = 1e-6
    with torch.no_grad():
      fd = (loss(u0 + eps * du, v0t + eps * dv) - loss(u0 - eps * du, v0t - eps * dv)) / (2 * eps)
    ad = (gu * du).sum() + (gv * dv).sum()
    assert abs(float(fd - ad)) <= 1e-6 * abs(float(ad)) + 1e-12


def test_jax_abs_convention():
  x = torch.tensor([0.0, -2.0, 3.0], dtype=torch.float64, requires_grad=True)
  (g,) = torch.autograd.grad(tf.jabs(x).sum(), x)
  np.testing.assert_array_equal(g.numpy(), [1.0, -1.0, 1.0])
