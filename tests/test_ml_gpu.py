"""GPU parity of the learned-interpolation path against the oracle (oracle/ml.py).

Seeded haiku-default tower weights (the reference ships no trained ones).
Tolerances (fp32): the tower's logits and the stencil coefficients, the
interpolated fields and one full step within 1e-5 of the oracle
(max-norm relative); the explicit terms, a difference quotient of fluxes,
within 1e-4; 25 steps within 1e-3 relative L2.  With the output layer
zeroed the interpolations are the multilinear ones and the step must equal
the linear-convection FVM step bit for bit.
"""
import functools

import numpy as np
import pytest
import torch

from oracle import fvm as ofvm
from oracle import grid_ops as go
from oracle import initial_conditions as oic
from oracle import ml as oml
from oracle.grid_ops import A, Grid as OGrid

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def _dev():
  return torch.device('cuda', 0)


def _maxrel(a, b):
  return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _rel_l2(a, b):
  return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _grid(n):
  import jax_cfd_amd.base as cfd
  return cfd.grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))


def _ogrid(n):
  return OGrid.from_domain((n, n), ((0, TWO_PI), (0, TWO_PI)))


def _forcing():
  import jax_cfd_amd.base as cfd
  return functools.partial(cfd.forcings.simple_turbulence_forcing, constant_magnitude=1.0,
                           constant_wavenumber=4, linear_coefficient=-0.1,
                           forcing_type='kolmogorov')


def _model(n, params, hidden, layers, equation='implicit_diffusion', dt=None, forced=True):
  from jax_cfd_amd import ml
  grid = _grid(n)
  if dt is None:
    dt = ofvm.stable_time_step(2.0, 0.5, 1e-3, _ogrid(n))
  tower = functools.partial(ml.towers.forward_tower_factory, num_hidden_channels=hidden,
                            num_hidden_layers=layers)
  conv = functools.partial(
      ml.advections.self_advection,
      advection_module=functools.partial(
          ml.advections.modular_self_advection,
          interpolation_module=functools.partial(
              ml.interpolations.FusedLearnedInterpolation, tags=('u', 'c'),
              tower_factory=tower, params=params)))
  specs = ml.physics_specifications.NavierStokesPhysicsSpecs(1.0, 1e-3,
                                                             _forcing() if forced else None)
  solver = (ml.equations.semi_implicit_navier_stokes if equation == 'semi_implicit'
            else ml.equations.implicit_diffusion_navier_stokes)
  return ml.equations.modular_navier_stokes_model(grid, dt, specs, equation_solver=solver,
                                                  convection_module=conv), dt


def _case(n, hidden=16, layers=2, seed=0, last_scale=1.0, bias_scale=0.05):
  shapes = oml.tower_shapes(120, num_hidden_channels=hidden, num_hidden_layers=layers)
  params = oml.init_tower_params(shapes, seed=seed, bias_scale=bias_scale, last_scale=last_scale)
  v0 = oic.filtered_velocity_field(seed, _ogrid(n), 2.0, 4)
  return params, v0


def _state(v0, n):
  import jax_cfd_amd.base as cfd
  grid = _grid(n)
  bc = cfd.boundaries.periodic_boundary_conditions(2)
  return tuple(cfd.grids.GridVariable(cfd.grids.GridArray(
      torch.as_tensor(u.data, device=_dev()), u.offset, grid), bc) for u in v0)


@pytest.mark.parametrize("n,hidden,layers", [(32, 16, 2), (64, 32, 2), (128, 64, 3), (64, 16, 0)])
def test_tower_logits_and_coefficients(n, hidden, layers):
  params, v0 = _case(n, hidden, layers)
  step, _ = _model(n, params, hidden, layers)
  plan = step.plan(_dev())
  u, v = (torch.as_tensor(x.data, device=_dev()) for x in v0)
  logits = plan.tower(u, v, coefficients=False)[0, :120].permute(1, 2, 0).cpu().numpy()
  want = oml.forward_tower(params, np.stack([x.data for x in v0], -1))
  err = _maxrel(logits, want)
  print(f"tower n={n} hidden={hidden} layers={layers + 1}: logits max rel err {err:.2e}")
  assert err <= 1e-5
  coef = plan.tower(u, v, coefficients=True)[0].permute(1, 2, 0).cpu().numpy()
  li = oml.FusedLearnedInterpolation(_ogrid(n), v0, params, tags=('u', 'c'))
  for k, (key, d) in enumerate(li.derivatives.items()):
    want_c = d.constraint(li.split_logits[key])
    e = _maxrel(coef[..., 16 * k:16 * (k + 1)], want_c)
    assert e <= 1e-5, (key, e)


def test_interpolations_match_the_oracle():
  from jax_cfd_amd import ml
  n = 64
  params, v0 = _case(n)
  state = _state(v0, n)
  interp = ml.interpolations.FusedLearnedInterpolation(
      _grid(n), 0.0, None, state, tags=('u', 'c'),
      tower_factory=functools.partial(ml.towers.forward_tower_factory, num_hidden_channels=16),
      params=params)
  li = oml.FusedLearnedInterpolation(_ogrid(n), v0, params, tags=('u', 'c'))
  for (in_off, target, tag) in li.derivatives:
    comp = 0 if in_off == (1.0, 0.5) else 1
    got = interp(state[comp], target, tag=tag).data.cpu().numpy()
    want = li(v0[comp], target, tag=tag).data
    e = _maxrel(got, want)
    print(f"interpolation {in_off}->{target} tag {tag}: max rel err {e:.2e}")
    assert e <= 1e-5
  with pytest.raises(KeyError):
    interp(state[0], (1.5, 0.5), tag=None)


@pytest.mark.parametrize("n,equation", [(64, 'semi_implicit'), (128, 'implicit_diffusion')])
def test_explicit_terms_and_one_step(n, equation):
  params, v0 = _case(n, seed=1)
  step, dt = _model(n, params, 16, 2, equation)
  forcing = go.Forcing(1.0, 4, -0.1)
  ostep = oml.LearnedInterpolationStep(_ogrid(n), dt, params, forcing=forcing, equation=equation)
  plan = step.plan(_dev())
  u, v = (torch.as_tensor(x.data, device=_dev()) for x in v0)
  if equation == 'semi_implicit':
    du, dv = plan.explicit_terms(u, v)
    for got, want in zip((du, dv), ostep.explicit_terms(v0)):
      e = _maxrel(got.cpu().numpy(), want.data)
      print(f"explicit terms n={n}: max rel err {e:.2e}")
      assert e <= 1e-4
  out = step(_state(v0, n))
  ref = ostep(v0)
  ref64 = oml.LearnedInterpolationStep(
      _ogrid(n), dt, [(w.astype(np.float64), b.astype(np.float64)) for w, b in params],
      forcing=forcing, equation=equation)(tuple(A(x.data.astype(np.float64), x.offset) for x in v0))
  for got, want, w64 in zip(out, ref, ref64):
    g = got.data.cpu().numpy()
    e = _maxrel(g, want.data)
    print(f"one {equation} step n={n}: max rel err {e:.2e} (vs fp64 {_maxrel(g, w64.data):.2e}; "
          f"fp32 oracle vs fp64 {_maxrel(want.data, w64.data):.2e})")
    assert e <= 1e-5


def test_rollout_rel_l2():
  """25 steps.  Untrained towers keep the stencils near the multilinear
  ones, whose centred fluxes under forward Euler grow without bound (the
  field blows up after ~40 steps here, in the oracle as on the GPU); a
  trained model is what makes long rollouts stable."""
  from jax_cfd_amd.base import funcutils
  n, steps = 64, 25
  params, v0 = _case(n, seed=2, last_scale=0.02)
  step, dt = _model(n, params, 16, 2)
  out = funcutils.repeated(step, steps)(_state(v0, n))
  ref = oml.LearnedInterpolationStep(_ogrid(n), dt, params, forcing=go.Forcing(1.0, 4, -0.1),
                                     equation='implicit_diffusion').rollout(v0, steps)
  for got, want in zip(out, ref):
    e = _rel_l2(got.data.cpu().numpy(), want.data)
    print(f"{steps} learned-interpolation steps n={n}: rel L2 {e:.2e}")
    assert e <= 1e-3


def test_zero_output_layer_is_linear_convection_bitwise():
  import jax_cfd_amd.base as cfd
  n = 64
  params, v0 = _case(n)
  params[-1] = (np.zeros_like(params[-1][0]), np.zeros_like(params[-1][1]))
  step, dt = _model(n, params, 16, 2, 'semi_implicit')
  grid = _grid(n)
  lin = cfd.equations.semi_implicit_navier_stokes(
      density=1.0, viscosity=1e-3, dt=dt, grid=grid,
      convect=lambda v: tuple(cfd.advection.advect_linear(u, v, dt) for u in v),
      forcing=_forcing()(grid))
  state = _state(v0, n)
  a = step.multistep(state, 5)
  b = lin.multistep(state, 5)
  for x, y in zip(a, b):
    assert torch.equal(x.data, y.data)
  ref = ofvm.FVMStep(_ogrid(n), dt=dt, viscosity=1e-3, forcing=go.Forcing(1.0, 4, -0.1),
                     convect='linear').rollout(v0, 5)
  for x, y in zip(a, ref):
    assert _maxrel(x.data.cpu().numpy(), y.data) <= 1e-5


def test_batch_equals_single_instances():
  n, batch = 64, 3
  params, _ = _case(n)
  step, _ = _model(n, params, 16, 2)
  fields = [oic.filtered_velocity_field(s, _ogrid(n), 2.0, 4) for s in range(batch)]
  u = torch.stack([torch.as_tensor(f[0].data) for f in fields]).to(_dev())
  v = torch.stack([torch.as_tensor(f[1].data) for f in fields]).to(_dev())
  ub, vb = step.plan(_dev(), batch).step(u, v, 3)
  single = step.plan(_dev())
  for s in range(batch):
    us, vs = single.step(u[s], v[s], 3)
    assert torch.equal(ub[s], us) and torch.equal(vb[s], vs)


def test_reference_architecture_256():
  """The tower of the reference's inference demo (7 convolutions: 2 -> 64,
  five 64 -> 64, 64 -> 120) on the forced 256^2 field: one step."""
  n = 256
  params, v0 = _case(n, hidden=64, layers=6, seed=4, last_scale=0.1)
  step, dt = _model(n, params, 64, 6)
  out = step(_state(v0, n))
  ref = oml.LearnedInterpolationStep(_ogrid(n), dt, params, forcing=go.Forcing(1.0, 4, -0.1),
                                     equation='implicit_diffusion')(v0)
  for got, want in zip(out, ref):
    e = _maxrel(got.data.cpu().numpy(), want.data)
    print(f"reference tower, 256^2 one step: max rel err {e:.2e}")
    assert e <= 1e-5
