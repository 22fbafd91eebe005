"""Host side of the learned-interpolation path, no GPU: the constraint
constants equal the (pinned) oracle's, the folded coefficient layer is the
constraint map applied to the logits layer, the interpolation keys follow
the split order, and the factory maps the reference's module composition
onto a learned-interpolation step (or raises for what is not built)."""
import functools

import numpy as np
import pytest

from oracle import ml as oml


def _ml():
  from jax_cfd_amd import ml
  return ml


def _grid(n=32):
  import jax_cfd_amd.base as cfd
  return cfd.grids.Grid((n, n), domain=((0, 2 * np.pi), (0, 2 * np.pi)))


def test_constraint_constants_match_the_oracle():
  ml = _ml()
  g = _grid()
  for off in g.cell_faces:
    for target in ml.interpolations.control_volume_offsets(off):
      got = ml.layers.SpatialDerivativeFromLogits((4, 4), off, target, (0, 0), g.step)
      want = oml.SpatialDerivativeFromLogits((4, 4), off, target, (0, 0), g.step)
      assert got.roll == want.roll == (0, 0)
      np.testing.assert_array_equal(got.constraint.bias, want.constraint.bias)
      np.testing.assert_array_equal(got.constraint.nullspace, want.constraint.nullspace)
  for stencil, order, acc in (([-2, -1, 0, 1, 2], 2, None), ([0, 1, 2, 3, 4], 4, None),
                              ([-2, -1, 0, 1, 2], 2, 1)):
    s = [np.array(stencil, float)]
    FD = ml.layers_util.Method.FINITE_DIFFERENCE
    np.testing.assert_allclose(
        ml.layers_util.polynomial_accuracy_coefficients(s, FD, [order], acc),
        oml.polynomial_accuracy_coefficients(s, oml.Method.FINITE_DIFFERENCE, [order], acc),
        atol=1e-12)
  assert ml.layers_util.patch_offsets((4, 4)) == oml.patch_shifts((4, 4))


def test_folded_layer_is_the_constraint_map():
  """conv with the folded layer == bias + conv(logits layer) . nullspace,
  per interpolation, in fp64 (linearity of the convolution)."""
  from jax_cfd_amd.ml import _li
  ml = _ml()
  g = _grid(8)
  rng = np.random.default_rng(0)
  keys, cons = [], []
  for off in g.cell_faces:
    for target in ml.interpolations.control_volume_offsets(off):
      for tag in ('u', 'c'):
        d = ml.layers.SpatialDerivativeFromLogits((4, 4), off, target, (0, 0), g.step)
        cons.append((d.constraint.nullspace, d.constraint.bias))
  w = rng.standard_normal((3, 3, 5, 120))
  b = rng.standard_normal(120)
  wc, bc = _li.fold_output_layer(w, b, cons)
  assert wc.shape == (3, 3, 5, 128) and bc.shape == (128,)
  x = rng.standard_normal((8, 8, 5))
  logits = oml.periodic_conv2d(x, w, b)
  coef = oml.periodic_conv2d(x, wc.astype(np.float64), bc.astype(np.float64))
  for k, (null, bias) in enumerate(cons):
    want = bias + logits[..., 15 * k:15 * (k + 1)] @ null
    np.testing.assert_allclose(coef[..., 16 * k:16 * (k + 1)], want, rtol=1e-5, atol=1e-5)
  with pytest.raises(ValueError):
    _li.fold_output_layer(w[..., :60], b[:60], cons)


def test_interp_keys_follow_the_split_order():
  from jax_cfd_amd.ml import _li
  # tags ('u', 'c'): interpolation (m, t, g) = (2 m + t) 2 + g
  assert _li.interp_keys(('u', 'c')) == (16, 0, 48, 64, 80, 32, 112, 96)
  assert _li.interp_keys(('c', 'u')) == (0, 16, 32, 80, 64, 48, 96, 112)
  with pytest.raises(NotImplementedError):
    _li.interp_keys((None,))


def _conv(tags=('u', 'c'), **kw):
  ml = _ml()
  return functools.partial(
      ml.advections.self_advection,
      advection_module=functools.partial(
          ml.advections.modular_self_advection,
          interpolation_module=functools.partial(ml.interpolations.FusedLearnedInterpolation,
                                                 tags=tags, **kw)))


def test_factory_builds_the_learned_step_and_rejects_the_rest():
  ml = _ml()
  import jax_cfd_amd.base as cfd
  g = _grid()
  forcing = functools.partial(cfd.forcings.simple_turbulence_forcing, constant_magnitude=1.0,
                              constant_wavenumber=4, linear_coefficient=-0.1,
                              forcing_type='kolmogorov')
  specs = ml.physics_specifications.NavierStokesPhysicsSpecs(1.0, 1e-3, forcing)
  step = ml.equations.modular_navier_stokes_model(g, 0.01, specs, convection_module=_conv())
  assert isinstance(step, ml.equations.LiStepFn)
  assert step.spec.diffusion == 1 and step.spec.linear_coefficient == -0.1
  assert [w.shape for w, _ in step.params] == [(3, 3, 2, 16), (3, 3, 16, 16), (3, 3, 16, 120)]
  step = ml.equations.modular_navier_stokes_model(
      g, 0.01, specs, equation_solver=ml.equations.semi_implicit_navier_stokes,
      convection_module=_conv())
  assert step.spec.diffusion == 0
  with pytest.raises(NotImplementedError):
    ml.equations.modular_navier_stokes_model(g, 0.01, specs, convection_module=_conv(tags=(None,)))
  with pytest.raises(NotImplementedError):
    ml.equations.modular_navier_stokes_model(
        g, 0.01, specs, convection_module=_conv(
            tower_factory=functools.partial(ml.towers.forward_tower_factory, kernel_size=5)))
  with pytest.raises(NotImplementedError):
    ml.equations.modular_navier_stokes_model(
        g, 0.01, specs, convection_module=_conv(
            tower_factory=functools.partial(ml.towers.forward_tower_factory,
                                            num_hidden_channels=24)))
  # the fixed-interpolation modules map onto the base schemes
  fixed = functools.partial(ml.advections.self_advection,
                            advection_module=functools.partial(
                                ml.advections.modular_advection,
                                c_interpolation_module=ml.interpolations.linear))
  step = ml.equations.modular_navier_stokes_model(
      g, 0.01, specs, equation_solver=ml.equations.semi_implicit_navier_stokes,
      convection_module=fixed)
  assert step.spec.convect == 2  # advect_linear


def test_tower_params_shape_check_and_init():
  ml = _ml()
  t = ml.towers.forward_tower_factory(120, 2, num_hidden_channels=64, num_hidden_layers=6)
  assert t.shapes[0] == (3, 3, 2, 64) and t.shapes[-1] == (3, 3, 64, 120) and len(t.shapes) == 7
  p = t.init_params(3)
  w0 = p[0][0]
  assert np.abs(w0).max() <= 2 * np.sqrt(1 / 18) / 0.87962566103423978 + 1e-6
  with pytest.raises(ValueError):
    t.check_params(p[:-1])
  ref = oml.init_tower_params(t.shapes, seed=3)
  for (a, b), (c, d) in zip(p, ref):
    np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(b, d)
