"""Pins the learned-interpolation oracle (oracle/ml.py) on the CPU.

Known answers from the reference's own tests (ml/layers_util_test.py,
ml/layers_test.py), a brute-force periodic cross-correlation, and what the
mathematics fixes: zero logits reduce every interpolation to multilinear
interpolation and the learned advection to advect_linear exactly; the
constraint space holds the classical stencils.
"""
import itertools

import numpy as np
import pytest

from oracle import fvm as ofvm
from oracle import grid_ops as go
from oracle import initial_conditions as oic
from oracle import ml
from oracle.grid_ops import A, Grid

FD = ml.Method.FINITE_DIFFERENCE
FV = ml.Method.FINITE_VOLUME


def test_exponents_up_to_degree():
  assert list(ml.exponents_up_to_degree(2, 2)) == [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 0)]


@pytest.mark.parametrize("order,method,ea,eb", [
    (1, FD, [[1, 1]], [1]), (1, FV, [[1, 1]], [1]),
    (2, FD, [[-1 / 2, 1 / 2], [1, 1]], [0, 1]), (2, FV, [[-1 / 2, 1 / 2], [1, 1]], [0, 1])])
def test_constraints_1d(order, method, ea, eb):
  a, b = ml.polynomial_accuracy_constraints([np.array([-0.5, 0.5])], method, [0], order, 1.0)
  np.testing.assert_allclose(a, ea)
  np.testing.assert_allclose(b, eb)


def test_constraints_2d_and_3d():
  a, b = ml.polynomial_accuracy_constraints([np.array([-0.5, 0.5])] * 2, FD, [0, 0], 2)
  assert a.shape == (3, 4) and b.shape == (3,)
  np.testing.assert_allclose(a.dot([1 / 4] * 4), b)
  np.testing.assert_allclose(a.dot([4 / 10, 1 / 10, 1 / 10, 4 / 10]), b)
  a, b = ml.polynomial_accuracy_constraints([np.array([-1, 0, 1])] * 2, FD, [1, 0], 1)
  assert a.shape == (3, 9)
  np.testing.assert_allclose(a.dot(np.array([-1, 0, -1, 0, 0, 0, 1, 0, 1]) / 4), b)
  assert np.linalg.norm(a.dot(np.array([-1, 0, -1, 1, 0, -1, 1, 0, 1]) / 4) - b) > 0.1
  a, b = ml.polynomial_accuracy_constraints([np.array([-1, 0, 1])] * 2, FD, [1, 1], 1)
  assert a.shape == (6, 9)
  np.testing.assert_allclose(a.dot(np.array([1, 0, -1, 0, 0, 0, -1, 0, 1]) / 4), b)
  stencils = [np.array([-1, 0, 1])] * 3
  a, b = ml.polynomial_accuracy_constraints(stencils, FD, [0, 0, 0], 2)
  assert a.shape == (4, 27)
  pts = list(itertools.product(*stencils))
  sa = np.zeros(27)
  sa[pts.index((0, 0, 0))] = 1.
  sb = np.zeros(27)
  sb[pts.index((0, -1, 0))], sb[pts.index((0, 0, 0))], sb[pts.index((0, 1, 0))] = .25, .5, .25
  np.testing.assert_allclose(a.dot(sa), b)
  np.testing.assert_allclose(a.dot(sb), b)
  a, b = ml.polynomial_accuracy_constraints(stencils, FD, [0, 0, 1], 2)
  assert a.shape == (10, 27)


@pytest.mark.parametrize("stencil,order,expected,acc", [
    ([-1, 0, 1], 1, [-1 / 2, 0, 1 / 2], None), ([-1, 0, 1], 2, [1, -2, 1], None),
    ([-2, -1, 0, 1, 2], 2, [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12], None),
    ([-2, -1, 0, 1, 2], 2, [0, 1, -2, 1, 0], 1), ([0, 1], 1, [-1, 1], None),
    ([0, 2], 1, [-0.5, 0.5], None), ([0, 0.5], 1, [-2, 2], None),
    ([0, 1, 2, 3, 4], 4, [1, -4, 6, -4, 1], None)])
def test_finite_difference_coefficients_1d(stencil, order, expected, acc):
  out = ml.polynomial_accuracy_coefficients([np.array(stencil)], FD, [order], acc)
  np.testing.assert_allclose(out, expected, atol=1e-12)


@pytest.mark.parametrize("stencils,orders,expected", [
    ([[-0.5, 0.5], [-0.5, 0.5]], [0, 0], [[0.25, 0.25], [0.25, 0.25]]),
    ([[-0.5, 0.5], [-0.5, 0.5]], [0, 1], [[-0.5, 0.5], [-0.5, 0.5]]),
    ([[-0.5, 0.5], [-0.5, 0.5]], [1, 1], [[1, -1], [-1, 1]]),
    ([[-1, 0, 1], [-0.5, 0.5]], [1, 0], [[-0.25, -0.25], [0, 0], [0.25, 0.25]])])
def test_finite_difference_coefficients_2d(stencils, orders, expected):
  args = ([np.array(s) for s in stencils], FD, orders)
  np.testing.assert_allclose(ml.polynomial_accuracy_coefficients(*args), expected, atol=1e-12)
  np.testing.assert_allclose(ml.polynomial_accuracy_coefficients(*args, accuracy_order=1),
                             expected, atol=1e-12)


@pytest.mark.parametrize("stencil,order,expected,acc", [
    ([-0.5, 0.5], 0, [1 / 2, 1 / 2], None), ([-1.5, -0.5, 0.5, 1.5], 0, [0, 1 / 2, 1 / 2, 0], 1),
    ([-1, 1], 0, [1 / 2, 1 / 2], None), ([-1.5, -0.5], 0, [-1 / 2, 3 / 2], None),
    ([-0.5, 0.5, 1.5], 0, [1 / 3, 5 / 6, -1 / 6], None), ([-0.5, 0.5], 1, [-1, 1], None),
    ([-1, 1], 1, [-1 / 2, 1 / 2], None), ([-1, 0, 1], 1, [-1 / 2, 0, 1 / 2], None),
    ([0.5, 1.5, 2.5], 1, [-2, 3, -1], None),
    ([-1.5, -0.5, 0.5, 1.5], 1, [1 / 12, -5 / 4, 5 / 4, -1 / 12], None),
    ([-.75, -0.25, 0.25, 0.75], 1, [1 / 6, -5 / 2, 5 / 2, -1 / 6], None)])
def test_finite_volume_coefficients_1d(stencil, order, expected, acc):
  step = stencil[1] - stencil[0]
  out = ml.polynomial_accuracy_coefficients([np.array(stencil)], FV, [order], acc, grid_step=step)
  np.testing.assert_allclose(out, expected, atol=1e-12)


def test_extract_patches_known_answers():
  x = np.arange(10, dtype=np.float32).reshape([10, 1])
  np.testing.assert_array_equal(ml.extract_patches(x, (4,))[0], [9, 0, 1, 2])
  np.testing.assert_array_equal(ml.extract_patches(x, (4,))[5], [4, 5, 6, 7])
  np.testing.assert_array_equal(ml.extract_patches(x, (5,))[0], [8, 9, 0, 1, 2])
  x = np.arange(16, dtype=np.float32).reshape([4, 4, 1])
  np.testing.assert_array_equal(ml.extract_patches(x, (2, 2))[0, 0], [0, 1, 4, 5])
  np.testing.assert_array_equal(ml.extract_patches(x, (2, 2))[1, 1], [5, 6, 9, 10])
  np.testing.assert_array_equal(ml.extract_patches(x, (3, 3))[0, 0],
                                [15, 12, 13, 3, 0, 1, 7, 4, 5])
  np.testing.assert_array_equal(ml.extract_patches(x, (3, 3))[1, 1], [0, 1, 2, 4, 5, 6, 8, 9, 10])
  x = np.arange(125, dtype=np.float32).reshape([5, 5, 5, 1])
  np.testing.assert_array_equal(ml.extract_patches(x, (2, 2, 2))[3, 3, 3],
                                [93, 94, 98, 99, 118, 119, 123, 124])


@pytest.mark.parametrize("sizes,orders", [((4, 4), (0, 0)), ((3,), (1,)), ((4, 3), (0, 1)),
                                          ((3, 3, 3), (0, 0, 0))])
def test_polynomial_constraint_outputs_satisfy_constraints(sizes, orders):
  """layers_test.py:280-337: random inputs give coefficients in the
  constraint set; the nullspace is orthogonal to the constraint rows."""
  h = 0.1
  stencils = [np.array([(s - (n - 1) / 2) * h for s in range(n)]) for n in sizes]
  mod = ml.PolynomialConstraint(stencils, orders, FV, (h,) * len(sizes))
  rng = np.random.default_rng(0)
  out = mod(rng.uniform(size=(5, 6, mod.subspace_size)))
  a, b = ml.polynomial_accuracy_constraints(stencils, FV, orders, 1, h)
  assert out.shape == (5, 6, int(np.prod(sizes)))
  np.testing.assert_allclose(np.tensordot(out, a, axes=[-1, -1]) - b, 0, atol=1e-8)
  np.testing.assert_allclose(a @ mod.nullspace.T, 0, atol=1e-10)


def test_constraint_space_holds_the_classical_face_stencil():
  """The 4-point finite-volume value at a face, (-1, 7, 7, -1)/12, lies in the
  order-1 constraint space of a (4,) stencil at shift 1/2: a least-squares
  combination of the nullspace reaches it exactly."""
  h = 0.3
  d = ml.SpatialDerivativeFromLogits((4,), (0.5,), (1.0,), (0,), (h,))
  target = np.array([-1, 7, 7, -1]) / 12
  logits, *_ = np.linalg.lstsq(d.constraint.nullspace.T, target - d.constraint.bias, rcond=None)
  np.testing.assert_allclose(d.constraint(logits), target, atol=1e-12)
  np.testing.assert_allclose(
      ml.polynomial_accuracy_coefficients(ml.get_stencils((4,), (0.5,), (h,)), FV, [0],
                                          grid_step=h).ravel(), target, atol=1e-12)


def _grid(n, lx=2 * np.pi):
  return Grid.from_domain((n, n), ((0, lx), (0, lx)))


def test_zero_logits_are_multilinear_interpolation():
  """Every (input, target) pair of the staggered velocity: the bias of the
  order-1 constraint is the multilinear interpolation, bitwise."""
  g = _grid(16)
  rng = np.random.default_rng(3)
  for off in g.cell_faces:
    c = A(rng.standard_normal(g.shape).astype(np.float32), off)
    for target in go.control_volume_offsets(c):
      d = ml.SpatialDerivativeFromLogits((4, 4), off, target, (0, 0), g.step)
      assert d.subspace_size == 15
      out = d(c.data[..., None], np.zeros(g.shape + (15,), np.float32))[..., 0]
      np.testing.assert_array_equal(out.astype(np.float32), go.linear(c, target).data)


def test_periodic_conv_brute_force_roll_and_equivariance():
  rng = np.random.default_rng(1)
  n0, n1, cin, cout = 6, 5, 3, 4
  x = rng.standard_normal((n0, n1, cin))
  w = rng.standard_normal((3, 3, cin, cout))
  b = rng.standard_normal(cout)
  out = ml.periodic_conv2d(x, w, b)
  want = np.zeros((n0, n1, cout))
  for i in range(n0):
    for j in range(n1):
      for a in range(3):
        for c in range(3):
          want[i, j] += x[(i + a - 1) % n0, (j + c - 1) % n1] @ w[a, c]
  np.testing.assert_allclose(out, want + b, rtol=1e-12, atol=1e-12)
  for a in range(3):
    for c in range(3):
      delta = np.zeros((3, 3, cin, cin))
      delta[a, c] = np.eye(cin)
      np.testing.assert_array_equal(ml.periodic_conv2d(x, delta, np.zeros(cin)),
                                    np.roll(x, (1 - a, 1 - c), axis=(0, 1)))
  np.testing.assert_allclose(ml.periodic_conv2d(np.roll(x, (2, -1), (0, 1)), w, b),
                             np.roll(out, (2, -1), (0, 1)), rtol=1e-12, atol=1e-12)


def test_tower_channels_and_relu():
  shapes = ml.tower_shapes(120, num_hidden_channels=8, num_hidden_layers=2)
  assert shapes == [(3, 3, 2, 8), (3, 3, 8, 8), (3, 3, 8, 120)]
  params = ml.init_tower_params(shapes, seed=0, bias_scale=0.1)
  x = np.random.default_rng(0).standard_normal((8, 8, 2)).astype(np.float32)
  h = np.maximum(ml.periodic_conv2d(x, *params[0]), 0)
  h = np.maximum(ml.periodic_conv2d(h, *params[1]), 0)
  np.testing.assert_array_equal(ml.forward_tower(params, x), ml.periodic_conv2d(h, *params[2]))
  w0 = params[0][0].astype(np.float64)
  std = np.sqrt(1 / 18) / 0.87962566103423978
  assert np.abs(w0).max() <= 2 * std + 1e-7


def test_zero_tower_output_gives_advect_linear_bitwise():
  """Output convolution zero: every interpolation is its bias, so the
  learned self-advection is convect_linear exactly; the whole step equals
  the oracle's linear-convection step."""
  n = 32
  g = _grid(n)
  v = oic.filtered_velocity_field(0, g, 2.0, 4)
  params = ml.init_tower_params(ml.tower_shapes(120, num_hidden_channels=8), seed=2)
  params[-1] = (np.zeros_like(params[-1][0]), np.zeros_like(params[-1][1]))
  conv = ml.learned_self_convection(g, v, params)
  for got, u in zip(conv, v):
    np.testing.assert_array_equal(got, go.advect_linear(u, v, g).data)
  dt = ofvm.stable_time_step(2.0, 0.5, 1e-3, g)
  forcing = go.Forcing(1.0, 4, -0.1)
  li = ml.LearnedInterpolationStep(g, dt, params, forcing=forcing)(v)
  lin = ofvm.FVMStep(g, dt=dt, viscosity=1e-3, forcing=forcing, convect='linear')(v)
  for a, b in zip(li, lin):
    np.testing.assert_array_equal(a.data, b.data)


def test_learned_interpolation_channel_order():
  """Logits split in the reference's insertion order: for u in v, for each
  control-volume face of u, for each tag."""
  g = _grid(8)
  v = tuple(A(np.zeros(g.shape, np.float32), o) for o in g.cell_faces)
  params = ml.init_tower_params(ml.tower_shapes(120, num_hidden_channels=4), seed=0)
  li = ml.FusedLearnedInterpolation(g, v, params, tags=('u', 'c'))
  keys = list(li.derivatives)
  assert keys == [((1.0, 0.5), (1.5, 0.5), 'u'), ((1.0, 0.5), (1.5, 0.5), 'c'),
                  ((1.0, 0.5), (1.0, 1.0), 'u'), ((1.0, 0.5), (1.0, 1.0), 'c'),
                  ((0.5, 1.0), (1.0, 1.0), 'u'), ((0.5, 1.0), (1.0, 1.0), 'c'),
                  ((0.5, 1.0), (0.5, 1.5), 'u'), ((0.5, 1.0), (0.5, 1.5), 'c')]
  with pytest.raises(KeyError):
    li(v[0], (1.5, 0.5), tag=None)
