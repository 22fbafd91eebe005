"""Pins the CPU oracle of the FVM path (no GPU needed).

Each test cites the reference test or closed form it restates.  A plausible
mistake in the oracle (a wrong shift sign, a transposed axis, a dropped term)
fails at least one of them.
"""
import numpy as np
import pytest
from scipy import interpolate

from oracle import fvm
from oracle import grid_ops as go
from oracle.grid_ops import A, Grid


# ------------------------------------------------------------- shift / FD KATs
@pytest.mark.parametrize("k,expected", [(-2, [13, 14, 11, 12]),
                                        (-1, [14, 11, 12, 13]),
                                        (0, [11, 12, 13, 14]),
                                        (1, [12, 13, 14, 11]),
                                        (2, [13, 14, 11, 12])])
def test_periodic_shift_kat(k, expected):
  """boundaries_test.py:168-209 known answers."""
  out = go.shift(A(np.array([11, 12, 13, 14]), (0,)), k, 0)
  np.testing.assert_array_equal(out.data, expected)
  assert out.offset == (k,)


def test_periodic_shift_matches_index_roll():
  """boundaries_test.py:150-166: shift == data[(i + o) % n] per axis."""
  shape = (3, 4, 5)
  data = np.arange(np.prod(shape)).reshape(shape)
  offs = (1, -2, 3)
  out = A(data, (0, 0.5, 1))
  for ax, o in enumerate(offs):
    out = go.shift(out, o, ax)
  idx = np.meshgrid(*[(np.arange(s) + o) % s for s, o in zip(shape, offs)],
                    indexing='ij')
  np.testing.assert_array_equal(out.data, data[tuple(idx)])
  assert out.offset == (1, -1.5, 4)


@pytest.mark.parametrize("method,neg,pos,expected_offset", [
    (go.backward_difference, -1, 0, -0.5),
    (go.forward_difference, 0, 1, 0.5)])
def test_finite_difference_indexing(method, neg, pos, expected_offset):
  """finite_differences_test.py:44-95 with explicit indices."""
  shape, step = (5, 6, 7), (.1, .3, 1.)
  u = A(np.arange(np.prod(shape), dtype=np.float64).reshape(shape), (0, 0, 0))
  mesh = np.meshgrid(*[np.arange(s) for s in shape], indexing='ij')
  for ax in range(3):
    out = method(u, ax, step[ax])
    ip = list(mesh)
    im = list(mesh)
    ip[ax] = np.roll(mesh[ax], -pos, axis=ax)
    im[ax] = np.roll(mesh[ax], -neg, axis=ax)
    expected = (u.data[tuple(ip)] - u.data[tuple(im)]) / step[ax]
    np.testing.assert_allclose(out.data, expected, rtol=1e-14)
    assert out.offset[ax] == expected_offset


@pytest.mark.parametrize("shape,f,g", [
    ((20, 20), lambda x, y: np.ones_like(x), lambda x, y: 0 * x),
    ((21, 21), lambda x, y: x * (x - 1.) + y * (y - 1.), lambda x, y: 4 + 0 * x),
    ((13, 13, 13), lambda x, y, z: x * (x - 1.) + y * (y - 1.) + z * (z - 1.),
     lambda x, y, z: 6 + 0 * x)])
def test_laplacian_closed_forms(shape, f, g):
  """finite_differences_test.py:137-162 (interior, periodic wrap trimmed)."""
  grid = Grid.from_step(shape, tuple(1. / s for s in shape))
  offset = (0,) * len(shape)
  mesh = grid.mesh(offset, np.float64)
  lap = go.laplacian(A(f(*mesh), offset), grid).data
  inner = tuple(slice(1, -1) for _ in shape)
  np.testing.assert_allclose(lap[inner], g(*mesh)[inner], atol=1e-3)


def test_diffusion_of_constant_is_zero():
  """diffusion_test.py:44-56."""
  grid = Grid.from_step((16, 16), 0.1)
  c = A(np.full((16, 16), 3.0, np.float32), (0.5, 0.5))
  assert np.abs(go.diffuse(c, 0.7, grid).data).max() == 0


def test_linear_interpolation_vs_scipy():
  """interpolation_test.py:55-98: linear == scipy interpn on interior points."""
  shape = (10, 12)
  grid = Grid.from_step(shape, (0.3, 0.2))
  rng = np.random.default_rng(0)
  data = rng.standard_normal(shape)
  c = A(data, (1.0, 0.5))
  target = (0.5, 1.0)
  out = go.linear(c, target)
  pts_src = [(np.arange(n) + o) * h for n, o, h in zip(shape, c.offset, grid.step)]
  mesh_t = grid.mesh(target, np.float64)
  xi = np.stack([m[1:-1, 1:-1].ravel() for m in mesh_t], -1)
  ref = interpolate.interpn(pts_src, data, xi)
  np.testing.assert_allclose(out.data[1:-1, 1:-1].ravel(), ref, rtol=1e-12)


def test_upwind_and_lax_wendroff_kat():
  """interpolation_test.py:123-174, 199-260 in 1-D: exact face values."""
  grid = Grid.from_step((6,), 1.0)
  c = A(np.array([1., 4., 2., 8., 5., 7.]), (0.5,))
  for sign in (1., -1.):
    u = A(np.full(6, sign * 0.5), (1.0,))
    up = go.upwind(c, (1.0,), (u,))
    expected = c.data if sign > 0 else np.roll(c.data, -1)
    np.testing.assert_array_equal(up.data, expected)
    dt = 0.4
    lw = go.lax_wendroff(c, (1.0,), (u,), dt, grid)
    cr = np.roll(c.data, -1)
    cour = dt * u.data
    if sign > 0:
      exp = c.data + 0.5 * (1 - cour) * (cr - c.data)
    else:
      exp = cr - 0.5 * (1 + cour) * (cr - c.data)
    np.testing.assert_allclose(lw.data, exp, rtol=1e-15)


# ------------------------------------------------------------- advection
def _gaussian(grid, offset, dtype=np.float64):
  mesh = grid.mesh(offset, dtype)
  size = [n * h for n, h in zip(grid.shape, grid.step)]
  return np.exp(-sum(((x / s - 0.5) ** 2) for x, s in zip(mesh, size)) * 50.)


def _unit_velocity(grid, sign=1.0, dtype=np.float64):
  return tuple(A(np.full(grid.shape, sign, dtype), off)
               for off in grid.cell_faces)


@pytest.mark.parametrize("shape,sign", [((101,), 1.), ((101,), -1.),
                                        ((40, 40), 1.), ((40, 40), -1.),
                                        ((16, 16, 16), 1.)])
def test_van_leer_same_as_van_leer_using_limiters(shape, sign):
  """advection_test.py:444-472: two formulations agree for a cell-centred
  scalar over 100 Euler steps (atol 1e-5)."""
  grid = Grid.from_step(shape, tuple(1. / s for s in shape))
  v = _unit_velocity(grid, sign)
  c0 = A(_gaussian(grid, grid.cell_center), grid.cell_center)
  dt = min(grid.step) / 100.
  c1, c2 = c0, c0
  for _ in range(100):
    c1 = A(c1.data + dt * go.advect_van_leer(c1, v, dt, grid).data, c1.offset)
    c2 = A(c2.data + dt * go.advect_van_leer_using_limiters(c2, v, dt, grid).data,
           c2.offset)
  np.testing.assert_allclose(c1.data, c2.data, atol=1e-5)
  # and the scheme moved the profile (guards against a zero advection term)
  assert np.abs(c1.data - c0.data).max() > 1e-3


def test_van_leer_tvd_and_mass():
  """advection_test.py:376-405 (TVD of a square wave) and periodic mass
  conservation (advection_test.py:407-442 pattern)."""
  shape = (101,)
  grid = Grid.from_step(shape, 1. / 101)
  x = grid.mesh(grid.cell_center, np.float64)[0]
  c = A(np.where((x > 0.3) & (x < 0.6), 1.0, 0.0), grid.cell_center)
  v = _unit_velocity(grid)
  dt = grid.step[0] / 100.
  tv0 = np.abs(np.diff(c.data, append=c.data[:1])).sum() + 1e-6
  m0 = c.data.sum()
  for _ in range(300):
    c = A(c.data + dt * go.advect_van_leer(c, v, dt, grid).data, c.offset)
    assert np.abs(np.diff(c.data, append=c.data[:1])).sum() <= tv0
  assert abs(c.data.sum() - m0) < 1e-9


def test_van_leer_unit_velocity_translates_gaussian():
  """advection_test.py:84-168: with Courant number 1 one Euler step of the
  upwind-limited scheme is an exact shift by one cell."""
  shape = (64,)
  grid = Grid.from_step(shape, 1. / 64)
  c = A(_gaussian(grid, grid.cell_center), grid.cell_center)
  v = _unit_velocity(grid)
  dt = grid.step[0]  # Courant number 1: correction prefactor (1 - C) = 0
  out = c.data + dt * go.advect_van_leer(c, v, dt, grid).data
  np.testing.assert_allclose(out, np.roll(c.data, 1), atol=1e-12)


def _phi(num, den):
  if abs(den) > 0:
    inv = den / (abs(den) + abs(num))
  else:
    inv = den / 1.0
  return num * (np.sign(den) + np.sign(num)) * inv


def _corr(U, num_pos, num_neg, den, dt, h):
  num = num_pos if U > 0 else num_neg
  return 0.5 * (1 - (dt / h) * abs(U)) * abs(U) * _phi(num, den)


def _brute_force_velocity_advection(u, v, dt, dx, dy):
  """Per-cell evaluation of the staggered-velocity Van Leer stencil written
  out by hand from advection.py:210-342 (independent of the array code)."""
  nx, ny = u.shape
  U = lambda i, j: u[i % nx, j % ny]
  V = lambda i, j: v[i % nx, j % ny]

  def fx_u(a, j):
    Uf = 0.5 * U(a - 1, j) + 0.5 * U(a, j)
    up = Uf * U(a - 1, j) if Uf > 0 else Uf * U(a, j)
    return up + _corr(Uf, U(a - 1, j) - U(a - 2, j), U(a + 1, j) - U(a, j),
                      U(a, j) - U(a - 1, j), dt, dx)

  def fy_u(a, b):
    Vf = 0.5 * V(a - 1, b) + 0.5 * V(a, b)
    up = Vf * U(a - 1, b) if Vf > 0 else Vf * U(a - 1, b + 1)
    # limiter stencil on column a (the stale-loop-variable behaviour)
    return up + _corr(Vf, U(a, b) - U(a, b - 1), U(a, b + 2) - U(a, b + 1),
                      U(a, b + 1) - U(a, b), dt, dy)

  def fx_v(a, b):
    Uf = 0.5 * U(a, b - 1) + 0.5 * U(a, b)
    up = Uf * V(a, b - 1) if Uf > 0 else Uf * V(a + 1, b - 1)
    # limiter stencil on row b
    return up + _corr(Uf, V(a, b) - V(a - 1, b), V(a + 2, b) - V(a + 1, b),
                      V(a + 1, b) - V(a, b), dt, dx)

  def fy_v(i, b):
    Vf = 0.5 * V(i, b - 1) + 0.5 * V(i, b)
    up = Vf * V(i, b - 1) if Vf > 0 else Vf * V(i, b)
    return up + _corr(Vf, V(i, b - 1) - V(i, b - 2), V(i, b + 1) - V(i, b),
                      V(i, b) - V(i, b - 1), dt, dy)

  adv_u = np.zeros_like(u)
  adv_v = np.zeros_like(v)
  for i in range(nx):
    for j in range(ny):
      adv_u[i, j] = -((fx_u(i + 1, j) - fx_u(i, j)) / dx
                      + (fy_u(i + 1, j) - fy_u(i + 1, j - 1)) / dy)
      adv_v[i, j] = -((fx_v(i, j + 1) - fx_v(i - 1, j + 1)) / dx
                      + (fy_v(i, j + 1) - fy_v(i, j)) / dy)
  return adv_u, adv_v


def test_van_leer_velocity_components_brute_force():
  """The oracle's staggered-velocity advection equals the per-cell stencil
  derived by hand from the reference source (rectangular, dx != dy, both
  velocity signs, some zero gradients to hit the den == 0 branch)."""
  rng = np.random.default_rng(3)
  shape = (7, 9)
  grid = Grid.from_step(shape, (0.3, 0.2))
  u = rng.standard_normal(shape)
  v = rng.standard_normal(shape)
  u[2, :] = u[3, :]  # flat stretches -> den == 0
  v[:, 4] = v[:, 5]
  vel = fvm.velocity(u, v, grid)
  dt = 0.05
  ou = go.advect_van_leer(vel[0], vel, dt, grid)
  ov = go.advect_van_leer(vel[1], vel, dt, grid)
  bu, bv = _brute_force_velocity_advection(u, v, dt, *grid.step)
  np.testing.assert_allclose(ou.data, bu, rtol=1e-12, atol=1e-12)
  np.testing.assert_allclose(ov.data, bv, rtol=1e-12, atol=1e-12)
  assert ou.offset == (1.0, 0.5) and ov.offset == (0.5, 1.0)


# ----------------------------------------------------------- pressure solve
def test_poisson_periodic_kat():
  """pressure_test.py:42-57, 209-213: laplacian(pinv(b)) == b - mean(b)."""
  rs = np.random.RandomState(0)
  b = rs.randn(4, 4).astype(np.float32)
  grid = Grid.from_domain((4, 4), ((0, 4), (0, 4)))
  diag = fvm.poisson_pinv_diagonal(grid, np.float32)
  x = np.fft.irfftn(diag * np.fft.rfftn(b), s=b.shape, axes=(0, 1))
  lap = go.laplacian(A(x.astype(np.float32), (1.0, 0.5)), grid).data
  np.testing.assert_allclose(lap, b - b.mean(), atol=1e-5)


def test_pinv_matches_dense_lstsq():
  """fast_diagonalization_test.py:58-130: the circulant pseudo-inverse equals
  the minimum-norm least-squares solve of the dense 2-D periodic Laplacian."""
  shape = (6, 8)
  grid = Grid.from_step(shape, (0.5, 0.25))
  n = np.prod(shape)
  dense = np.zeros((n, n))
  for k in range(n):
    e = np.zeros(n)
    e[k] = 1
    dense[:, k] = go.laplacian(A(e.reshape(shape), (0.5, 0.5)), grid).data.ravel()
  rng = np.random.default_rng(1)
  b = rng.standard_normal(shape)
  ref = np.linalg.lstsq(dense, b.ravel(), rcond=None)[0].reshape(shape)
  diag = fvm.poisson_pinv_diagonal(grid, np.float64)
  x = np.fft.irfftn(diag * np.fft.rfftn(b), s=shape, axes=(0, 1))
  np.testing.assert_allclose(x, ref, atol=1e-10)


@pytest.mark.parametrize("solve", ["fast_diag", "cg"])
def test_projection_divergence_free(solve):
  """pressure_test.py:59-110: 1 + 0.3 N(0,1) velocity, projected, div ~ 0."""
  shape = (32, 32)
  grid = Grid.from_step(shape, (1., 1.))
  rng = np.random.default_rng(222)
  v = tuple(A((1. + .3 * rng.standard_normal(shape)).astype(np.float64), o)
            for o in grid.cell_faces)
  s = fvm.solve_fast_diag if solve == "fast_diag" else (
      lambda vv, g: fvm.solve_cg(vv, g, 1e-10, 1e-10))
  out = fvm.projection(v, grid, s)
  div = go.divergence(out, grid).data
  np.testing.assert_allclose(div, 0., atol=1e-8)
  # projection removes only a gradient: curl unchanged
  np.testing.assert_allclose(go.curl_2d(out, grid).data,
                             go.curl_2d(v, grid).data, atol=1e-10)


def test_cg_matches_fast_diag():
  shape = (16, 16)
  grid = Grid.from_step(shape, 2 * np.pi / 16)
  rng = np.random.default_rng(5)
  v = tuple(A(rng.standard_normal(shape), o) for o in grid.cell_faces)
  q1 = fvm.solve_fast_diag(v, grid).data
  q2, iters = fvm.solve_cg(v, grid, 1e-12, 1e-12, return_iterations=True)
  # the solutions differ by the null space (a constant)
  np.testing.assert_allclose(q1 - q1.mean(), q2.data - q2.data.mean(), atol=1e-9)
  assert 0 < iters < 10 * 256


# ------------------------------------------------------------- full step
def test_stable_time_step_config_values():
  """dt of the 64^2 (v_max 2) and 256^2 (v_max 7) configurations, CFL 0.5."""
  g64 = Grid.from_domain((64, 64), ((0, 2 * np.pi), (0, 2 * np.pi)))
  g256 = Grid.from_domain((256, 256), ((0, 2 * np.pi), (0, 2 * np.pi)))
  assert fvm.stable_time_step(2.0, 0.5, 1e-3, g64) == pytest.approx(
      0.02454369260617026, rel=1e-15)
  assert fvm.stable_time_step(7.0, 0.5, 1e-3, g256) == pytest.approx(
      0.0017531209004407328, rel=1e-15)
  with pytest.raises(ValueError):
    fvm.stable_time_step(1e-3, 0.5, 10.0, g64)


def _sinusoidal_velocity(grid):
  """equations_test.py:41-49."""
  size = [n * h for n, h in zip(grid.shape, grid.step)]
  vs = [np.sin(2. * np.pi * g / s)
        for g, s in zip(grid.mesh(grid.cell_center, np.float64), size)]
  vs = vs[1:] + vs[:1]
  return tuple(A(v, o) for v, o in zip(vs, grid.cell_faces))


def test_divergence_and_momentum_invariants():
  """equations_test.py:84-163 with the Van Leer convection of this path:
  divergence stays ~0 and momentum is conserved without forcing."""
  grid = Grid.from_step((32, 32), (1., 1.))
  step = fvm.FVMStep(grid, dt=1e-2, viscosity=1e-4, convect='van_leer')
  v0 = fvm.projection(_sinusoidal_velocity(grid), grid)
  v = step.rollout(v0, 100)
  div = go.divergence(v, grid).data
  assert np.abs(div).max() < 1e-10
  for a, b in zip(v0, v):
    assert abs(a.data.sum() - b.data.sum()) < 1e-9


def test_forcing_momentum_budget():
  """equations_test.py:150-163: momentum changes by sum(forcing) * dt * steps
  (constant forcing, zero linear coefficient, starting from rest)."""
  grid = Grid.from_domain((32, 32), ((0, 2 * np.pi), (0, 2 * np.pi)))
  forcing = go.Forcing(scale=1.0, wavenumber=1, linear_coefficient=0.0)
  zeros = tuple(A(np.zeros((32, 32)), o) for o in grid.cell_faces)
  fsum = sum(f.sum() for f in forcing(zeros, grid))
  step = fvm.FVMStep(grid, dt=1e-2, viscosity=1e-3, forcing=forcing)
  v = step.rollout(zeros, 20)
  mom = sum(u.data.sum() for u in v)
  np.testing.assert_allclose(mom, fsum * 0.01 * 20, atol=1e-9)


def _taylor_green(grid, t, nu):
  """validation_problems.py:58-106."""
  scale = np.exp(-2 * nu * t)
  ux, uy = grid.mesh(grid.cell_faces[0], np.float64)
  vx, vy = grid.mesh(grid.cell_faces[1], np.float64)
  return (A(scale * np.cos(ux) * np.sin(uy), grid.cell_faces[0]),
          A(-scale * np.sin(vx) * np.cos(vy), grid.cell_faces[1]))


def test_taylor_green_convergence():
  """validation_test.py:34-169 pattern with the Van Leer step: the error
  against the closed form shrinks at (close to) second order in dx."""
  nu, t_end = 1e-2, 0.5
  errs = []
  for n in (16, 32):
    grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
    dt = 0.05 * grid.step[0]
    steps = int(round(t_end / dt))
    step = fvm.FVMStep(grid, dt=dt, viscosity=nu, convect='van_leer')
    v = step.rollout(_taylor_green(grid, 0, nu), steps)
    ref = _taylor_green(grid, steps * dt, nu)
    errs.append(max(np.abs(a.data - b.data).max() for a, b in zip(v, ref)))
  assert errs[1] < 0.1
  assert errs[0] / errs[1] > 2.5


# ------------------------------------------------------ implicit diffusion
def _dense_laplacian(grid):
  """Kronecker sum of the periodic 1-D second-difference matrices."""
  mats = []
  for n, h in zip(grid.shape, grid.step):
    m = -2 * np.eye(n) + np.roll(np.eye(n), 1, 0) + np.roll(np.eye(n), -1, 0)
    mats.append(m / h ** 2)
  nx, ny = grid.shape
  return np.kron(mats[0], np.eye(ny)) + np.kron(np.eye(nx), mats[1])


def test_implicit_diffusion_keeps_constants():
  """diffusion_test.py:58-76: constant velocity is unchanged (atol 1e-6)."""
  grid = Grid.from_step((100, 100), (1., 1.))
  v = tuple(A(np.ones((100, 100), np.float32), o) for o in grid.cell_faces)
  out = fvm.solve_fast_diag_diffusion(v, 1.0, 0.1, grid)
  for a, b in zip(out, v):
    np.testing.assert_allclose(a.data, b.data, atol=1e-6)


def test_implicit_diffusion_residual():
  """diffusion_test.py:31-42: x - nu dt lap(x) == b for random b (4x4,
  step 1, nu = 1, dt = 0.5), with the oracle's periodic Laplacian."""
  grid = Grid.from_step((4, 4), (1., 1.))
  b = np.random.RandomState(0).randn(4, 4)
  (x,) = fvm.solve_fast_diag_diffusion((A(b, (1, .5)),), 1.0, 0.5, grid)
  resid = x.data - 1.0 * 0.5 * go.laplacian(x, grid).data
  np.testing.assert_allclose(resid, b, atol=1e-12)


def test_implicit_diffusion_dense_brute_force():
  """(1 - nu dt L)^-1 b by a dense solve on a rectangular periodic grid."""
  grid = Grid.from_domain((8, 6), ((0, 2.), (0, 3.)))
  b = np.random.RandomState(1).randn(8, 6)
  nu, dt = 0.3, 0.7
  dense = np.linalg.solve(np.eye(48) - nu * dt * _dense_laplacian(grid), b.ravel())
  (x,) = fvm.solve_fast_diag_diffusion((A(b, (.5, 1)),), nu, dt, grid)
  np.testing.assert_allclose(x.data.ravel(), dense, atol=1e-12)


def test_implicit_diffusion_fourier_mode():
  """A single mode cos(2 pi k i / n) is scaled by 1 / (1 - nu dt lambda_k),
  lambda_k = (2 cos(2 pi k / n) - 2) / h^2."""
  n, k, nu, dt = 32, 3, 0.05, 0.2
  grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  h = grid.step[0]
  i = np.arange(n)
  mode = np.cos(2 * np.pi * k * i / n)[:, None] * np.ones((1, n))
  (x,) = fvm.solve_fast_diag_diffusion((A(mode, (1, .5)),), nu, dt, grid)
  lam = (2 * np.cos(2 * np.pi * k / n) - 2) / h ** 2
  np.testing.assert_allclose(x.data, mode / (1 - nu * dt * lam), atol=1e-12)


def test_cg_diffusion_keeps_constants():
  """diffusion_test.py:58-76 with solve_cg: r0 = 0 for a constant field, so
  no iteration runs and x = x0 = b exactly."""
  grid = Grid.from_step((100, 100), (1., 1.))
  v = tuple(A(np.ones((100, 100), np.float32), o) for o in grid.cell_faces)
  out, iters = fvm.solve_cg_diffusion(v, 1.0, 0.1, grid, atol=1e-6, return_iterations=True)
  assert iters == [0, 0]
  for a, b in zip(out, v):
    np.testing.assert_array_equal(a.data, b.data)


def test_cg_diffusion_dense_brute_force_and_fast_diag():
  """The CG solution of (1 - nu dt L) x = b against a dense solve (float64,
  tight tolerance) and, in float32, against the fast-diagonal solve; a wrong
  sign of the nu dt L term or a dropped x0 would miss both."""
  grid = Grid.from_domain((8, 6), ((0, 2.), (0, 3.)))
  b = np.random.RandomState(1).randn(8, 6)
  nu, dt = 0.3, 0.7
  dense = np.linalg.solve(np.eye(48) - nu * dt * _dense_laplacian(grid), b.ravel())
  ((x,), iters) = fvm.solve_cg_diffusion((A(b, (.5, 1)),), nu, dt, grid, rtol=1e-13,
                                         atol=0.0, return_iterations=True)
  np.testing.assert_allclose(x.data.ravel(), dense, atol=1e-10)
  assert 0 < iters[0] <= 48
  grid = Grid.from_domain((32, 32), ((0, 2 * np.pi), (0, 2 * np.pi)))
  u = np.random.RandomState(2).randn(32, 32).astype(np.float32)
  (xc,) = fvm.solve_cg_diffusion((A(u, (1, .5)),), 0.05, 0.2, grid, atol=1e-6)
  (xf,) = fvm.solve_fast_diag_diffusion((A(u, (1, .5)),), 0.05, 0.2, grid)
  assert xc.data.dtype == np.float32
  assert np.abs(xc.data - xf.data).max() < 1e-5
  # true residual (float64) of the float32 iterate: the recursive one is
  # <= atol, the float32 rounding of 1024 entries adds ~1e-6 each
  x64 = A(xc.data.astype(np.float64), xc.offset)
  resid = x64.data - (0.05 * 0.2) * go.laplacian(x64, grid).data - u
  assert np.linalg.norm(resid) <= 1e-4 * np.linalg.norm(u)


def test_cg_diffusion_step_matches_fast_diag_step():
  """ImplicitDiffusionStep with diffusion='cg' tracks the fast-diagonal
  step (the same system, solved to atol 1e-6)."""
  grid = Grid.from_domain((32, 32), ((0, 2 * np.pi), (0, 2 * np.pi)))
  v0 = fvm.projection(_sinusoidal_velocity(grid), grid)
  v0 = tuple(A(a.data.astype(np.float32), a.offset) for a in v0)
  a = fvm.ImplicitDiffusionStep(grid, dt=0.05, viscosity=1e-2).rollout(v0, 5)
  b = fvm.ImplicitDiffusionStep(grid, dt=0.05, viscosity=1e-2, diffusion='cg').rollout(v0, 5)
  for x, y in zip(a, b):
    assert np.abs(x.data - y.data).max() < 1e-5


def test_implicit_diffusion_step_invariants():
  """equations_test.py:198-245 (the implicit_diffusion_navier_stokes cases):
  divergence stays ~0 and momentum is conserved without forcing."""
  grid = Grid.from_step((32, 32), (1., 1.))
  step = fvm.ImplicitDiffusionStep(grid, dt=1e-2, viscosity=1e-4)
  v0 = fvm.projection(_sinusoidal_velocity(grid), grid)
  v = step.rollout(v0, 50)
  assert np.abs(go.divergence(v, grid).data).max() < 1e-10
  for a, b in zip(v0, v):
    assert abs(a.data.sum() - b.data.sum()) < 1e-9


# ------------------------------------------------------ Runge-Kutta steppers
@pytest.mark.parametrize("name,atol", [('forward_euler', 1e-2), ('midpoint_rk2', 3e-5),
                                       ('heun_rk2', 3e-5), ('classic_rk4', 4e-7)])
def test_rk_harmonic_oscillator(name, atol):
  """time_stepping_test.py:27-90: x' = (x1, -x0) from (1, 1) with the
  identity projection, dt 1e-2, trajectory of 5 x 20 steps against the
  closed form, with the reference's tolerance for each stepper (fp64)."""
  a, b = fvm.TABLEAUS[name]
  F = lambda x: (x[1], -x[0])  # noqa: E731
  step = fvm.navier_stokes_rk(a, b, F, lambda x: x, 1e-2)
  x = (np.float64(1.0), np.float64(1.0))
  theta, r = np.arctan(1.0), np.sqrt(2.0)
  for outer in range(5):
    for _ in range(20):
      x = step(x)
    t = 1e-2 * 20 * (outer + 1)
    np.testing.assert_allclose(x, (r * np.sin(t + theta), r * np.cos(t + theta)),
                               atol=atol, rtol=0)


def test_rk_orders_of_accuracy():
  """Global error on the oscillator falls as dt^p: p = 1, 2, 2, 4."""
  for name, order in (('forward_euler', 1), ('midpoint_rk2', 2), ('heun_rk2', 2),
                      ('classic_rk4', 4)):
    a, b = fvm.TABLEAUS[name]
    errs = []
    for dt in (0.02, 0.01):
      step = fvm.navier_stokes_rk(a, b, lambda x: (x[1], -x[0]), lambda x: x, dt)
      x = (1.0, 0.0)
      for _ in range(int(round(1.0 / dt))):
        x = step(x)
      errs.append(np.hypot(x[0] - np.cos(1.0), x[1] + np.sin(1.0)))
    assert 0.8 * order < np.log2(errs[0] / errs[1]) < 1.25 * order + 0.3, (name, errs)


def test_rk_forward_euler_equals_euler_step():
  """The generic driver with b = [1] is the fused Euler step bit for bit."""
  grid = Grid.from_step((16, 16), (1., 1.))
  v0 = fvm.projection(_sinusoidal_velocity(grid), grid)
  e = fvm.FVMStep(grid, dt=1e-2, viscosity=1e-3)
  a, b = fvm.TABLEAUS['forward_euler']
  F = lambda xs: tuple(k.data for k in e.explicit_terms(tuple(A(x, u.offset) for x, u in zip(xs, v0))))  # noqa: E731
  P = lambda xs: tuple(u.data for u in fvm.projection(tuple(A(x, u.offset) for x, u in zip(xs, v0)), grid))  # noqa: E731
  out = fvm.navier_stokes_rk(a, b, F, P, 1e-2)(tuple(u.data for u in v0))
  for x, y in zip(out, e(v0)):
    np.testing.assert_array_equal(x, y.data)


def test_rk4_step_invariants():
  """equations_test.py:84-163 pattern with classic_rk4: divergence ~0 and
  momentum conserved."""
  grid = Grid.from_step((32, 32), (1., 1.))
  step = fvm.FVMStep(grid, dt=1e-2, viscosity=1e-4, time_stepper='classic_rk4')
  v0 = fvm.projection(_sinusoidal_velocity(grid), grid)
  v = step.rollout(v0, 20)
  assert np.abs(go.divergence(v, grid).data).max() < 1e-10
  for a_, b_ in zip(v0, v):
    assert abs(a_.data.sum() - b_.data.sum()) < 1e-9


# ------------------------------------------------------ linear advection
def test_linear_advection_translates_gaussian():
  """advection_test.py:84-98 (linear_1D): 1000 Euler steps at Courant 0.01
  with unit velocity move a smooth profile by 10 cells (atol 5e-2)."""
  shape = (101,)
  grid = Grid.from_step(shape, 1. / 101)
  c = A(_gaussian(grid, grid.cell_center), grid.cell_center)
  v = _unit_velocity(grid)
  dt = 0.01 * grid.step[0]
  c0 = c
  for _ in range(1000):
    c = A(c.data + dt * go.advect_linear(c, v, grid).data, c.offset)
  np.testing.assert_allclose(c.data, np.roll(c0.data, 10), atol=5e-2)


def test_convect_linear_matches_hand_written_stencil():
  """convect_linear (advection.py:184-207) written out per cell from the
  face values: u-face (u[i]+u[i+1])/2 squared, the corner product
  (u[i,j]+u[i,j+1])/2 * (v[i,j]+v[i+1,j])/2, their backward differences."""
  rng = np.random.default_rng(3)
  nx, ny = 6, 5
  grid = Grid.from_domain((nx, ny), ((0, 1.5), (0, 2.5)))
  dx, dy = grid.step
  u = rng.standard_normal((nx, ny))
  v = rng.standard_normal((nx, ny))
  vel = (A(u, grid.cell_faces[0]), A(v, grid.cell_faces[1]))
  got_u = go.advect_linear(vel[0], vel, grid).data
  got_v = go.advect_linear(vel[1], vel, grid).data
  U = lambda i, j: u[i % nx, j % ny]  # noqa: E731
  V = lambda i, j: v[i % nx, j % ny]  # noqa: E731
  uu = lambda i, j: (0.5 * U(i, j) + 0.5 * U(i + 1, j)) ** 2  # noqa: E731  u-face of CV(u)
  vv = lambda i, j: (0.5 * V(i, j) + 0.5 * V(i, j + 1)) ** 2  # noqa: E731  v-face of CV(v)
  uv = lambda i, j: (0.5 * U(i, j) + 0.5 * U(i, j + 1)) * (0.5 * V(i, j) + 0.5 * V(i + 1, j))  # noqa: E731
  for i in range(nx):
    for j in range(ny):
      eu = -((uu(i, j) - uu(i - 1, j)) / dx + (uv(i, j) - uv(i, j - 1)) / dy)
      ev = -((uv(i, j) - uv(i - 1, j)) / dx + (vv(i, j) - vv(i, j - 1)) / dy)
      assert abs(got_u[i, j] - eu) <= 1e-12 * max(1.0, abs(eu))
      assert abs(got_v[i, j] - ev) <= 1e-12 * max(1.0, abs(ev))


def test_taylor_green_linear_second_order():
  """validation_test.py:34-169 with convect_linear: the error against the
  closed form falls at second order (ratio ~4 per halving of h)."""
  nu, t_end = 1e-2, 0.5
  errs = []
  for n in (16, 32):
    grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
    dt = 0.05 * grid.step[0]
    steps = int(round(t_end / dt))
    step = fvm.FVMStep(grid, dt=dt, viscosity=nu, convect='linear')
    v = step.rollout(_taylor_green(grid, 0, nu), steps)
    ref = _taylor_green(grid, steps * dt, nu)
    errs.append(max(np.abs(a.data - b.data).max() for a, b in zip(v, ref)))
  assert errs[0] / errs[1] > 3.5


# ------------------------------------------------------ upwind advection
def test_upwind_advection_translates_gaussian():
  """advection_test.py:99-104 (upwind_1D): 100 Euler steps at Courant 0.5
  with unit velocity move the profile by 50 cells (atol 7e-2)."""
  shape = (101,)
  grid = Grid.from_step(shape, 1. / 101)
  x = (np.arange(101) - 51) / 101.  # advection_test.py:30-34 profile
  c = A(np.exp(-30. * x ** 2), grid.cell_center)
  v = _unit_velocity(grid)
  dt = 0.5 * grid.step[0]
  c0 = c
  for _ in range(100):
    c = A(c.data + dt * go.advect_upwind(c, v, grid).data, c.offset)
  np.testing.assert_allclose(c.data, np.roll(c0.data, 50), atol=7e-2)


def test_upwind_velocity_advection_matches_hand_written_stencil():
  """advect_upwind of each velocity component, written out per cell: face
  velocities averaged, the upwind neighbour chosen by their sign."""
  rng = np.random.default_rng(4)
  nx, ny = 6, 5
  grid = Grid.from_domain((nx, ny), ((0, 1.5), (0, 2.5)))
  dx, dy = grid.step
  u = rng.standard_normal((nx, ny))
  v = rng.standard_normal((nx, ny))
  vel = (A(u, grid.cell_faces[0]), A(v, grid.cell_faces[1]))
  got_u = go.advect_upwind(vel[0], vel, grid).data
  got_v = go.advect_upwind(vel[1], vel, grid).data
  U = lambda i, j: u[i % nx, j % ny]  # noqa: E731
  V = lambda i, j: v[i % nx, j % ny]  # noqa: E731

  def up(vel_face, lo, hi):
    return (lo if vel_face > 0 else hi) * vel_face

  fxu = lambda i, j: up(0.5 * U(i, j) + 0.5 * U(i + 1, j), U(i, j), U(i + 1, j))  # noqa: E731
  fyu = lambda i, j: up(0.5 * V(i, j) + 0.5 * V(i + 1, j), U(i, j), U(i, j + 1))  # noqa: E731
  fxv = lambda i, j: up(0.5 * U(i, j) + 0.5 * U(i, j + 1), V(i, j), V(i + 1, j))  # noqa: E731
  fyv = lambda i, j: up(0.5 * V(i, j) + 0.5 * V(i, j + 1), V(i, j), V(i, j + 1))  # noqa: E731
  for i in range(nx):
    for j in range(ny):
      eu = -((fxu(i, j) - fxu(i - 1, j)) / dx + (fyu(i, j) - fyu(i, j - 1)) / dy)
      ev = -((fxv(i, j) - fxv(i - 1, j)) / dx + (fyv(i, j) - fyv(i, j - 1)) / dy)
      assert abs(got_u[i, j] - eu) <= 1e-12 * max(1.0, abs(eu))
      assert abs(got_v[i, j] - ev) <= 1e-12 * max(1.0, abs(ev))


# ------------------------------------------------- semi-Lagrangian advection
def test_map_coordinates_wrap_linear_vs_scipy():
  """The oracle's map_coordinates(order=1, mode='wrap') against SciPy's
  independent implementation (mode 'grid-wrap', the periodic one) at random
  points far outside the array, 1-D to 3-D."""
  from scipy import ndimage
  rng = np.random.default_rng(7)
  for shape in [(11,), (6, 9), (4, 5, 3)]:
    a = rng.standard_normal(shape)
    pts = [rng.uniform(-3 * n, 3 * n, size=50) for n in shape]
    want = ndimage.map_coordinates(a, pts, order=1, mode='grid-wrap')
    np.testing.assert_allclose(go.map_coordinates_wrap_linear(a, pts), want, atol=1e-12)


def test_semilagrangian_integer_displacement_is_a_roll():
  """dt v = 2 h along axis 0 and -3 h along axis 1: every departure point
  is a grid point, so the step is an exact periodic roll."""
  grid = Grid.from_domain((16, 12), ((0, 2.), (0, 3.)))
  rng = np.random.default_rng(8)
  c = A(rng.standard_normal((16, 12)), grid.cell_center)
  hx, hy = grid.step
  dt = 0.5
  v = (A(np.full((16, 12), 2 * hx / dt), grid.cell_faces[0]),
       A(np.full((16, 12), -3 * hy / dt), grid.cell_faces[1]))
  out = go.advect_step_semilagrangian(c, v, dt, grid)
  np.testing.assert_allclose(out.data, np.roll(np.roll(c.data, 2, 0), -3, 1), atol=1e-12)


def test_semilagrangian_half_cell_is_two_point_average():
  """Courant 0.5 with unit velocity along axis 0: out[i] = (c[i-1] + c[i]) / 2,
  with u interpolated to the cell centre from its faces."""
  grid = Grid.from_step((10, 6), (0.1, 0.2))
  c = A(np.random.default_rng(9).standard_normal((10, 6)), grid.cell_center)
  v = (A(np.ones((10, 6)), grid.cell_faces[0]), A(np.zeros((10, 6)), grid.cell_faces[1]))
  out = go.advect_step_semilagrangian(c, v, 0.05, grid)
  np.testing.assert_allclose(out.data, 0.5 * (np.roll(c.data, 1, 0) + c.data), atol=1e-12)


@pytest.mark.parametrize("shape", [(101,), (128, 128)])
def test_semilagrangian_translates_gaussian(shape):
  """advection_test.py:140-170 (semilagrangian_1D, and 2-D with the
  velocity along axis 0): 100 steps at Courant 0.5 move the profile by 50
  cells (atol 7e-2)."""
  grid = Grid.from_step(shape, tuple(1. / s for s in shape))
  n = shape[0]
  x = (np.arange(n) - (n // 2 + 1)) / n  # advection_test.py:30-34 profile
  prof = np.exp(-30. * x ** 2).astype(np.float32)
  c = A(np.broadcast_to(prof.reshape((n,) + (1,) * (len(shape) - 1)), shape).copy(),
        grid.cell_center)
  v = tuple(A(np.full(shape, 1.0 if ax == 0 else 0.0, np.float32), off)
            for ax, off in enumerate(grid.cell_faces))
  dt = 0.5 * min(grid.step)
  c0 = c
  for _ in range(100):
    c = go.advect_step_semilagrangian(c, v, dt, grid)
  assert c.data.dtype == np.float32
  np.testing.assert_allclose(c.data, np.roll(c0.data, 50, axis=0), atol=7e-2)
  with pytest.raises(ValueError):
    go.advect_step_semilagrangian(
        c, v, dt, Grid.from_domain(shape, tuple((1., 2.) for _ in shape)))
