"""The step factories read the reference-style operator callables on the
host (no GPU needed): supported compositions map to a kernel spec, others
raise NotImplementedError (never a silent fallback)."""
import functools

import numpy as np
import pytest

import jax_cfd_amd.base as cfd
from jax_cfd_amd import _lib

TWO_PI = 2 * np.pi


def _grid(n=64):
  return cfd.grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))


def _vl(dt):
  return lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v)


def test_forced_configuration_spec():
  grid = _grid(256)
  dt = cfd.equations.stable_time_step(7.0, 0.5, 1e-3, grid)
  assert dt == pytest.approx(0.0017531209004407328, rel=1e-15)
  forcing = cfd.forcings.simple_turbulence_forcing(
      grid, constant_magnitude=1.0, constant_wavenumber=4,
      linear_coefficient=-0.1, forcing_type='kolmogorov')
  step = cfd.equations.semi_implicit_navier_stokes(
      density=1.0, viscosity=1e-3, dt=dt, grid=grid, convect=_vl(dt),
      forcing=forcing)
  s = step.spec
  assert s.n == 256 and s.viscosity == 1e-3 and s.linear_coefficient == -0.1
  y = (np.arange(256, dtype=np.float32) + np.float32(0.5)) * np.float32(grid.step[1])
  np.testing.assert_array_equal(np.asarray(s.forcing_x, np.float32),
                                np.sin(4 * y).astype(np.float32))
  assert s.convect == _lib.CONVECT_VAN_LEER


def test_unforced_configuration_spec():
  grid = _grid(64)
  step = cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, 0.02, grid,
                                                   convect=_vl(0.02))
  assert step.spec.linear_coefficient is None and step.spec.forcing_x is None


def test_kolmogorov_only_and_linear_only():
  grid = _grid(64)
  f1 = cfd.forcings.kolmogorov_forcing(grid, scale=2.0, k=3)
  s1 = cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, 0.02, grid,
                                                 convect=_vl(0.02), forcing=f1).spec
  assert s1.linear_coefficient == 0.0 and max(s1.forcing_x) > 1.9
  f2 = cfd.forcings.linear_forcing(grid, 0.3)
  s2 = cfd.equations.semi_implicit_navier_stokes(1.0, None, 0.02, grid,
                                                 convect=_vl(0.02), forcing=f2).spec
  assert s2.linear_coefficient == 0.3 and s2.viscosity is None
  assert not any(s2.forcing_x)


def test_default_convection_is_van_leer_using_limiters():
  """equations.py:89-92: convect=None means advect_van_leer_using_limiters."""
  grid = _grid(64)
  s = cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, 0.02, grid).spec
  assert s.convect == _lib.CONVECT_VAN_LEER_LIMITERS
  s2 = cfd.equations.semi_implicit_navier_stokes(
      1.0, 1e-3, 0.02, grid, convect=lambda v: tuple(
          cfd.advection.advect_van_leer_using_limiters(u, v, 0.02) for u in v)).spec
  assert s2 == s


@pytest.mark.parametrize("kw", [
    dict(convect=lambda v: tuple(cfd.advection.advect_van_leer(v[0], v, 0.02) for _ in v)),
    dict(convect=lambda v: tuple(cfd.advection.advect_van_leer(u, v, 0.01) for u in v)),
    dict(convect=_vl(0.02), forcing=cfd.forcings.kolmogorov_forcing(_grid(), swap_xy=True)),
    dict(convect=_vl(0.02), time_stepper=lambda equation, dt: None),
    dict(convect=_vl(0.02), time_stepper=functools.partial(
        cfd.time_stepping.navier_stokes_rk,
        cfd.time_stepping.ButcherTableau(a=[[1], [1, 1], [1, 1, 1], [1, 1, 1, 1]],
                                         b=[1, 1, 1, 1, 1]))),
    dict(convect=_vl(0.02), diffuse=lambda c, nu: c.array),
])
def test_unsupported_compositions_raise(kw):
  with pytest.raises(NotImplementedError):
    cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, 0.02, _grid(), **kw)


def test_non_square_grid_raises():
  grid = cfd.grids.Grid((64, 32), domain=((0, TWO_PI), (0, TWO_PI)))
  with pytest.raises(NotImplementedError):
    cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, 0.02, grid,
                                              convect=_vl(0.02))


def test_grid_matches_reference_geometry():
  """grids.py:499-658: steps, faces and fp32 meshes."""
  grid = cfd.grids.Grid((8, 8), domain=((0, TWO_PI), (0, TWO_PI)))
  assert grid.step == (TWO_PI / 8, TWO_PI / 8)
  assert grid.cell_faces == ((1.0, 0.5), (0.5, 1.0))
  x, y = grid.mesh((1.0, 0.5))
  assert x.dtype == np.float32
  np.testing.assert_allclose(x[:, 0], (np.arange(8) + 1) * TWO_PI / 8, rtol=1e-6)
  np.testing.assert_allclose(y[0], (np.arange(8) + .5) * TWO_PI / 8, rtol=1e-6)
  kx, ky = grid.rfft_mesh()
  assert kx.shape == (8, 5)
  np.testing.assert_allclose(2 * np.pi * ky[0], np.arange(5), atol=1e-5)
  np.testing.assert_allclose(2 * np.pi * kx[:, 0], [0, 1, 2, 3, -4, -3, -2, -1], atol=1e-5)


def test_implicit_diffusion_factory_spec():
  """implicit_diffusion_navier_stokes (equations.py:154-195): default
  convection is advect_van_leer_using_limiters; the diffusion solve maps to
  the implicit fast-diagonal plan with the kinematic viscosity as given."""
  grid = _grid(128)
  step = cfd.equations.implicit_diffusion_navier_stokes(2.0, 1e-3, 0.01, grid)
  s = step.spec
  assert s.diffusion == _lib.DIFFUSION_IMPLICIT_FAST_DIAG
  assert s.convect == _lib.CONVECT_VAN_LEER_LIMITERS
  assert s.viscosity == 1e-3 and s.density == 2.0 and s.linear_coefficient is None
  step = cfd.equations.implicit_diffusion_navier_stokes(
      1.0, 1e-3, 0.01, grid, convect=_vl(0.01),
      diffusion_solve=functools.partial(cfd.diffusion.solve_fast_diag, implementation='rfft'),
      pressure_solve=functools.partial(cfd.pressure.solve_cg, rtol=1e-5))
  assert step.spec.convect == _lib.CONVECT_VAN_LEER
  assert step.spec.pressure == _lib.PRESSURE_CG and step.spec.cg_rtol == 1e-5
  with pytest.raises(NotImplementedError):
    cfd.equations.implicit_diffusion_navier_stokes(
        1.0, 1e-3, 0.01, grid, diffusion_solve=lambda v, nu, dt: v)
  with pytest.raises(NotImplementedError):
    cfd.equations.implicit_diffusion_navier_stokes(
        1.0, 1e-3, 0.01, grid,
        diffusion_solve=functools.partial(cfd.diffusion.solve_fast_diag,
                                          implementation='matmul'))


def test_implicit_diffusion_cg_factory_spec():
  """diffusion_solve=solve_cg (diffusion.py:137-167) maps to the CG
  diffusion plan with its keywords; maxiter None is the 10 n^2 default."""
  grid = _grid(64)
  step = cfd.equations.implicit_diffusion_navier_stokes(
      1.0, 1e-3, 0.01, grid,
      diffusion_solve=functools.partial(cfd.diffusion.solve_cg, atol=1e-6, maxiter=512))
  s = step.spec
  assert s.diffusion == _lib.DIFFUSION_IMPLICIT_CG
  assert (s.diff_cg_rtol, s.diff_cg_atol, s.diff_cg_maxiter) == (1e-6, 1e-6, 512)
  s = cfd.equations.implicit_diffusion_navier_stokes(
      1.0, 1e-3, 0.01, grid, diffusion_solve=cfd.diffusion.solve_cg).spec
  assert s.diffusion == _lib.DIFFUSION_IMPLICIT_CG and s.diff_cg_maxiter == 0
  with pytest.raises(NotImplementedError):
    cfd.equations.implicit_diffusion_navier_stokes(
        1.0, 1e-3, 0.01, grid,
        diffusion_solve=functools.partial(cfd.diffusion.solve_cg, tol=1e-3))


@pytest.mark.parametrize("stepper,stages,a,b", [
    ('midpoint_rk2', 2, (0.5,), (0.0, 1.0)),
    ('heun_rk2', 2, (1.0,), (0.5, 0.5)),
    ('classic_rk4', 4, (0.5, 0.0, 0.5, 0.0, 0.0, 1.0), (1 / 6, 1 / 3, 1 / 3, 1 / 6))])
def test_runge_kutta_time_steppers_map_to_tableau(stepper, stages, a, b):
  """time_stepping.py:109-158: named steppers carry their tableaux into the
  plan; forward Euler stays on the fused kernels."""
  grid = _grid(64)
  step = cfd.equations.semi_implicit_navier_stokes(
      1.0, 1e-3, 0.01, grid, convect=_vl(0.01),
      time_stepper=getattr(cfd.time_stepping, stepper))
  assert (step.spec.rk_stages, step.spec.rk_a, step.spec.rk_b) == (stages, a, b)
  euler = cfd.equations.semi_implicit_navier_stokes(1.0, 1e-3, 0.01, grid, convect=_vl(0.01))
  assert euler.spec.rk_stages == 0
  custom = cfd.equations.semi_implicit_navier_stokes(
      1.0, 1e-3, 0.01, grid, convect=_vl(0.01),
      time_stepper=functools.partial(cfd.time_stepping.navier_stokes_rk,
                                     cfd.time_stepping.ButcherTableau(a=[], b=[2.0])))
  assert (custom.spec.rk_stages, custom.spec.rk_b) == (1, (2.0,))
  with pytest.raises(NotImplementedError):  # Runge-Kutta with the CG solve is not built
    cfd.equations.semi_implicit_navier_stokes(
        1.0, 1e-3, 0.01, grid, convect=_vl(0.01), pressure_solve=cfd.pressure.solve_cg,
        time_stepper=getattr(cfd.time_stepping, stepper))
