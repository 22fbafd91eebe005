"""Taylor-Green vortex against its closed form on the GPU
(jax_cfd/base/validation_test.py:34-169 pattern; the problem is
validation_problems.py:58-106: u = e^{-2 nu t} cos x sin y,
v = -e^{-2 nu t} sin x cos y on [0, 2 pi]^2).

The reference validates with convect_linear; the kernels here build the
Van Leer schemes, so the check is the one the mathematics fixes for them:
with dt = 0.05 h the error against the exact solution falls faster than
first order in the grid step at every halving of h (max-norm ratio > 2;
the fp32 oracle shows 3.06 then 2.55 from 64^2 to 256^2: second order in
smooth regions, the limiter clipping the extrema) and is below 1.5e-4 at
512^2, for forward Euler, classic RK4 and implicit diffusion -- a
size-independent property at sizes the fp32 oracle is too slow to run.  A
dropped term or a wrong sign stops the convergence.
"""

import functools

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def _cfd():
  import jax_cfd_amd.base as cfd
  return cfd


def _taylor_green(grid, t, nu):
  scale = np.exp(-2 * nu * t)
  cfd = _cfd()
  bc = cfd.boundaries.periodic_boundary_conditions(2)
  out = []
  for comp, off in enumerate(grid.cell_faces):
    x, y = grid.mesh(off)
    x, y = x.astype(np.float64), y.astype(np.float64)
    f = scale * np.cos(x) * np.sin(y) if comp == 0 else -scale * np.sin(x) * np.cos(y)
    out.append(f)
  state = tuple(cfd.grids.GridVariable(cfd.grids.GridArray(
      torch.as_tensor(f.astype(np.float32), device=torch.device('cuda', 0)), off, grid), bc)
      for f, off in zip(out, grid.cell_faces))
  return state, out


def _solver(kind, nu, dt, grid):
  cfd = _cfd()
  convect = lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v)  # noqa: E731
  if kind == 'euler':
    return cfd.equations.semi_implicit_navier_stokes(1.0, nu, dt, grid, convect=convect)
  if kind == 'rk4':
    return cfd.equations.semi_implicit_navier_stokes(
        1.0, nu, dt, grid, convect=convect, time_stepper=cfd.time_stepping.classic_rk4)
  return cfd.equations.implicit_diffusion_navier_stokes(1.0, nu, dt, grid, convect=convect)


@pytest.mark.parametrize("kind", ['euler', 'rk4', 'implicit_diffusion'])
def test_taylor_green_converges(kind):
  nu, t_end = 1e-2, 0.5
  errs = []
  for n in (64, 128, 256, 512):
    grid = _cfd().grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))
    dt = 0.05 * grid.step[0]
    steps = int(round(t_end / dt))
    print(f"taylor-green {kind} n={n} steps={steps}", flush=True)
    v0, _ = _taylor_green(grid, 0.0, nu)
    step = _solver(kind, nu, dt, grid)
    v = _cfd().funcutils.repeated(step, steps)(v0)
    _, exact = _taylor_green(grid, steps * dt, nu)
    err = max(float(np.abs(c.data.cpu().numpy().astype(np.float64) - e).max())
              for c, e in zip(v, exact))
    errs.append(err)
  ratios = [a / b for a, b in zip(errs, errs[1:])]
  print(f"  max errors {['%.2e' % e for e in errs]}, ratios {['%.2f' % r for r in ratios]}")
  assert all(r > 2.0 for r in ratios)
  assert errs[-1] < 1.5e-4


# The reference's own validation cases (validation_test.py:34-131), with its
# solvers (convect_linear), time steps (dynamic_time_step at its Courant
# numbers), end times and tolerances (np.testing.assert_allclose with the
# given atol and the default rtol 1e-7, computed vs analytic).
REFERENCE_CASES = [
    # name, n, viscosity, time stepper, implicit diffusion, courant, time, atol
    ('semi_implicit', 1024, 1e-3, 'forward_euler', False, 0.1, 10.0, 1e-5),
    ('semi_implicit_rk1', 512, 1e-2, 'forward_euler', False, 0.1, 40.0, 6e-6),
    ('semi_implicit_rk4', 512, 1e-2, 'classic_rk4', False, 0.1, 40.0, 8e-7),
    ('implicit_diffusion_rfft', 1024, 1e-3, None, True, 0.1, 10.0, 4e-5),
    ('implicit_diffusion_viscous', 1024, 0.5, None, True, 0.5, 1.0, 6e-4),
    # diffusion_solve=partial(solve_cg, atol=1e-6, maxiter=512)
    ('implicit_diffusion_cg', 1024, 1e-3, None, 'cg', 0.1, 10.0, 3e-5),
]


@pytest.mark.parametrize('case', REFERENCE_CASES, ids=[c[0] for c in REFERENCE_CASES])
def test_reference_taylor_green_cases(case):
  name, n, nu, stepper, implicit, courant, time, atol = case
  cfd = _cfd()
  grid = cfd.grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))
  v0, _ = _taylor_green(grid, 0.0, nu)
  dt = cfd.equations.dynamic_time_step(v0, courant, nu, grid, bool(implicit))
  steps = int(np.ceil(time / dt))
  print(f"reference case {name}: n={n} dt={dt:.4e} steps={steps}", flush=True)
  if implicit:
    solve = (functools.partial(cfd.diffusion.solve_cg, atol=1e-6, maxiter=512)
             if implicit == 'cg' else cfd.diffusion.solve_fast_diag)
    step = cfd.equations.implicit_diffusion_navier_stokes(
        1.0, nu, dt, grid, convect=cfd.advection.convect_linear, diffusion_solve=solve)
  else:
    step = cfd.equations.semi_implicit_navier_stokes(
        1.0, nu, dt, grid, convect=cfd.advection.convect_linear,
        time_stepper=getattr(cfd.time_stepping, stepper))
  v = cfd.funcutils.repeated(step, steps)(v0)
  _, exact = _taylor_green(grid, time, nu)
  for c, e in zip(v, exact):
    got = c.data.cpu().numpy().astype(np.float64)
    print(f"  max |error| {np.abs(got - e).max():.2e} (atol {atol})")
    np.testing.assert_allclose(got, e, atol=atol, rtol=1e-7)
