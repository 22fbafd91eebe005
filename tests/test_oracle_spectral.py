"""Pins the CPU oracle of the pseudospectral path (no GPU needed)."""
import numpy as np
import pytest

from oracle import dft
from oracle import spectral as sp
from oracle.grid_ops import Forcing, Grid

TWO_PI = 2 * np.pi


def _grid(n):
  return Grid.from_domain((n, n), ((0, TWO_PI), (0, TWO_PI)))


class _ODE:
  def __init__(self, explicit_terms, implicit_terms, implicit_solve):
    self.explicit_terms = explicit_terms
    self.implicit_terms = implicit_terms
    self.implicit_solve = implicit_solve


def _harmonic(x0, t):
  theta = np.arctan(x0[0] / x0[1])
  r = np.linalg.norm(x0)
  return r * np.stack([np.sin(t + theta), np.cos(t + theta)])


# spectral/time_stepping_test.py:42-122, tolerance column of crank_nicolson_rk4
CNRK4_PROBLEMS = [
    ("zero", lambda x: 0 * x, lambda x: 0 * x, lambda x, e: x, np.ones(10),
     lambda x0, t: x0, 1e-12),
    ("constant", lambda x: 5 * np.ones_like(x), lambda x: 0 * x,
     lambda x, e: x, np.ones(3), lambda x0, t: x0 + 5 * t, 1e-12),
    ("linear_explicit", lambda x: x, lambda x: 0 * x, lambda x, e: x,
     np.arange(3.0), lambda x0, t: np.arange(3) * np.exp(t), 1e-9),
    ("linear_implicit", lambda x: 0 * x, lambda x: x,
     lambda x, e: x / (1 - e), np.arange(3.0),
     lambda x0, t: np.arange(3) * np.exp(t), 1e-5),
    ("linear_semi_implicit", lambda x: x / 2, lambda x: x / 2,
     lambda x, e: x / (1 - e / 2), np.arange(3.0),
     lambda x0, t: np.arange(3.0) * np.exp(t), 1e-6),
    ("harmonic_explicit", lambda x: np.stack([x[1], -x[0]]), np.zeros_like,
     lambda x, e: x, np.ones(2), _harmonic, 5e-11),
    ("harmonic_implicit", np.zeros_like, lambda x: np.stack([x[1], -x[0]]),
     lambda x, e: np.stack([x[0] + e * x[1], x[1] - e * x[0]]) / (1 + e ** 2),
     np.ones(2), _harmonic, 1e-6),
]


@pytest.mark.parametrize("name,f,g,ginv,x0,closed,tol", CNRK4_PROBLEMS,
                         ids=[p[0] for p in CNRK4_PROBLEMS])
def test_cnrk4_closed_forms(name, f, g, ginv, x0, closed, tol):
  step = sp.crank_nicolson_rk4(_ODE(f, g, ginv), 1e-2)
  x = x0
  inner = 10 if name in ("zero", "constant") else 20
  for outer in range(5):
    for _ in range(inner):
      x = step(x)
    t = 1e-2 * inner * (outer + 1)
    np.testing.assert_allclose(x, closed(x0, t), atol=tol, rtol=0)


# the crank_nicolson_rk3 column of the same table (time_stepping_test.py:42-122,
# tolerances index 2 of ALL_TIME_STEPPERS, 124-130)
CNRK3_TOL = {"zero": 1e-12, "constant": 1e-12, "linear_explicit": 1e-6,
             "linear_implicit": 1e-5, "linear_semi_implicit": 2e-6,
             "harmonic_explicit": 6e-8, "harmonic_implicit": 2e-6}


def _closed_form_errors(step, x0, closed, name, dt=1e-2):
  x = x0
  inner = 10 if name in ("zero", "constant") else 20
  errs = []
  for outer in range(5):
    for _ in range(inner):
      x = step(x)
    errs.append(float(np.abs(x - closed(x0, dt * inner * (outer + 1))).max()))
  return max(errs)


@pytest.mark.parametrize("name,f,g,ginv,x0,closed,tol", CNRK4_PROBLEMS,
                         ids=[p[0] for p in CNRK4_PROBLEMS])
def test_cnrk3_closed_forms(name, f, g, ginv, x0, closed, tol):
  step = sp.crank_nicolson_rk3(_ODE(f, g, ginv), 1e-2)
  assert _closed_form_errors(step, x0, closed, name) <= CNRK3_TOL[name]


def test_cnrk3_closed_form_detects_wrong_coefficients():
  """A tableau off by 1e-3 in one coefficient fails the harmonic-oscillator
  bound that the Williamson coefficients meet."""
  name, f, g, ginv, x0, closed, _ = CNRK4_PROBLEMS[5]
  ode = _ODE(f, g, ginv)
  good = sp.low_storage_rk_crank_nicolson(sp.CNRK3_ALPHAS, sp.CNRK3_BETAS,
                                          sp.CNRK3_GAMMAS, ode, 1e-2)
  bad = sp.low_storage_rk_crank_nicolson(
      sp.CNRK3_ALPHAS, sp.CNRK3_BETAS[:2] + (sp.CNRK3_BETAS[2] + 1e-3,),
      sp.CNRK3_GAMMAS, ode, 1e-2)
  assert _closed_form_errors(good, x0, closed, name) <= CNRK3_TOL[name]
  assert _closed_form_errors(bad, x0, closed, name) > 10 * CNRK3_TOL[name]


@pytest.mark.parametrize("n,rows_pos,rows_neg,cols", [
    (64, 21, 21, 22), (128, 42, 43, 43), (256, 85, 85, 86),
    (512, 170, 171, 171)])
def test_brick_wall_filter_sets(n, rows_pos, rows_neg, cols):
  """spectral/utils.py:128-134 with Python int/floor: rows [0, a//2) and
  [n + (-a)//2, n), a = int(2n/3); columns [0, int(2(n/2+1)/3))."""
  f = sp.brick_wall_filter_2d(_grid(n))
  assert f.shape == (n, n // 2 + 1)
  keep_rows = np.where(f[:, 0] == 1)[0]
  expected = np.r_[np.arange(rows_pos), np.arange(n - rows_neg, n)]
  np.testing.assert_array_equal(keep_rows, expected)
  assert f[0].sum() == cols and f[0, cols - 1] == 1 and f[0, cols] == 0
  assert f.sum() == (rows_pos + rows_neg) * cols


def test_nonlinear_term_closed_form():
  """omega = cos x + cos 2y has psi = cos x + cos(2y)/4, v = (psi_y, -psi_x);
  -(v . grad omega) = 1.5 sin x sin 2y.  Pins signs, axes, the stream
  function solve and the filter (modes well inside the 2/3 band)."""
  n = 32
  grid = _grid(n)
  x, y = grid.mesh((0, 0), np.float64)
  w = np.cos(x) + np.cos(2 * y)
  eq = sp.NavierStokes2D(1e-3, grid, dtype=np.float64)
  out = eq.explicit_terms(np.fft.rfftn(w))
  np.testing.assert_allclose(np.fft.irfftn(out, s=(n, n), axes=(0, 1)),
                             1.5 * np.sin(x) * np.sin(2 * y), atol=1e-12)


def test_forcing_curl_closed_form():
  """curl of (sin 4y, 0): -2 pi i ky * rfft2(sin 4y) is -2 n^2 at [0, 4]."""
  n = 64
  eq = sp.ForcedNavierStokes2D(1e-3, _grid(n), smooth=True, dtype=np.float64)
  fc = eq.forcing_curl
  assert fc[0, 4] == pytest.approx(-2 * n * n, rel=1e-12)
  mask = np.ones_like(fc, bool)
  mask[0, 4] = False
  assert np.abs(fc[mask]).max() < 1e-9


def test_vorticity_to_velocity_taylor_green():
  """spectral/utils_test.py:61-85 pattern: Taylor-Green vorticity
  -2 cos x cos y maps back to u = cos x sin y, v = -sin x cos y."""
  n = 32
  grid = _grid(n)
  x, y = grid.mesh((0, 0), np.float64)
  w_hat = np.fft.rfftn(-2 * np.cos(x) * np.cos(y))
  vxh, vyh = sp.vorticity_to_velocity(grid, np.float64)(w_hat)
  np.testing.assert_allclose(np.fft.irfftn(vxh, s=(n, n), axes=(0, 1)),
                             np.cos(x) * np.sin(y), atol=1e-12)
  np.testing.assert_allclose(np.fft.irfftn(vyh, s=(n, n), axes=(0, 1)),
                             -np.sin(x) * np.cos(y), atol=1e-12)


def test_taylor_green_decay_cnrk4():
  """spectral/equations_test.py:263-303 pattern (smaller grid): the vortex is
  a steady Euler solution, so CN-RK4 must give exp(-2 nu t) decay."""
  n, nu = 32, 1e-2
  grid = _grid(n)
  x, y = grid.mesh((0, 0), np.float64)
  w0 = -2 * np.cos(x) * np.cos(y)
  eq = sp.NavierStokes2D(nu, grid, drag=0.0, dtype=np.float64)
  step = sp.crank_nicolson_rk4(eq, 0.01)
  w = sp.rollout(step, np.fft.rfftn(w0), 50)
  np.testing.assert_allclose(np.fft.irfftn(w, s=(n, n), axes=(0, 1)),
                             w0 * np.exp(-2 * nu * 0.5), atol=1e-9)


def test_drag_and_forcing_steady_balance():
  """With forcing curl F at mode (0, 4) the linear steady state is
  -F / L there; the step preserves it (the mode is a steady Euler flow)."""
  n = 32
  grid = _grid(n)
  eq = sp.ForcedNavierStokes2D(1e-2, grid, smooth=True, dtype=np.float64)
  w = np.zeros((n, n // 2 + 1), np.complex128)
  w[0, 4] = -eq.forcing_curl[0, 4] / eq.linear_term[0, 4]
  step = sp.crank_nicolson_rk4(eq, 0.01)
  out = sp.rollout(step, w, 20)
  np.testing.assert_allclose(out, w, atol=1e-9 * abs(w[0, 4]))


def test_fp32_rfft_matches_dft():
  rng = np.random.default_rng(0)
  x = rng.standard_normal((16, 16)).astype(np.float32)
  np.testing.assert_allclose(np.fft.rfft2(x), dft.rfft2(x), atol=1e-4)
  xh = np.fft.rfft2(x)
  np.testing.assert_allclose(dft.irfft2(xh, 16), x, atol=1e-5)


def test_fp32_step_tracks_fp64():
  """The fp32 oracle stays within fp32 round-off of the fp64 run (10 steps,
  forced turbulence at 64^2)."""
  n = 64
  grid = _grid(n)
  rng = np.random.default_rng(7)
  w = np.fft.rfftn(rng.standard_normal((n, n)))
  w[np.abs(w) < 0] = 0
  filt = sp.brick_wall_filter_2d(grid, np.float64)
  w = w * filt
  outs = []
  for dt_ in (np.float32, np.float64):
    eq = sp.ForcedNavierStokes2D(1e-3, grid, smooth=True, dtype=dt_)
    ct = np.complex64 if dt_ == np.float32 else np.complex128
    step = sp.crank_nicolson_rk4(eq, 1e-3)
    outs.append(sp.rollout(step, w.astype(ct), 10))
  assert outs[0].dtype == np.complex64
  err = np.abs(outs[0] - outs[1]).max() / np.abs(outs[1]).max()
  assert err < 1e-5
