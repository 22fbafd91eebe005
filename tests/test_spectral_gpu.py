"""GPU parity of the pseudospectral path against the oracle (through the C ABI).

Tolerances (fp32/complex64): max-norm relative error <= 1e-5 for one call,
relative L2 <= 1e-3 after 100 steps; the fp64 oracle is reported beside.
"""
import numpy as np
import pytest
import torch

from oracle import initial_conditions as oic
from oracle import spectral as osp
from oracle.grid_ops import Grid as OGrid

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def _dev():
  return torch.device("cuda", 0)


def _maxrel(a, b):
  return float(np.abs(a - b).max() / np.abs(b).max())


def _rel_l2(a, b):
  return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _ogrid(n):
  return OGrid.from_domain((n, n), ((0, TWO_PI), (0, TWO_PI)))


def _sp():
  import jax_cfd_amd.spectral as sp
  return sp


def _grid(n):
  import jax_cfd_amd.base as cfd
  return cfd.grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))


def _w0(n, seed, vmax=1.0):
  """rfftn(curl_2d(filtered velocity)) -- the spectral runs' initial state."""
  v = oic.filtered_velocity_field(seed, _ogrid(n), vmax, 4)
  return oic.vorticity_hat_from_velocity(v, _ogrid(n)).astype(np.complex64)


def _eqs(n, forced=True):
  og = _ogrid(n)
  if forced:
    return (osp.ForcedNavierStokes2D(1e-3, og, True),
            osp.ForcedNavierStokes2D(1e-3, og, True, dtype=np.float64),
            _sp().equations.ForcedNavierStokes2D(1e-3, _grid(n), True))
  return (osp.NavierStokes2D(1e-3, og, drag=0.1),
          osp.NavierStokes2D(1e-3, og, drag=0.1, dtype=np.float64),
          _sp().equations.NavierStokes2D(1e-3, _grid(n), drag=0.1))


@pytest.mark.parametrize("n,forced", [(32, True), (64, True), (128, False), (256, True),
                                      (512, True), (1024, True)])
def test_explicit_terms_parity(n, forced):
  w = _w0(n, n)
  o32, _, eq = _eqs(n, forced)
  ref = o32.explicit_terms(w)
  out = eq.explicit_terms(torch.as_tensor(w, device=_dev())).cpu().numpy()
  err = _maxrel(out, ref)
  print(f"spectral explicit terms n={n}: max rel err {err:.2e}")
  assert err <= 1e-5


def test_explicit_terms_drop_imaginary_dc_nyquist():
  """irfftn ignores the imaginary parts of the DC/Nyquist columns after the
  axis-0 inverse; a non-Hermitian input must give the reference's answer."""
  n = 64
  rng = np.random.default_rng(1)
  w = _w0(n, 2)
  w = (w + 3.0 * (rng.standard_normal(w.shape) + 1j * rng.standard_normal(w.shape))
       ).astype(np.complex64)
  o32, _, eq = _eqs(n)
  ref = o32.explicit_terms(w)
  out = eq.explicit_terms(torch.as_tensor(w, device=_dev())).cpu().numpy()
  assert _maxrel(out, ref) <= 1e-5


@pytest.mark.parametrize("n", [64, 512, 1024])
def test_one_step_parity(n):
  w = _w0(n, 3)
  o32, o64, eq = _eqs(n)
  dt = 1e-3
  ref = osp.crank_nicolson_rk4(o32, dt)(w)
  ref64 = osp.crank_nicolson_rk4(o64, dt)(w.astype(np.complex128))
  step = _sp().time_stepping.crank_nicolson_rk4(eq, dt)
  out = step(torch.as_tensor(w, device=_dev())).cpu().numpy()
  err = _maxrel(out, ref)
  print(f"spectral one step n={n}: max rel err {err:.2e} (vs fp64 {_maxrel(out, ref64):.2e};"
        f" fp32 oracle vs fp64 {_maxrel(ref, ref64):.2e})")
  assert err <= 1e-5


@pytest.mark.parametrize("n,steps", [(128, 100), (256, 100)])
def test_multi_step_rel_l2(n, steps):
  w = _w0(n, 5, vmax=2.0)
  o32, o64, eq = _eqs(n)
  dt = 1e-3
  ref = osp.rollout(osp.crank_nicolson_rk4(o32, dt), w, steps)
  ref64 = osp.rollout(osp.crank_nicolson_rk4(o64, dt), w.astype(np.complex128), steps)
  from jax_cfd_amd.base import funcutils
  step = _sp().time_stepping.crank_nicolson_rk4(eq, dt)
  out = funcutils.repeated(step, steps)(torch.as_tensor(w, device=_dev())).cpu().numpy()
  err = _rel_l2(out, ref)
  print(f"spectral {steps} steps n={n}: rel L2 {err:.2e} (vs fp64 {_rel_l2(out, ref64):.2e};"
        f" fp32 oracle vs fp64 {_rel_l2(ref, ref64):.2e})")
  assert err <= 1e-3


# Low-storage CN-RK tableaux other than CN-RK4 (spectral/time_stepping.py:117-183):
# Williamson's CN-RK3, a one-stage tableau (explicit Euler with Crank-Nicolson
# for the linear term) and an eight-stage tableau (the largest built).
_TABLEAUX = {
    "rk3": (osp.CNRK3_ALPHAS, osp.CNRK3_BETAS, osp.CNRK3_GAMMAS),
    "one_stage": ((0.0, 1.0), (0.0,), (1.0,)),
    "eight_stage": (tuple(k / 8 for k in range(9)), (0.0,) + (-0.25,) * 7,
                    (0.125,) + (0.15,) * 7),
}


@pytest.mark.parametrize("name,n", [("rk3", 64), ("rk3", 256), ("one_stage", 64),
                                    ("eight_stage", 128)])
def test_low_storage_tableau_parity(name, n):
  alphas, betas, gammas = _TABLEAUX[name]
  w = _w0(n, 3)
  o32, o64, eq = _eqs(n)
  dt = 1e-3
  ref_step = osp.low_storage_rk_crank_nicolson(alphas, betas, gammas, o32, dt)
  ts = _sp().time_stepping
  step = (ts.crank_nicolson_rk3(eq, dt) if name == "rk3" else
          ts.low_storage_runge_kutta_crank_nicolson(alphas, betas, gammas, eq, dt))
  print(f"tableau {name} n={n}", flush=True)
  out = step(torch.as_tensor(w, device=_dev())).cpu().numpy()
  err = _maxrel(out, ref_step(w))
  print(f"  one step: max rel err {err:.2e}")
  assert err <= 1e-5
  from jax_cfd_amd.base import funcutils
  out = funcutils.repeated(step, 50)(torch.as_tensor(w, device=_dev())).cpu().numpy()
  err = _rel_l2(out, osp.rollout(ref_step, w, 50))
  print(f"  50 steps: rel L2 {err:.2e}")
  assert err <= 1e-4
  plan = step.plan()
  assert plan.LAUNCHES_PER_STEP["ns_rows"] == len(betas)


def test_tableau_limits_raise():
  _, _, eq = _eqs(64)
  ts = _sp().time_stepping
  with pytest.raises(NotImplementedError):
    ts.low_storage_runge_kutta_crank_nicolson(
        tuple(range(10)), (0.0,) * 9, (1.0,) * 9, eq, 1e-3)
  with pytest.raises(NotImplementedError):
    ts.crank_nicolson_rk2(eq, 1e-3)


def test_batch_equals_single_instances():
  n, b = 64, 3
  ws = np.stack([_w0(n, s) for s in range(b)])
  _, _, eq = _eqs(n)
  step = _sp().time_stepping.crank_nicolson_rk4(eq, 1e-3)
  batched = step.multistep(torch.as_tensor(ws, device=_dev()), 5)
  for i in range(b):
    single = step.multistep(torch.as_tensor(ws[i], device=_dev()), 5)
    assert torch.equal(batched[i], single)


def test_graph_replay_matches_individual_launches():
  n, b = 64, 2
  ws = torch.as_tensor(np.stack([_w0(n, s) for s in range(b)]), device=_dev())
  _, _, eq = _eqs(n)
  plan = eq.plan(1e-3, b, _dev())
  plan.set_graphs(False)
  ref = plan.step(ws, 70)
  plan.set_graphs(True)
  for _ in range(3):
    out = plan.step(ws, 70)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    plan.explicit_terms(out)  # unrelated eager work between replays


def test_taylor_green_decay():
  """spectral/equations_test.py:263-303 pattern: exp(-2 nu t) decay."""
  n, nu = 64, 1e-2
  og = _ogrid(n)
  x, y = og.mesh((0, 0), np.float64)
  w0 = (-2 * np.cos(x) * np.cos(y))
  eq = _sp().equations.NavierStokes2D(nu, _grid(n), drag=0.0)
  step = _sp().time_stepping.crank_nicolson_rk4(eq, 0.01)
  w = step.multistep(torch.as_tensor(np.fft.rfftn(w0).astype(np.complex64), device=_dev()), 50)
  out = np.fft.irfftn(w.cpu().numpy(), s=(n, n), axes=(0, 1))
  np.testing.assert_allclose(out, w0 * np.exp(-2 * nu * 0.5), atol=1e-5)


def test_reference_taylor_green_accuracy_1024():
  """spectral/equations_test.py:263-303 as written: Taylor-Green at 1024^2,
  viscosity 1e-3, NavierStokes2D without forcing or drag, crank_nicolson_rk4
  with dt = stable_time_step(7, 0.5, 1e-3, grid, implicit_diffusion=True),
  ceil(0.11 / dt) steps from rfftn(curl_2d(v(0))), compared with
  curl_2d(v(0.11)) at atol 1e-3."""
  n, nu, time = 1024, 1e-3, 0.11
  grid = _grid(n)
  import jax_cfd_amd.base as cfd

  def vorticity(t):
    # staggered Taylor-Green velocity (validation_problems.py:80-103), fp32,
    # then curl_2d (finite_differences.py:186-193) at offset (1, 1)
    scale = np.exp(-2 * nu * t)
    (ux, uy), (vx, vy) = [tuple(m.astype(np.float32) for m in grid.mesh(off))
                          for off in grid.cell_faces]
    u = (np.float32(scale) * np.cos(ux) * np.sin(uy)).astype(np.float32)
    v = (-np.float32(scale) * np.sin(vx) * np.cos(vy)).astype(np.float32)
    dx, dy = grid.step
    return ((np.roll(v, -1, 0) - v) / np.float32(dx)
            - (np.roll(u, -1, 1) - u) / np.float32(dy)).astype(np.float32)

  dt = cfd.equations.stable_time_step(7, 0.5, nu, grid, implicit_diffusion=True)
  steps = int(np.ceil(time / dt))
  eq = _sp().equations.NavierStokes2D(nu, grid, drag=0.0)
  step = _sp().time_stepping.crank_nicolson_rk4(eq, dt)
  w0 = np.fft.rfftn(vorticity(0.0)).astype(np.complex64)
  print(f"taylor-green spectral 1024^2: dt {dt:.6g}, {steps} steps", flush=True)
  w = step.multistep(torch.as_tensor(w0, device=_dev()), steps)
  out = np.fft.irfftn(w.cpu().numpy(), s=(n, n), axes=(0, 1))
  want = vorticity(time)
  err = float(np.abs(out - want).max())
  print(f"  max |error| {err:.3e} (atol 1e-3)")
  np.testing.assert_allclose(out, want, atol=1e-3)


@pytest.mark.parametrize("n,b", [(64, 1), (256, 2), (1024, 1)])
def test_rfft2_irfft2(n, b):
  from jax_cfd_amd.spectral import _ns2d
  plan = _ns2d.Ns2dPlan(_sp().equations.NavierStokes2D(1e-3, _grid(n)).spec(0.0), b, _dev())
  rng = np.random.default_rng(n)
  x = rng.standard_normal((b, n, n) if b > 1 else (n, n)).astype(np.float32)
  xh = plan.rfft2(torch.as_tensor(x, device=_dev())).cpu().numpy()
  ref = np.fft.rfftn(x.astype(np.float64), axes=(-2, -1))
  assert _maxrel(xh, ref) <= 1e-5
  back = plan.irfft2(torch.as_tensor(xh, device=_dev())).cpu().numpy()
  assert _maxrel(back, x) <= 1e-5
