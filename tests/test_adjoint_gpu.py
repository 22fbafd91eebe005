"""Hand-written FVM adjoint on the GPU vs torch.autograd on the oracle's
torch restatement (the jax.grad stand-in; same differentiation rules) and vs
central finite differences.  Tolerance vs the autograd reference: 1e-4
relative (L2 over the whole gradient)."""
import numpy as np
import pytest
import torch

from oracle import fvm as ofvm
from oracle import grid_ops as go
from oracle import initial_conditions as oic
from oracle import torch_fvm as tf
from oracle.grid_ops import Grid as OGrid

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def _dev():
  return torch.device("cuda", 0)


def _setup(n, seed=1, vmax=7.0):
  import jax_cfd_amd.base as cfd
  ogrid = OGrid.from_domain((n, n), ((0, TWO_PI), (0, TWO_PI)))
  grid = cfd.grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))
  dt = ofvm.stable_time_step(vmax, 0.5, 1e-3, ogrid)
  step = cfd.equations.semi_implicit_navier_stokes(
      1.0, 1e-3, dt, grid,
      convect=lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v),
      forcing=cfd.forcings.simple_turbulence_forcing(grid, 1.0, 4, -0.1, 'kolmogorov'))
  fx = go.Forcing(1.0, 4, -0.1).constant_terms(ogrid, np.float64)[0][0]
  tstep = tf.TorchFVMStep(ogrid, dt, 1e-3, 1.0, -0.1, fx)
  v0 = oic.filtered_velocity_field(seed, ogrid, vmax, 4)
  return grid, ogrid, step, tstep, v0


def _rel(a, b):
  return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [32, 64, 256])
def test_step_vjp_matches_autograd(n):
  grid, ogrid, step, tstep, v0 = _setup(n)
  rng = np.random.default_rng(n)
  lu, lv = (rng.standard_normal((n, n)).astype(np.float32) for _ in range(2))
  plan = step.plan(_dev())
  d = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=_dev())  # noqa: E731
  gu, gv = plan.step_vjp(d(v0[0].data), d(v0[1].data), d(lu), d(lv))
  u = torch.tensor(v0[0].data.astype(np.float64), requires_grad=True)
  v = torch.tensor(v0[1].data.astype(np.float64), requires_grad=True)
  uo, vo = tstep(u, v)
  ru, rv = torch.autograd.grad((uo * torch.tensor(lu, dtype=torch.float64)).sum()
                               + (vo * torch.tensor(lv, dtype=torch.float64)).sum(), (u, v))
  eu, ev = _rel(gu.cpu().numpy(), ru.numpy()), _rel(gv.cpu().numpy(), rv.numpy())
  print(f"step vjp n={n}: rel err u {eu:.2e} v {ev:.2e}")
  assert eu <= 1e-4 and ev <= 1e-4


def _loss_setup(n, steps, every):
  from jax_cfd_amd import adjoint
  grid, ogrid, step, tstep, v0 = _setup(n, seed=1)
  # target: curl of the state reached from the seed-2 field
  vt = oic.filtered_velocity_field(2, ogrid, 7.0, 4)
  d = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=_dev())  # noqa: E731
  plan = step.plan(_dev())
  ut, vt_ = plan.step(d(vt[0].data), d(vt[1].data), steps)
  wt = adjoint.curl_2d(ut, vt_, grid).detach()
  roll = adjoint.rollout(step, steps, every)

  def loss(u0, v0_):
    u, v = roll(u0, v0_)
    return ((adjoint.curl_2d(u, v, grid) - wt) ** 2).mean()
  return grid, ogrid, tstep, v0, d, loss, wt


def test_rollout_gradient_matches_autograd_and_checkpointing():
  n, steps = 64, 24
  grid, ogrid, tstep, v0, d, loss, wt = _loss_setup(n, steps, 8)
  u0 = d(v0[0].data).requires_grad_(True)
  w0 = d(v0[1].data).requires_grad_(True)
  gu, gv = torch.autograd.grad(loss(u0, w0), (u0, w0))
  # autograd through the fp64 torch restatement
  u = torch.tensor(v0[0].data.astype(np.float64), requires_grad=True)
  v = torch.tensor(v0[1].data.astype(np.float64), requires_grad=True)
  uu, vv = u, v
  for _ in range(steps):
    uu, vv = tstep(uu, vv)
  wt64 = torch.tensor(wt.cpu().numpy().astype(np.float64))
  L = ((tf.curl_2d(uu, vv, ogrid) - wt64) ** 2).mean()
  ru, rv = torch.autograd.grad(L, (u, v))
  eu, ev = _rel(gu.cpu().numpy(), ru.numpy()), _rel(gv.cpu().numpy(), rv.numpy())
  print(f"rollout grad n={n} steps={steps}: rel err u {eu:.2e} v {ev:.2e}")
  assert eu <= 1e-4 and ev <= 1e-4
  # checkpoint spacing changes nothing (recomputation is deterministic)
  for every in (1, 5, 64):
    _, _, _, _, _, loss_k, _ = _loss_setup(n, steps, every)
    u1 = d(v0[0].data).requires_grad_(True)
    w1 = d(v0[1].data).requires_grad_(True)
    hu, hv = torch.autograd.grad(loss_k(u1, w1), (u1, w1))
    assert torch.equal(hu, gu) and torch.equal(hv, gv)


def test_gradient_c4_config():
  """256^2, 64 steps, checkpoints every 8, MSE of vorticity against the
  seed-2 trajectory's: (a) against torch.autograd of the fp64 restatement
  (<= 1e-4 relative); (b) 16 random directions of central differences with
  eps 1e-3 on the fp32 GPU forward.  The fp32 differences carry cancellation
  noise of ~1e-7 |L| / eps, so (b) is judged on the 16 directional
  derivatives together (relative L2), within 2e-2."""
  n, steps = 256, 64
  grid, ogrid, tstep, v0, d, loss, wt = _loss_setup(n, steps, 8)
  u0 = d(v0[0].data).requires_grad_(True)
  w0 = d(v0[1].data).requires_grad_(True)
  gu, gv = torch.autograd.grad(loss(u0, w0), (u0, w0))
  # (a) jax.grad stand-in
  u = torch.tensor(v0[0].data.astype(np.float64), requires_grad=True)
  v = torch.tensor(v0[1].data.astype(np.float64), requires_grad=True)
  uu, vv = u, v
  for _ in range(steps):
    uu, vv = tstep(uu, vv)
  wt64 = torch.tensor(wt.cpu().numpy().astype(np.float64))
  ru, rv = torch.autograd.grad(((tf.curl_2d(uu, vv, ogrid) - wt64) ** 2).mean(), (u, v))
  eu, ev = _rel(gu.cpu().numpy(), ru.numpy()), _rel(gv.cpu().numpy(), rv.numpy())
  print(f"256^2 64-step gradient vs fp64 autograd: rel err u {eu:.2e} v {ev:.2e}")
  assert eu <= 1e-4 and ev <= 1e-4
  # (b) central differences
  gen = torch.Generator(device=_dev()).manual_seed(0)
  fds, ads = [], []
  with torch.no_grad():
    for _ in range(16):
      du = torch.randn((n, n), device=_dev(), generator=gen)
      dv = torch.randn((n, n), device=_dev(), generator=gen)
      eps = 1e-3
      fds.append(float((loss(u0 + eps * du, w0 + eps * dv)
                        - loss(u0 - eps * du, w0 - eps * dv)) / (2 * eps)))
      ads.append(float((gu * du).sum() + (gv * dv).sum()))
  fds, ads = np.array(fds), np.array(ads)
  agg = float(np.linalg.norm(fds - ads) / np.linalg.norm(ads))
  print(f"256^2 64-step gradient, central differences: relative L2 over 16 directions {agg:.2e}; "
        f"per-direction median {np.median(np.abs(fds - ads) / np.abs(ads)):.2e}")
  assert agg <= 2e-2


def _kink_fields(n, kind):
  """Fields in exact arithmetic (small integers and halves: every face
  average and difference is exact in fp32 and fp64, so both make the same
  branch choices) that put faces on every kink of advect_van_leer: limiter
  denominators and numerators exactly 0 (flat runs), face velocities exactly
  0 with nonzero limiter values (+1/-1 alternations next to ramps), and
  (kind 'bands') whole bands of zero velocity."""
  pa = np.array([0, 0, 0, 1, 1, 2, 3, 4, 4, 4, -1, 1, -1, 1, 0, 0,
                 2, 2, 2, 1, 0, -1, -2, -2, .5, .5, 0, 0, 3, 3, 3, 0], np.float64)
  pb = np.array([1, 1, 1, 1, 2, 3, 4, -1, 1, -1, 1, -1, 0, 0, 0, .5,
                 1, 1.5, 2, 2, 2, 0, 0, -3, -2, -1, 0, 1, 1, 1, 0, 0], np.float64)
  reps = n // pa.size
  a, b = np.tile(pa, reps), np.tile(pb, reps)
  u = a[:, None] + b[None, :]
  v = b[:, None] - a[None, ::-1]
  if kind == 'bands':
    u[n // 4: n // 2, :] = 0.0
    v[:, n // 2: 3 * n // 4] = 0.0
  return u.astype(np.float32), v.astype(np.float32)


def _kink_counts(u, v):
  """Faces of the x-direction flux of u at each kink (the other three flux
  families have the same structure): U = 0, den = 0, num = 0, and U = 0 with
  num and den of the same sign (a nonzero limiter, where d|U|/dU matters)."""
  um1 = np.roll(u, 1, 0)
  U = 0.5 * um1 + 0.5 * u
  den = u - um1
  num = np.where(U > 0, um1 - np.roll(u, 2, 0), np.roll(u, -1, 0) - u)
  live = (U == 0) & (num * den > 0)
  return dict(U0=int((U == 0).sum()), den0=int((den == 0).sum()),
              num0=int((num == 0).sum()), U0_phi=int(live.sum()))


@pytest.mark.parametrize("kind", ["stripes", "bands"])
def test_step_vjp_at_limiter_kinks(kind):
  """jcfd_fvm_step_vjp where the Van Leer limiter is not differentiable:
  the GPU adjoint must apply jax's rules at the kinks (where routes the
  cotangent to the taken branch, sign' = 0, the guarded quotient at den = 0,
  d|x|/dx = +1 at 0) and match torch.autograd through the fp64 stand-in,
  which uses the same rules, to <= 1e-4.  The same stand-in with
  d|x|/dx = 0 at 0 (torch.abs) must differ from it by far more than that,
  so the comparison decides the rule at U = 0."""
  n = 64
  grid, ogrid, step, tstep, _ = _setup(n)
  u0, v0 = _kink_fields(n, kind)
  counts = _kink_counts(u0, v0)
  print(f"kinks ({kind}): {counts}", flush=True)
  assert all(c > 0 for c in counts.values())
  rng = np.random.default_rng(7)
  lu, lv = (rng.standard_normal((n, n)).astype(np.float32) for _ in range(2))
  plan = step.plan(_dev())
  d = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=_dev())  # noqa: E731
  gu, gv = plan.step_vjp(d(u0), d(v0), d(lu), d(lv))
  gu, gv = gu.cpu().numpy(), gv.cpu().numpy()

  def stand_in():
    u = torch.tensor(u0.astype(np.float64), requires_grad=True)
    v = torch.tensor(v0.astype(np.float64), requires_grad=True)
    uo, vo = tstep(u, v)
    return torch.autograd.grad((uo * torch.tensor(lu, dtype=torch.float64)).sum()
                               + (vo * torch.tensor(lv, dtype=torch.float64)).sum(), (u, v))
  ru, rv = (x.numpy() for x in stand_in())
  assert np.isfinite(gu).all() and np.isfinite(gv).all()
  eu, ev = _rel(gu, ru), _rel(gv, rv)
  jabs = tf.jabs
  try:
    tf.jabs = torch.abs  # the other rule at 0
    au, av = (x.numpy() for x in stand_in())
  finally:
    tf.jabs = jabs
  alt = max(_rel(au, ru), _rel(av, rv))
  print(f"step vjp at kinks ({kind}): rel err u {eu:.2e} v {ev:.2e}; "
        f"stand-in with d|x|/dx(0) = 0 differs by {alt:.2e}")
  assert eu <= 1e-4 and ev <= 1e-4
  assert alt >= 1e-3
