"""implicit_diffusion_navier_stokes on the GPU (equations.py:154-195 with
diffusion.solve_fast_diag, diffusion.py:170-212, or diffusion.solve_cg,
diffusion.py:137-167) against the oracle's restatement: the diffusion solve
alone, one step (max relative error <= 1e-5), 100 steps (relative L2 <= 1e-3,
with the fp32 oracle's own distance to fp64 reported beside), batched ==
single instance and graph replay == eager launches (bitwise).  The CG solve
stops on its own fp32 residual, so its iteration count may differ from the
oracle's by one; its tolerances are written in each test."""
import functools

import numpy as np
import pytest
import torch

from oracle import fvm as ofvm
from oracle import grid_ops as go
from oracle import initial_conditions as oic
from oracle.grid_ops import Grid as OGrid

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def _dev():
  return torch.device('cuda', 0)


def _cfd():
  import jax_cfd_amd.base as cfd
  return cfd


def _grid(n):
  return _cfd().grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))


def _ogrid(n):
  return OGrid.from_domain((n, n), ((0, TWO_PI), (0, TWO_PI)))


def _state(v, n):
  cfd = _cfd()
  g = _grid(n)
  bc = cfd.boundaries.periodic_boundary_conditions(2)
  return tuple(cfd.grids.GridVariable(cfd.grids.GridArray(
      torch.as_tensor(np.ascontiguousarray(u.data, np.float32), device=_dev()), o, g), bc)
      for u, o in zip(v, g.cell_faces))


def _maxrel(a, b):
  return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _rel_l2(a, b):
  return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_solve_fast_diag_constants_and_oracle():
  """diffusion_test.py:58-76 (constants unchanged, atol 1e-6) and a random
  field against the oracle's solve."""
  cfd = _cfd()
  n = 128
  g = _grid(n)
  bc = cfd.boundaries.periodic_boundary_conditions(2)
  ones = tuple(cfd.grids.GridVariable(cfd.grids.GridArray(
      torch.ones((n, n), device=_dev()), o, g), bc) for o in g.cell_faces)
  out = cfd.diffusion.solve_fast_diag(ones, 1.0, 0.1)
  for o in out:
    assert float(torch.max(torch.abs(o.data - 1))) <= 1e-6
  rng = np.random.default_rng(3)
  raw = [rng.standard_normal((n, n)).astype(np.float32) for _ in range(2)]
  og = _ogrid(n)
  ref = ofvm.solve_fast_diag_diffusion(ofvm.velocity(*raw, og), 0.05, 0.02, og)
  got = cfd.diffusion.solve_fast_diag(_state(ofvm.velocity(*raw, og), n), 0.05, 0.02)
  for a, b in zip(got, ref):
    err = _maxrel(a.data.cpu().numpy(), b.data)
    print(f"diffusion solve n={n}: max rel err {err:.2e}")
    assert err <= 1e-6


def test_solve_cg_constants_oracle_and_fast_diag():
  """diffusion_test.py:58-76 with solve_cg (constants: no iteration, x = b
  exactly), then a random field against the oracle's CG solve (max rel
  <= 2e-5) and the GPU's fast-diagonal solve of the same system."""
  cfd = _cfd()
  from jax_cfd_amd import _lib
  from jax_cfd_amd.base import _fvm
  n = 128
  g = _grid(n)
  bc = cfd.boundaries.periodic_boundary_conditions(2)
  ones = tuple(cfd.grids.GridVariable(cfd.grids.GridArray(
      torch.ones((n, n), device=_dev()), o, g), bc) for o in g.cell_faces)
  out = cfd.diffusion.solve_cg(ones, 1.0, 0.1, atol=1e-6)
  for o in out:
    assert torch.equal(o.data, torch.ones_like(o.data))
  rng = np.random.default_rng(4)
  raw = [rng.standard_normal((n, n)).astype(np.float32) for _ in range(2)]
  og = _ogrid(n)
  ref, iters = ofvm.solve_cg_diffusion(ofvm.velocity(*raw, og), 0.05, 0.02, og,
                                       rtol=1e-7, atol=1e-6, return_iterations=True)
  state = _state(ofvm.velocity(*raw, og), n)
  got = cfd.diffusion.solve_cg(state, 0.05, 0.02, rtol=1e-7, atol=1e-6)
  plan = _fvm.plan_for(_fvm.spec_from_grid(
      g, dt=0.02, viscosity=0.05, diffusion=_lib.DIFFUSION_IMPLICIT_CG,
      **cfd.diffusion.cg_spec_fields(rtol=1e-7, atol=1e-6)), _dev())
  fast = cfd.diffusion.solve_fast_diag(state, 0.05, 0.02)
  for m, (a, b, f) in enumerate(zip(got, ref, fast)):
    err = _maxrel(a.data.cpu().numpy(), b.data)
    dfast = float(torch.max(torch.abs(a.data - f.data)))
    k = plan.diffusion_cg_iterations(m)
    print(f"CG diffusion n={n} component {m}: {k} iterations (oracle {iters[m]}), "
          f"max rel err {err:.2e}, max |CG - fast diag| {dfast:.2e}")
    assert abs(k - iters[m]) <= 1 and k > 0
    assert err <= 2e-5
    assert dfast <= 1e-4


def _setup(n, convect, pressure='fast_diag', diffusion='fast_diag'):
  cfd = _cfd()
  og = _ogrid(n)
  v0 = oic.filtered_velocity_field(5, og, 7.0, 4)
  dt = ofvm.stable_time_step(7.0, 0.5, 1e-3, og, implicit_diffusion=True)
  forcing = go.Forcing(scale=1.0, wavenumber=4, linear_coefficient=-0.1)
  ostep = ofvm.ImplicitDiffusionStep(og, dt, viscosity=1e-3, forcing=forcing,
                                     convect=convect, pressure=pressure,
                                     diffusion=diffusion)
  g = _grid(n)
  conv = None if convect == 'van_leer_limiters' else (
      lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v))
  solve = (cfd.pressure.solve_fast_diag if pressure == 'fast_diag'
           else functools.partial(cfd.pressure.solve_cg, rtol=1e-6, atol=1e-6))
  dsolve = (cfd.diffusion.solve_fast_diag if diffusion == 'fast_diag'
            else cfd.diffusion.solve_cg)
  step = cfd.equations.implicit_diffusion_navier_stokes(
      1.0, 1e-3, dt, g, convect=conv, pressure_solve=solve, diffusion_solve=dsolve,
      forcing=cfd.forcings.simple_turbulence_forcing(g, 1.0, 4, -0.1, 'kolmogorov'))
  return og, v0, ostep, step


@pytest.mark.parametrize("n,convect", [(64, 'van_leer_limiters'), (256, 'van_leer'),
                                       (256, 'van_leer_limiters')])
def test_one_and_100_steps(n, convect):
  og, v0, ostep, step = _setup(n, convect)
  out = step(_state(v0, n))
  ref = ostep(v0)
  for o, r in zip(out, ref):
    err = _maxrel(o.data.cpu().numpy(), r.data)
    print(f"implicit diffusion one step n={n} {convect}: max rel err {err:.2e}")
    assert err <= 1e-5
  out = _cfd().funcutils.repeated(step, 100)(_state(v0, n))
  ref32 = ostep.rollout(v0, 100)
  v64 = tuple(go.A(u.data.astype(np.float64), u.offset) for u in v0)
  ref64 = ostep.rollout(v64, 100)
  for o, r, r64 in zip(out, ref32, ref64):
    err = _rel_l2(o.data.cpu().numpy(), r.data)
    e64 = _rel_l2(r.data, r64.data)
    print(f"implicit diffusion 100 steps n={n}: rel L2 {err:.2e} (fp32 oracle vs fp64 {e64:.2e})")
    assert err <= 1e-3


def test_cg_pressure_variant_one_step():
  og, v0, ostep, step = _setup(64, 'van_leer_limiters', pressure='cg')
  out = step(_state(v0, 64))
  ref = ostep(v0)
  for o, r in zip(out, ref):
    err = _maxrel(o.data.cpu().numpy(), r.data)
    print(f"implicit diffusion + CG one step: max rel err {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("n,pressure", [(64, 'fast_diag'), (64, 'cg'), (256, 'fast_diag')])
def test_cg_diffusion_steps(n, pressure):
  """diffusion_solve=solve_cg (default tolerances): one step max rel <= 1e-5,
  20 steps rel L2 <= 1e-4 against the oracle's CG-diffusion step."""
  og, v0, ostep, step = _setup(n, 'van_leer_limiters', pressure=pressure, diffusion='cg')
  out = step(_state(v0, n))
  ref = ostep(v0)
  for o, r in zip(out, ref):
    err = _maxrel(o.data.cpu().numpy(), r.data)
    print(f"CG diffusion one step n={n} pressure {pressure}: max rel err {err:.2e}")
    assert err <= 1e-5
  out = _cfd().funcutils.repeated(step, 20)(_state(v0, n))
  ref = ostep.rollout(v0, 20)
  for o, r in zip(out, ref):
    err = _rel_l2(o.data.cpu().numpy(), r.data)
    print(f"CG diffusion 20 steps n={n} pressure {pressure}: rel L2 {err:.2e}")
    assert err <= 1e-4


def test_cg_diffusion_rejects_batches():
  from jax_cfd_amd.base import _fvm
  _, _, _, step = _setup(64, 'van_leer_limiters', diffusion='cg')
  with pytest.raises(NotImplementedError):
    _fvm.FvmPlan(step.spec, _dev(), batch=2)


def test_batched_and_graph_replay_bitwise():
  from jax_cfd_amd.base import _fvm
  n = 128
  og, _, _, step = _setup(n, 'van_leer_limiters')
  fields = [oic.filtered_velocity_field(s, og, 7.0, 4) for s in range(3)]
  u = torch.stack([torch.as_tensor(f[0].data, device=_dev()) for f in fields])
  v = torch.stack([torch.as_tensor(f[1].data, device=_dev()) for f in fields])
  plan = _fvm.FvmPlan(step.spec, _dev(), batch=3)
  plan.set_graphs(True)
  ug, vg = plan.step(u, v, 40)
  plan.set_graphs(False)
  ue, ve = plan.step(u, v, 40)
  assert torch.equal(ug, ue) and torch.equal(vg, ve)
  single = step.plan(_dev())
  for s in range(3):
    us, vs = single.step(u[s], v[s], 40)
    assert torch.equal(us, ug[s]) and torch.equal(vs, vg[s])
