"""advect_step_semilagrangian on the GPU (advection.py:345-382) against the
oracle's restatement: seeded fields at the four staggered offsets, with
departure points up to several cells away, offsets whose velocity
interpolation is a nonzero whole shift, and a domain with dx != dy (max
relative error <= 1e-6; the kernel keeps the oracle's operation order, so
most values are identical),
and the reference's Gaussian translation case in 2-D (100 steps at Courant
0.5 move the profile 50 cells, atol 7e-2; advection_test.py:140-170)."""
import numpy as np
import pytest
import torch

from oracle import grid_ops as go
from oracle.grid_ops import A, Grid as OGrid

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def _dev():
  return torch.device('cuda', 0)


def _cfd():
  import jax_cfd_amd.base as cfd
  return cfd


def _var(x, offset, grid):
  cfd = _cfd()
  return cfd.grids.GridVariable(
      cfd.grids.GridArray(torch.as_tensor(np.ascontiguousarray(x), device=_dev()), offset, grid),
      cfd.boundaries.periodic_boundary_conditions(2))


# The staggered offsets, plus offsets whose difference to u's (1, 0.5) or
# v's (0.5, 1) offset is a nonzero whole number on some axis, so the
# velocity interpolation takes its pure-shift branch with a nonzero shift
# (interpolation.py:52-56): (0, 0.5) and (0, 0) shift u by -1 along x,
# (1.5, 0.5) shifts v by +1 along x, (0.5, 2) shifts v by +1 along y.
OFFSETS = [(0.5, 0.5), (1.0, 0.5), (0.5, 1.0), (1.0, 1.0),
           (0.0, 0.5), (0.0, 0.0), (1.5, 0.5), (0.5, 2.0)]
# square cells, and a domain of lengths (2 pi, pi) so dx != dy
DOMAINS = {'square': ((0, TWO_PI), (0, TWO_PI)), 'rect': ((0, TWO_PI), (0, np.pi))}


@pytest.mark.parametrize("n,domain", [(64, 'square'), (256, 'square'), (64, 'rect')])
@pytest.mark.parametrize("offset", OFFSETS)
def test_semilagrangian_parity(n, domain, offset):
  cfd = _cfd()
  dom = DOMAINS[domain]
  grid = cfd.grids.Grid((n, n), domain=dom)
  og = OGrid.from_domain((n, n), dom)
  rng = np.random.default_rng(n + int(4 * offset[0] + 2 * offset[1]))
  c = rng.standard_normal((n, n)).astype(np.float32)
  u = (3 * rng.standard_normal((n, n))).astype(np.float32)
  v = (3 * rng.standard_normal((n, n))).astype(np.float32)
  dt = 0.7 * og.step[0]  # departure points up to ~10 cells away
  ref = go.advect_step_semilagrangian(
      A(c, offset), (A(u, og.cell_faces[0]), A(v, og.cell_faces[1])), dt, og).data
  print(f"semilagrangian n={n} domain={dom} offset={offset}", flush=True)
  out = cfd.advection.advect_step_semilagrangian(
      _var(c, offset, grid), (_var(u, (1.0, 0.5), grid), _var(v, (0.5, 1.0), grid)), dt)
  assert out.offset == offset
  got = out.data.cpu().numpy()
  err = float(np.abs(got - ref).max() / np.abs(ref).max())
  same = float(np.mean(got == ref))
  print(f"  max rel err {err:.2e}, identical fraction {same:.4f}")
  assert err <= 1e-6


def test_semilagrangian_translates_gaussian():
  cfd = _cfd()
  n = 128
  grid = cfd.grids.Grid((n, n), domain=((0, 1), (0, 1)))
  x = (np.arange(n) - (n // 2 + 1)) / n  # advection_test.py:30-34 profile
  prof = np.exp(-30. * x ** 2).astype(np.float32)
  c0 = np.broadcast_to(prof[:, None], (n, n)).copy()
  vel = (_var(np.ones((n, n), np.float32), (1.0, 0.5), grid),
         _var(np.zeros((n, n), np.float32), (0.5, 1.0), grid))
  c = _var(c0, (0.5, 0.5), grid)
  dt = 0.5 / n
  for _ in range(100):
    c = cfd.advection.advect_step_semilagrangian(c, vel, dt)
  err = np.abs(c.data.cpu().numpy() - np.roll(c0, 50, axis=0)).max()
  print(f"semilagrangian Gaussian 100 steps: max |error| {err:.2e} (atol 7e-2)")
  assert err <= 7e-2


def test_semilagrangian_rejects():
  cfd = _cfd()
  n = 32
  grid = cfd.grids.Grid((n, n), domain=((1, 2), (0, 1)))
  z = np.zeros((n, n), np.float32)
  with pytest.raises(ValueError):
    cfd.advection.advect_step_semilagrangian(
        _var(z, (0.5, 0.5), grid), (_var(z, (1.0, 0.5), grid), _var(z, (0.5, 1.0), grid)), 0.1)
  from jax_cfd_amd.base import _fvm
  grid = cfd.grids.Grid((n, n), domain=((0, 1), (0, 1)))
  plan = _fvm.plan_for(_fvm.spec_from_grid(grid), _dev())
  t = torch.zeros((n, n), device=_dev())
  with pytest.raises(ValueError):
    from jax_cfd_amd import _lib
    with _lib.dlpack_args() as dl:
      rc = plan._lib.jcfd_fvm_advect_semilagrangian(
          plan._handle, dl(t), 0.5, 0.5, dl(t), dl(t), 0.1, dl(t), _lib.stream_handle(_dev()))
    _lib.check(rc, 'jcfd_fvm_advect_semilagrangian')
