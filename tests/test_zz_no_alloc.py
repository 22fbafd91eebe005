"""jcfd_fvm_step_noalloc against jcfd_fvm_step, through the C ABI: the same
kernels with the step's scratch in the library's static per-GPU array, so
results must be identical bit for bit.  Two calls in a row, the second on a
smaller grid (the static array keeps the first call's contents), then a
larger one again; an implicit-diffusion plan; nsteps = 0; and the two
rejections (scratch above the static maximum, a plan kind the entry point
does not run)."""
import dataclasses

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
  return torch.device('cuda', 0)


def _fam(n):
  from jax_cfd_amd import serving
  return serving.FvmFamily(n, _dev())


def _batch(fam, batch, seed0):
  u = torch.stack([fam.initial_state(seed0 + s)[0] for s in range(batch)])
  v = torch.stack([fam.initial_state(seed0 + s)[1] for s in range(batch)])
  return (u[0], v[0]) if batch == 1 else (u, v)


def _check(plan, u, v, nsteps, label):
  print(f"no-alloc step: {label}, n={plan.spec.n}, batch={plan.batch}, {nsteps} steps",
        flush=True)
  want = plan.step(u.clone(), v.clone(), nsteps)
  got = plan.step_noalloc(u.clone(), v.clone(), nsteps)
  torch.cuda.synchronize()
  assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1]), label


def test_noalloc_equals_step_two_calls_then_larger_again():
  from jax_cfd_amd.base import _fvm
  big = _fam(256)
  plan_big = _fvm.FvmPlan(big.spec, _dev(), batch=2)
  _check(plan_big, *_batch(big, 2, 0), 5, "first call")
  small = _fam(64)
  plan_small = _fvm.FvmPlan(small.spec, _dev(), batch=1)
  _check(plan_small, *_batch(small, 1, 9), 3, "second call, smaller input")
  _check(plan_big, *_batch(big, 2, 4), 2, "third call, larger again")


def test_noalloc_implicit_diffusion_and_zero_steps():
  from jax_cfd_amd import _lib
  from jax_cfd_amd.base import _fvm
  fam = _fam(128)
  spec = dataclasses.replace(fam.spec, diffusion=_lib.DIFFUSION_IMPLICIT_FAST_DIAG)
  plan = _fvm.FvmPlan(spec, _dev(), batch=1)
  u, v = _batch(fam, 1, 3)
  _check(plan, u, v, 4, "implicit fast-diagonal diffusion")
  uo, vo = plan.step_noalloc(u, v, 0)
  assert torch.equal(uo, u) and torch.equal(vo, v)


def test_noalloc_rejects():
  from jax_cfd_amd import _lib
  from jax_cfd_amd.base import _fvm
  fam = _fam(1024)
  plan = _fvm.FvmPlan(fam.spec, _dev(), batch=2)  # 12 * 2 * 1024^2 bytes > 16 MiB
  z = torch.zeros((2, 1024, 1024), device=_dev())
  with pytest.raises(ValueError, match="static maximum"):
    plan.step_noalloc(z, z.clone(), 1)
  small = _fam(64)
  cg = _fvm.FvmPlan(dataclasses.replace(small.spec, pressure=_lib.PRESSURE_CG), _dev())
  u, v = _batch(small, 1, 0)
  with pytest.raises(NotImplementedError):
    cg.step_noalloc(u, v, 1)
