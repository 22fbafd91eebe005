"""Mixed-size group stepping (jcfd_fvm_group_step): one launch per kernel
class (one workgroup per work item) over instances of every grid size.  Each instance must equal,
bit for bit, the plain single-instance step of its own plan; slots not
listed must stay untouched; graph replay must equal eager launches; member
sets may change between calls."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = (32, 64, 128, 256, 512)


def _dev():
  return torch.device('cuda', 0)


def _family(n):
  from jax_cfd_amd import serving
  return serving.FvmFamily(n, _dev())


def _fields(fam, batch, seed0):
  u = torch.empty((batch, fam.n, fam.n), dtype=torch.float32, device=_dev())
  v = torch.empty_like(u)
  for s in range(batch):
    u0, v0 = fam.initial_state(seed0 + s)
    u[s].copy_(u0)
    v[s].copy_(v0)
  return u, v


def _plain(fam, u0, v0, nsteps):
  import jax_cfd_amd.base as cfd
  plan = cfd._fvm.FvmPlan(fam.spec, _dev())
  return plan.step(u0.clone(), v0.clone(), nsteps)


def _setup(batch=3):
  import jax_cfd_amd.base as cfd
  members = []
  for k, n in enumerate(SIZES):
    fam = _family(n)
    plan = cfd._fvm.FvmPlan(fam.spec, _dev(), batch=batch)
    u, v = _fields(fam, batch, 100 * k)
    members.append((fam, plan, u, v))
  return members


@pytest.mark.parametrize("graphs", [True, False])
def test_group_equals_plain_per_instance(graphs):
  from jax_cfd_amd.base import _fvm
  members = _setup()
  init = [(u.clone(), v.clone()) for (_, _, u, v) in members]
  group = _fvm.FvmGroup(_dev())
  group.set_graphs(graphs)
  # slot 1 of every plan is left out; 40 steps = one 32-step graph + 8 eager
  slots = [0, 2]
  nsteps = 40
  print(f"group step: sizes {SIZES}, slots {slots}, {nsteps} steps, graphs={graphs}", flush=True)
  group.step([(plan, u, v, slots) for (_, plan, u, v) in members], nsteps)
  torch.cuda.synchronize()
  for (fam, _, u, v), (u0, v0) in zip(members, init):
    for s in slots:
      uo, vo = _plain(fam, u0[s], v0[s], nsteps)
      assert torch.equal(u[s], uo) and torch.equal(v[s], vo), (fam.n, s)
    assert torch.equal(u[1], u0[1]) and torch.equal(v[1], v0[1]), fam.n


def test_group_member_set_changes_and_graph_equals_eager():
  from jax_cfd_amd.base import _fvm
  runs = []
  for graphs in (True, False):
    members = _setup()
    group = _fvm.FvmGroup(_dev())
    group.set_graphs(graphs)
    # call 1: every slot of every size; call 2: a smaller, different set
    # (re-uploaded work list, smaller largest grid); call 3: back to all
    group.step([(plan, u, v, [0, 1, 2]) for (_, plan, u, v) in members], 33)
    group.step([(plan, u, v, [1]) for (fam, plan, u, v) in members if fam.n <= 128], 35)
    group.step([(plan, u, v, [2, 0]) for (_, plan, u, v) in members], 7)
    torch.cuda.synchronize()
    runs.append([(u.clone(), v.clone()) for (_, _, u, v) in members])
  for (ug, vg), (ue, ve) in zip(*runs):
    assert torch.equal(ug, ue) and torch.equal(vg, ve)
  # against the plain path: slot 1 of the small sizes ran 33 + 35 steps,
  # every other slot 33 + 7
  members = _setup()
  for (fam, _, u0, v0), (ug, vg) in zip(members, runs[0]):
    for s in range(3):
      nsteps = 33 + (35 if (s == 1 and fam.n <= 128) else 0) + (7 if s != 1 else 0)
      uo, vo = _plain(fam, u0[s], v0[s], nsteps)
      assert torch.equal(ug[s], uo) and torch.equal(vg[s], vo), (fam.n, s, nsteps)


def test_group_rejects_bad_members():
  import jax_cfd_amd.base as cfd
  from jax_cfd_amd.base import _fvm
  fam = _family(64)
  plan = cfd._fvm.FvmPlan(fam.spec, _dev(), batch=2)
  u, v = _fields(fam, 2, 0)
  group = _fvm.FvmGroup(_dev())
  with pytest.raises(ValueError):
    group.step([(plan, u, v, [0, 0])], 1)             # repeated slot
  with pytest.raises(ValueError):
    group.step([(plan, u, v, [2])], 1)                # out of range
  with pytest.raises(ValueError):
    group.step([(plan, u, v, [0]), (plan, u, v, [0])], 1)  # same instance twice
  with pytest.raises(ValueError):
    group.step([(plan, u[:, :32], v, [0])], 1)        # wrong shape
  big = _family(1024)
  plan_big = cfd._fvm.FvmPlan(big.spec, _dev(), batch=1)
  ub = torch.zeros((1, 1024, 1024), device=_dev())
  with pytest.raises(NotImplementedError):
    group.step([(plan_big, ub, ub.clone(), [0])], 1)  # grid not built for groups
  # implicit-diffusion plans (diffusion.solve_cg, diffusion.solve_fast_diag)
  # would be stepped without their diffusion solve: rejected, not run
  import dataclasses
  from jax_cfd_amd import _lib
  for mode in (_lib.DIFFUSION_IMPLICIT_CG, _lib.DIFFUSION_IMPLICIT_FAST_DIAG):
    spec = dataclasses.replace(fam.spec, diffusion=mode, viscosity=1e-3)
    plan_impl = cfd._fvm.FvmPlan(spec, _dev(), batch=1)
    ui, vi = _fields(fam, 1, 7)
    with pytest.raises(NotImplementedError):
      group.step([(plan_impl, ui, vi, [0])], 1)
