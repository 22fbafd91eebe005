"""Mixed-size spectral group stepping (jcfd_ns2d_group_step): one launch
per kernel class and tableau stage (one workgroup per work item) over instances of every grid size.
Each instance must equal, bit for bit, the plain single-instance step of its
own plan; slots not listed must stay untouched; graph replay must equal eager
launches; member sets may change between calls."""
import dataclasses

import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = (32, 64, 128, 256, 512)


def _dev():
  return torch.device('cuda', 0)


def _family(n):
  from jax_cfd_amd import serving
  return serving.SpectralFamily(n, _dev())


def _states(fam, batch, seed0):
  return torch.stack([fam.initial_state(seed0 + s)[0] for s in range(batch)])


def _plan(spec, batch):
  from jax_cfd_amd.spectral import _ns2d
  return _ns2d.Ns2dPlan(spec, batch, _dev())


def _plain(spec, w0, nsteps):
  return _plan(spec, 1).step(w0.clone(), nsteps)


def _setup(sizes=SIZES, batch=3, tableau=None):
  members = []
  for k, n in enumerate(sizes):
    fam = _family(n)
    spec = dataclasses.replace(fam.spec, tableau=tableau)
    members.append((spec, _plan(spec, batch), _states(fam, batch, 100 * k)))
  return members


def _group():
  from jax_cfd_amd.spectral import _ns2d
  return _ns2d.Ns2dGroup(_dev())


@pytest.mark.parametrize("graphs", [True, False])
def test_group_equals_plain_per_instance(graphs):
  members = _setup()
  init = [w.clone() for (_, _, w) in members]
  group = _group()
  group.set_graphs(graphs)
  # slot 1 of every plan is left out; 40 steps = one 32-step graph + 8 eager
  slots, nsteps = [0, 2], 40
  print(f"spectral group: sizes {SIZES}, slots {slots}, {nsteps} steps, graphs={graphs}",
        flush=True)
  group.step([(plan, w, slots) for (_, plan, w) in members], nsteps)
  torch.cuda.synchronize()
  for (spec, _, w), w0 in zip(members, init):
    for s in slots:
      assert torch.equal(w[s], _plain(spec, w0[s], nsteps)), (spec.n, s)
    assert torch.equal(w[1], w0[1]), spec.n


def test_group_1024_with_small_members():
  """The largest size (one column per tile, one row per block) beside 64^2."""
  members = _setup(sizes=(1024, 64), batch=1)
  init = [w.clone() for (_, _, w) in members]
  print("spectral group: 1x1024^2 + 1x64^2, 3 steps", flush=True)
  _group().step([(plan, w, [0]) for (_, plan, w) in members], 3)
  torch.cuda.synchronize()
  for (spec, _, w), w0 in zip(members, init):
    assert torch.equal(w[0], _plain(spec, w0[0], 3)), spec.n


def test_group_member_set_changes_and_graph_equals_eager():
  runs = []
  for graphs in (True, False):
    members = _setup()
    group = _group()
    group.set_graphs(graphs)
    print(f"spectral group member-set changes, graphs={graphs}", flush=True)
    # call 1: every slot of every size; call 2: a smaller, different set
    # (re-uploaded work list, smaller largest grid); call 3: back to all
    group.step([(plan, w, [0, 1, 2]) for (_, plan, w) in members], 33)
    group.step([(plan, w, [1]) for (spec, plan, w) in members if spec.n <= 128], 35)
    group.step([(plan, w, [2, 0]) for (_, plan, w) in members], 7)
    torch.cuda.synchronize()
    runs.append([w.clone() for (_, _, w) in members])
  for wg, we in zip(*runs):
    assert torch.equal(wg, we)
  # against the plain path: slot 1 of the small sizes ran 33 + 35 steps,
  # every other slot 33 + 7
  members = _setup()
  for (spec, _, w0), wg in zip(members, runs[0]):
    for s in range(3):
      nsteps = 33 + (35 if (s == 1 and spec.n <= 128) else 0) + (7 if s != 1 else 0)
      assert torch.equal(wg[s], _plain(spec, w0[s], nsteps)), (spec.n, s, nsteps)


def test_group_three_stage_tableau():
  """crank_nicolson_rk3 (spectral/time_stepping.py:173-183) on every member."""
  rk3 = ((0, 1 / 3, 3 / 4, 1), (0, -5 / 9, -153 / 128), (1 / 3, 15 / 16, 8 / 15))
  members = _setup(sizes=(64, 256), batch=2, tableau=rk3)
  init = [w.clone() for (_, _, w) in members]
  print("spectral group: rk3 tableau, 64^2 + 256^2, 5 steps", flush=True)
  _group().step([(plan, w, [0, 1]) for (_, plan, w) in members], 5)
  torch.cuda.synchronize()
  for (spec, _, w), w0 in zip(members, init):
    for s in range(2):
      assert torch.equal(w[s], _plain(spec, w0[s], 5)), (spec.n, s)


def test_group_timing_and_shape():
  members = _setup(sizes=(64, 128), batch=2)
  group = _group()
  group.step([(plan, w, [0, 1]) for (_, plan, w) in members], 2)
  shape = group.launch_shape()
  # every item is 1024 complex words: column tiles of 1024 / n columns (16
  # for 64^2, capped at n / 2), row blocks of 1024 / n rows (capped at n / 4)
  assert shape['ns_cols_inverse']['items'] == 2 * (32 // 16) + 2 * (64 // 8)
  assert shape['ns_rows']['items'] == 2 * (64 // 16) + 2 * (128 // 8)
  assert all(v['grid'] > 0 and v['bytes'] > 0 for v in shape.values())
  ms = group.time_kernels(3)
  assert all(v > 0 for v in ms.values())


def test_group_rejects_bad_members():
  fam = _family(64)
  plan = _plan(fam.spec, 2)
  w = _states(fam, 2, 0)
  group = _group()
  with pytest.raises(ValueError):
    group.step([(plan, w, [0, 0])], 1)              # repeated slot
  with pytest.raises(ValueError):
    group.step([(plan, w, [2])], 1)                 # out of range
  with pytest.raises(ValueError):
    group.step([(plan, w, [0]), (plan, w, [0])], 1)  # same instance twice
  with pytest.raises(ValueError):
    group.step([(plan, w[:, :32], [0])], 1)         # wrong shape
  r4 = _plan(fam.spec, 1)
  r4.set_tiles(r4.tiles['cols'], r4.tiles['rows'], 4)
  with pytest.raises(NotImplementedError):
    group.step([(r4, w[:1].clone(), [0])], 1)       # radix-4 schedule
  rk3 = ((0, 1 / 3, 3 / 4, 1), (0, -5 / 9, -153 / 128), (1 / 3, 15 / 16, 8 / 15))
  other = _plan(dataclasses.replace(fam.spec, tableau=rk3), 1)
  with pytest.raises(NotImplementedError):
    group.step([(plan, w, [0]), (other, w[:1].clone(), [0])], 1)  # stage counts differ
