#!/usr/bin/env python3
"""Throughput of jax-cfd's FVM time step on one MI355X (gfx950).

Workload: the forced 256x256 periodic MAC-grid run -- fp32, viscosity 1e-3,
density 1, Kolmogorov forcing sin(4y) scale 1 with linear drag 0.1,
advect_van_leer, explicit diffusion, fast-diagonalisation projection,
dt = stable_time_step(v_max=7, CFL=0.5), initial velocity
filtered_velocity_field(seed 1, peak wavenumber 4, v_max 7).  A bench
"step" is one FVM time step of that instance; the default K = 2048 steps is
one full rollout.  `value` is cell-steps/s over the K timed steps with the
state already resident in HBM.

Also reported: per-kernel HIP-event durations and the HBM roofline fraction
of the dominant kernel (algorithmic bytes / mean launch time / 8 TB/s), and
the CPU oracle (NumPy restatement of the reference, one core) timed on a
bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "jax-cfd_amd")):
  if _p not in sys.path:
    sys.path.insert(0, _p)

METRIC = ("served vs plain cell-steps/s at rel-L2≤1e-3; p50/p99 latency; "
          "HBM %peak")
HBM_PEAK_GBS = 8000.0
N, SEED, VMAX, PEAK, NU, CFL = 256, 1, 7.0, 4, 1e-3, 0.5


def _args():
  ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=2048)
  ap.add_argument("--warmup", type=int, default=64)
  ap.add_argument("--kernel-reps", type=int, default=200)
  ap.add_argument("--cpu-seconds", type=float, default=15.0,
                  help="wall-clock budget of the CPU oracle sample (0 = skip)")
  ap.add_argument("--batched-steps", type=int, default=1000,
                  help="steps of the 64 x 512^2 spectral batch (0 = skip)")
  ap.add_argument("--served-requests", type=int, default=400,
                  help="requests of the served Poisson stream (0 = skip)")
  ap.add_argument("--group-steps", type=int, default=256,
                  help="steps of the mixed-size group line (0 skips it)")
  ap.add_argument("--adjoint", type=int, default=1, help="time the adjoint gradient of a 256^2 64-step rollout (0 skips)")
  ap.add_argument("--cg-steps", type=int, default=100,
                  help="steps of the solve_cg variant of the 256^2 run (0 = skip)")
  ap.add_argument("--copy-mib", type=int, default=0,
                  help="first copy this many MiB device-to-device once: a kernel of "
                       "known traffic that calibrates rocprofv3's FETCH_SIZE/WRITE_SIZE")
  ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                  help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; "
                       "both made with --copy-mib) into HBM bytes per dispatch, then exit")
  return ap.parse_args()


def _pmc_rows(root, counter):
  """[(dispatch id, kernel name, counter value)] from rocprofv3's
  *counter_collection.csv files under `root`."""
  import csv
  out = []
  for dirpath, _, files in os.walk(root):
    for fn in files:
      if not fn.endswith("counter_collection.csv"):
        continue
      with open(os.path.join(dirpath, fn), newline="") as f:
        for row in csv.DictReader(f):
          if row.get("Counter_Name") == counter:
            out.append((int(row["Dispatch_Id"]), _kernel_key(row["Kernel_Name"], row["Grid_Size"]),
                        float(row["Counter_Value"])))
  return out


def _kernel_key(name, grid):
  """'fvm_rows_forward<256, 1, 0, 4> grid=65536': the kernel's short name and
  its launch size in work-items (one launch configuration)."""
  short = name.replace("void ", "", 1).replace("jcfd::(anonymous namespace)::", "")
  depth, cut = 0, len(short)
  for i, ch in enumerate(short):  # cut at the parameter list, not inside <...>
    depth += (ch == "<") - (ch == ">")
    if ch == "(" and depth == 0:
      cut = i
      break
  return f"{short[:cut]} grid={int(grid)}"


def _pmc_summary(fetch_dir, write_dir, copy_mib):
  """HBM bytes per dispatch of every kernel: FETCH_SIZE and WRITE_SIZE (KiB
  per dispatch) of the two runs, each scaled so that the calibration copy
  (--copy-mib: the process's first large dispatch, which reads and writes
  copy_mib MiB) reads and writes exactly its size."""
  fetch, write = _pmc_rows(fetch_dir, "FETCH_SIZE"), _pmc_rows(write_dir, "WRITE_SIZE")
  copy_kib = copy_mib * 1024.0

  def calibration(rows):  # the first dispatch of the run that moves a sizeable part of it
    return min((r for r in rows if r[2] >= 0.25 * copy_kib), key=lambda r: r[0])
  f_ref, w_ref = calibration(fetch), calibration(write)
  f_scale, w_scale = copy_kib / f_ref[2], copy_kib / w_ref[2]
  per = {}
  for rows, key, scale in ((fetch, "fetch_bytes", f_scale), (write, "write_bytes", w_scale)):
    for did, name, val in rows:
      if (did, name) in ((f_ref[0], f_ref[1]), (w_ref[0], w_ref[1])):
        continue  # the calibration copy itself
      per.setdefault(name, {}).setdefault(key, []).append(val * 1024.0 * scale)
  kernels = {k: {"dispatches": len(v.get("fetch_bytes", [])),
                 "fetch_bytes": sum(v["fetch_bytes"]) / len(v["fetch_bytes"]),
                 "write_bytes": sum(v["write_bytes"]) / len(v["write_bytes"])}
             for k, v in per.items() if v.get("fetch_bytes") and v.get("write_bytes")}
  return {"copy_mib": copy_mib, "fetch_calibration": {"kernel": f_ref[1], "kib": f_ref[2]},
          "write_calibration": {"kernel": w_ref[1], "kib": w_ref[2]},
          "fetch_scale": f_scale, "write_scale": w_scale, "kernels": kernels}


def _traffic(key):
  """HBM bytes (fetched + written) per launch of one launch configuration
  (see _kernel_key) from the committed rocprofv3 counter summary, or None."""
  path = os.path.join(ROOT, "profiles", "traffic.json")
  try:
    with open(path) as f:
      k = json.load(f)["kernels"].get(key)
  except (OSError, ValueError, KeyError):
    return None
  return None if k is None else round(k["fetch_bytes"] + k["write_bytes"])


def _fvm_kernel_key(n, tiles):
  """_kernel_key of the plain plan's stencil kernel for its tuned tiles."""
  rb, radix = tiles["rows_forward"], tiles["radix"]
  name = (f"fvm_rows_forward_march<{n}, {rb}, {radix}>" if rb >= 16
          else f"fvm_rows_forward<{n}, {rb}, 0, {radix}>")
  return f"{name} grid={n // rb * 256}"


def _cpu_baseline(seconds):
  import numpy as np
  from oracle import fvm as ofvm
  from oracle import grid_ops as go
  from oracle import initial_conditions as oic
  grid = go.Grid.from_domain((N, N), ((0, 2 * np.pi), (0, 2 * np.pi)))
  v = oic.filtered_velocity_field(SEED, grid, VMAX, PEAK)
  dt = ofvm.stable_time_step(VMAX, CFL, NU, grid)
  step = ofvm.FVMStep(grid, dt=dt, viscosity=NU,
                      forcing=go.Forcing(1.0, 4, -0.1))
  v = step(v)  # warm-up
  steps, t0 = 0, time.perf_counter()
  while time.perf_counter() - t0 < seconds or steps < 3:
    v = step(v)
    steps += 1
  wall = time.perf_counter() - t0
  return {"value": N * N * steps / wall, "unit": "cell-steps/s", "cores": 1,
          "kind": "port",
          "sample": f"{steps} steps of the {N}x{N} forced FVM workload, NumPy "
                    f"fp32 restatement of the reference step, single thread, "
                    f"{wall:.1f} s"}


def _batched_spectral(dev, batch, n, steps, reps):
  """64 x 512^2 ForcedNavierStokes2D (viscosity 1e-3, drag 0.1, k = 4),
  CN-RK4, dt 1e-3, seeds 0..batch-1, one batched launch sequence per step
  replayed from hipGraphs."""
  import torch
  from jax_cfd_amd import serving
  fam = serving.SpectralFamily(n, dev)
  w = torch.stack([fam.initial_state(s)[0] for s in range(batch)])
  plan = fam.make_plan(batch)
  from jax_cfd_amd import autotune
  tiles = autotune.tune_ns2d(plan, radices=(8, 4))
  plan.step(w, 8, out=w)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  plan.step(w, steps, out=w)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  assert torch.isfinite(torch.view_as_real(w)).all()
  kms = plan.time_kernels(w, reps)
  kb = plan.kernel_bytes()
  per_step = {k: kms[k] * plan.LAUNCHES_PER_STEP[k] for k in kms}
  dom = max(per_step, key=per_step.get)
  ach = kb[dom] / (kms[dom] * 1e-3) / 1e9
  return {
      "workload": f"spectral {batch}x{n}^2 forced, CN-RK4 dt 1e-3",
      "tiles": {k: tiles[k] for k in ("cols", "rows", "radix")},
      "value": batch * n * n * steps / wall, "unit": "cell-steps/s",
      "steps": steps, "ms_per_step": wall / steps * 1e3,
      "kernels": {k: {"ms_per_launch": round(kms[k], 5), "launches_per_step":
                      plan.LAUNCHES_PER_STEP[k], "algorithmic_bytes": kb[k],
                      "GBps": round(kb[k] / (kms[k] * 1e-3) / 1e9, 1)} for k in kms},
      "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1),
                   "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(ach / HBM_PEAK_GBS, 4)},
  }


def _cg_variant(dev, steps):
  """The forced 256^2 FVM run with solve_cg (rtol = atol = 1e-6) in place of
  the fast-diagonal solve."""
  import functools
  import numpy as np
  import torch
  import jax_cfd_amd.base as cfd
  grid = cfd.grids.Grid((N, N), domain=((0, 2 * np.pi), (0, 2 * np.pi)))
  dt = cfd.equations.stable_time_step(VMAX, CFL, NU, grid)
  step = cfd.equations.semi_implicit_navier_stokes(
      density=1.0, viscosity=NU, dt=dt, grid=grid,
      convect=lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v),
      forcing=cfd.forcings.simple_turbulence_forcing(grid, 1.0, 4, -0.1, "kolmogorov"),
      pressure_solve=functools.partial(cfd.pressure.solve_cg, rtol=1e-6, atol=1e-6))
  v0 = cfd.initial_conditions.filtered_velocity_field(SEED, grid, VMAX, PEAK)
  plan = step.plan(dev)
  u, v = v0[0].data.clone(), v0[1].data.clone()
  plan.step(u, v, 2, out=(u, v))
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  plan.step(u, v, steps, out=(u, v))
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  return {"workload": f"fvm {N}x{N} forced, solve_cg rtol=atol=1e-6", "steps": steps,
          "value": N * N * steps / wall, "unit": "cell-steps/s",
          "ms_per_step": wall / steps * 1e3,
          "cg_iterations_last_step": plan.cg_iterations}


def _plain_step_seconds(dev, solver, n, steps=40):
  """Seconds per step of one instance on the plain path (one instance per
  launch sequence, no graphs, tile shapes tuned for one instance)."""
  import torch
  from jax_cfd_amd import serving
  from jax_cfd_amd import autotune
  fam = serving.FAMILIES[solver](n, dev)
  plan = fam.make_plan(1)
  plan.set_graphs(False)
  (autotune.tune_fvm if solver == "fvm" else autotune.tune_ns2d)(plan, radices=(8,))
  init = fam.initial_state(1)
  if solver == "fvm":
    u, v = init
    run = lambda k: plan.step(u, v, k, out=(u, v))  # noqa: E731
  else:
    (w,) = init
    run = lambda k: plan.step(w, k, out=w)  # noqa: E731
  run(4)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  run(steps)
  torch.cuda.synchronize()
  return (time.perf_counter() - t0) / steps


def _served(dev, count, rate, seed):
  """Poisson stream of mixed FVM / spectral rollouts served with continuous
  batching; the plain figure replays the same mix one instance at a time."""
  import numpy as np
  from jax_cfd_amd import serving
  reqs = serving.poisson_requests(count, rate=rate, seed=seed)
  srv = serving.RolloutServer(hbm_budget_bytes=200e9, tune=True)
  warm = [serving.RolloutRequest(10 ** 6 + i, s, n, 40, i)
          for i, (s, n) in enumerate((s, n) for s in ("fvm", "spectral")
                                     for n in (64, 128, 256, 512))]
  srv.run(warm, realtime=False)  # creates the pools (plans, states)
  t0 = time.perf_counter()
  res = srv.run(reqs, realtime=True)
  wall = time.perf_counter() - t0
  summ = serving.summarize(res, wall)
  plain_s = {}
  for s in ("fvm", "spectral"):
    for n in (64, 128, 256, 512):
      plain_s[(s, n)] = _plain_step_seconds(dev, s, n)
  done = [r for r in res if r.status == "done"]
  plain_wall = sum(r.steps * plain_s[(r.solver, r.n)] for r in done)
  plain_rate = summ["cell_steps"] / plain_wall
  # results agree with the plain path (bitwise in the tests); check a few here
  check = {}
  for r in done:
    if (r.solver, r.n) in check or len(check) >= 4:
      continue
    fam = serving.FAMILIES[r.solver](r.n, dev)
    plan = fam.make_plan(1)
    init = fam.initial_state([q for q in reqs if q.request_id == r.request_id][0].seed)
    if r.solver == "fvm":
      out = [t.cpu().numpy() for t in plan.step(init[0], init[1], r.steps)]
    else:
      out = [plan.step(init[0], r.steps).cpu().numpy()]
    num = sum(float(np.linalg.norm(a - b) ** 2) for a, b in zip(r.state, out))
    den = sum(float(np.linalg.norm(b) ** 2) for b in out)
    check[(r.solver, r.n)] = (num / den) ** 0.5
  summ.update({
      "workload": f"{count} requests, Poisson {rate}/s (seed {seed}), solver in "
                  "{fvm, spectral}, n in {64,128,256,512}, steps U[100,2000], "
                  "200 GB budget",
      "plain_cell_steps_per_s": plain_rate,
      "served_vs_plain": summ["cell_steps_per_s"] / plain_rate,
      "plain_ms_per_step": {f"{s}{n}": round(v * 1e3, 4) for (s, n), v in plain_s.items()},
      "max_rel_l2_vs_plain": max(check.values()) if check else None,
  })
  return summ


def _group_mixed(dev, mix, steps):
  """Uneven sizes in one persistent launch per kernel class (jcfd_fvm_group)
  against the same instances as one batched launch sequence per size on
  concurrent streams (tuned tiles).  FVM physics of the 256^2 run at every
  size, seeds 0.. per size."""
  import torch
  from jax_cfd_amd import autotune, serving
  from jax_cfd_amd.base import _fvm
  members, pools = [], []
  for n, b in mix.items():
    fam = serving.FvmFamily(n, dev)
    u = torch.stack([fam.initial_state(s)[0] for s in range(b)])
    v = torch.stack([fam.initial_state(s)[1] for s in range(b)])
    gplan = fam.make_plan(b)
    members.append((gplan, u, v, list(range(b))))
    pplan = fam.make_plan(b)
    autotune.tune_fvm(pplan, radices=(8,))
    pools.append((pplan, u.clone(), v.clone(),
                  torch.arange(b, dtype=torch.int32, device=dev), torch.cuda.Stream(device=dev)))
  cells = sum(n * n * b for n, b in mix.items())
  group = _fvm.FvmGroup(dev)
  group.step(members, 32)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  group.step(members, steps)
  torch.cuda.synchronize()
  g_wall = time.perf_counter() - t0
  shape = group.launch_shape()

  def run_pools(k):
    for plan, u, v, idx, st in pools:
      st.wait_stream(torch.cuda.current_stream())
      with torch.cuda.stream(st):
        plan.step_slots(u, v, idx, k)
    for *_, st in pools:
      torch.cuda.current_stream().wait_stream(st)
  run_pools(32)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  run_pools(steps)
  torch.cuda.synchronize()
  p_wall = time.perf_counter() - t0
  same = all(torch.equal(m[1], p[1]) and torch.equal(m[2], p[2]) for m, p in zip(members, pools))
  tk = group.time_kernels(10)  # advances the group's members further
  return {"workload": "fvm mixed sizes " + ", ".join(f"{b}x{n}^2" for n, b in mix.items()),
          "steps": steps, "group_value": cells * steps / g_wall,
          "pools_value": cells * steps / p_wall, "unit": "cell-steps/s",
          "group_ms_per_step": g_wall / steps * 1e3,
          "bitwise_equal_to_pools": bool(same),
          "phases": {k: {"grid": shape[k]["grid"], "items": shape[k]["items"],
                         "ms_per_launch": round(tk[k], 5),
                         "GBps": round(shape[k]["bytes"] / (tk[k] * 1e-3) / 1e9, 1)}
                     for k in tk}}


def _adjoint(dev, n=256, steps=64, every=8, reps=3):
  """Gradient of the vorticity MSE after 64 steps of the forced 256^2 run
  with respect to the initial velocity (hand-written adjoint, checkpoints
  every 8 steps), against the forward rollout alone."""
  import numpy as np
  import torch
  import jax_cfd_amd.base as cfd
  from jax_cfd_amd import adjoint
  grid = cfd.grids.Grid((n, n), domain=((0, 2 * np.pi), (0, 2 * np.pi)))
  dt = cfd.equations.stable_time_step(VMAX, CFL, NU, grid)
  step = cfd.equations.semi_implicit_navier_stokes(
      density=1.0, viscosity=NU, dt=dt, grid=grid,
      convect=lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v),
      forcing=cfd.forcings.simple_turbulence_forcing(grid, 1.0, 4, -0.1, "kolmogorov"))
  v0 = cfd.initial_conditions.filtered_velocity_field(SEED, grid, VMAX, PEAK)
  vt = cfd.initial_conditions.filtered_velocity_field(2, grid, VMAX, PEAK)
  plan = step.plan(dev)
  ut, vt_ = plan.step(vt[0].data, vt[1].data, steps)
  target = adjoint.curl_2d(ut, vt_, grid).detach()
  roll = adjoint.rollout(step, steps, every)

  def grad():
    u0 = v0[0].data.clone().requires_grad_(True)
    w0 = v0[1].data.clone().requires_grad_(True)
    u, v = roll(u0, w0)
    loss = ((adjoint.curl_2d(u, v, grid) - target) ** 2).mean()
    return torch.autograd.grad(loss, (u0, w0))
  grad()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(reps):
    g = grad()
  torch.cuda.synchronize()
  g_s = (time.perf_counter() - t0) / reps
  assert all(torch.isfinite(x).all() for x in g)
  t0 = time.perf_counter()
  for _ in range(reps):
    plan.step(v0[0].data, v0[1].data, steps)
  torch.cuda.synchronize()
  f_s = (time.perf_counter() - t0) / reps
  return {"workload": f"fvm {n}^2, {steps}-step rollout, checkpoints every {every}, "
                      "d(vorticity MSE)/d(initial velocity)",
          "ms_per_gradient": g_s * 1e3, "ms_forward_only": f_s * 1e3,
          "gradient_over_forward": g_s / f_s}


def main():
  args = _args()
  if args.pmc_summary:
    print(json.dumps(_pmc_summary(*args.pmc_summary, args.copy_mib), indent=1))
    return
  if args.gpus != 1:
    print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here "
          "(independent instances shard across GPUs with no communication)",
          file=sys.stderr)
    sys.exit(2)
  import numpy as np
  import torch
  import jax_cfd_amd.base as cfd

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  if args.copy_mib > 0:
    src = torch.ones(args.copy_mib * 2 ** 18, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.mul(src, 2.0, out=dst)  # one elementwise kernel: reads and writes copy_mib MiB
    torch.cuda.synchronize()
    del src, dst
  grid = cfd.grids.Grid((N, N), domain=((0, 2 * np.pi), (0, 2 * np.pi)))
  dt = cfd.equations.stable_time_step(VMAX, CFL, NU, grid)
  forcing = cfd.forcings.simple_turbulence_forcing(
      grid, constant_magnitude=1.0, constant_wavenumber=4,
      linear_coefficient=-0.1, forcing_type="kolmogorov")
  step = cfd.equations.semi_implicit_navier_stokes(
      density=1.0, viscosity=NU, dt=dt, grid=grid,
      convect=lambda v: tuple(cfd.advection.advect_van_leer(u, v, dt) for u in v),
      forcing=forcing)
  v0 = cfd.initial_conditions.filtered_velocity_field(SEED, grid, VMAX, PEAK)
  plan = step.plan(dev)
  u, v = v0[0].data.clone(), v0[1].data.clone()
  from jax_cfd_amd import autotune
  # tile choices are cached per (device, size, batch): a second run of the same
  # command (e.g. under a profiler) steps with the same kernels
  tuned = autotune.tune_fvm(plan, u.clone(), v.clone(), radices=(8, 4))
  tiles = {k: tuned[k] for k in ("rows_forward", "poisson_cols", "rows_project", "radix")}
  print(f"[bench] workload fvm {N}x{N} forced, dt={dt:.6g}; warmup {args.warmup}; "
        f"tuned tiles {tiles}", file=sys.stderr, flush=True)
  plan.step(u, v, args.warmup, out=(u, v))
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  plan.step(u, v, args.steps, out=(u, v))
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  assert torch.isfinite(u).all() and torch.isfinite(v).all()
  value = N * N * args.steps / wall
  print(f"[bench] {args.steps} steps in {wall:.4f} s -> {value:.4g} cell-steps/s",
        file=sys.stderr, flush=True)

  kms = plan.time_kernels(u, v, args.kernel_reps)
  kbytes = plan.kernel_bytes()
  dominant = max(kms, key=kms.get)
  achieved = kbytes[dominant] / (kms[dominant] * 1e-3) / 1e9
  roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
              "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
              "traffic": (_traffic(_fvm_kernel_key(N, tiles))
                          if dominant == "fvm_rows_forward" else None),
              "kernel": dominant,
              "bytes_per_launch": kbytes[dominant],
              "ms_per_launch": round(kms[dominant], 6)}
  kernels = {k: {"ms_per_launch": round(kms[k], 6), "algorithmic_bytes": kbytes[k],
                 "GBps": round(kbytes[k] / (kms[k] * 1e-3) / 1e9, 2)} for k in kms}
  print(f"[bench] kernels {json.dumps(kernels)}", file=sys.stderr, flush=True)

  cg = _cg_variant(dev, args.cg_steps) if args.cg_steps > 0 else None
  if cg:
    print(f"[bench] cg variant {json.dumps(cg)}", file=sys.stderr, flush=True)
  batched = None
  if args.batched_steps > 0:
    batched = _batched_spectral(dev, 64, 512, args.batched_steps, 20)
    print(f"[bench] batched {json.dumps(batched)}", file=sys.stderr, flush=True)
  group = None
  if args.group_steps > 0:
    group = _group_mixed(dev, {64: 64, 128: 32, 256: 16, 512: 8}, args.group_steps)
    print(f"[bench] group {json.dumps(group)}", file=sys.stderr, flush=True)
  adj = _adjoint(dev) if args.adjoint else None
  if adj:
    print(f"[bench] adjoint {json.dumps(adj)}", file=sys.stderr, flush=True)
  served = None
  if args.served_requests > 0:
    served = _served(dev, args.served_requests, 200.0, 0)
    print(f"[bench] served {json.dumps(served)}", file=sys.stderr, flush=True)

  cpu = _cpu_baseline(args.cpu_seconds) if args.cpu_seconds > 0 else None
  out = {
      "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
      "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
      "scaling": "weak", "vs_baseline": None, "dtype": "f32",
      "data": "synthetic (seeded filtered_velocity_field)",
      "config": {"workload": "fvm-256x256-kolmogorov-forced-vanleer-fastdiag",
                 "n": N, "dt": dt, "viscosity": NU, "seed": SEED, "v_max": VMAX,
                 "path": "plain (one instance, 3 kernels per step)", "tiles": tiles},
      "roofline": roofline, "cpu_baseline": cpu, "kernels": kernels,
      "batched": batched, "served": served, "cg_variant": cg, "group": group,
      "adjoint": adj,
  }
  print(json.dumps(out), flush=True)


if __name__ == "__main__":
  main()
