#!/usr/bin/env python3
"""Served vs plain throughput of jax-cfd's time steps on one MI355X (gfx950).

Headline workload: a served Poisson stream of rollout requests (400 at
200/s, seed 0), each an FVM or a pseudospectral rollout on a grid of
64^2..512^2 with 100..2000 steps from its own seeded initial field: FVM =
the forced run (viscosity 1e-3, Kolmogorov sin(4y) forcing with drag 0.1,
advect_van_leer, explicit diffusion, fast-diagonal projection, dt =
stable_time_step(7, 0.5)); spectral = ForcedNavierStokes2D (viscosity 1e-3,
drag 0.1, k = 4, 2/3 dealiasing) with CN-RK4 at dt 1e-3.  `value` is the
served cell-steps/s (arrival to result on the host, initial fields built on
the GPU); beside it the same mix run one instance at a time on the plain
path (served_vs_plain), p50/p99 latency, the relative L2 distance of served
results to the fp32 and fp64 oracles, and the HBM roofline of the stream's
dominant kernel.

Also reported: the plain forced 256^2 FVM path (`--steps` steps per call,
repeated; its ms_per_step and its own roofline), the solve_cg variant, the
64 x 512^2 batched spectral run, a mixed-size FVM group, the adjoint, and
the CPU baseline: the oracle (NumPy restatement of the reference steps) on
the host's cores, timed before the GPU is used.
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
  ap.add_argument("--cpu-run-seconds", type=float, default=0.25,
                  help="length of one timed run of the CPU baseline (5 runs per "
                       "solver and size; 0 skips the baseline)")
  ap.add_argument("--batched-steps", type=int, default=1000,
                  help="steps of the 64 x 512^2 spectral batch (0 = skip)")
  ap.add_argument("--served-requests", type=int, default=400,
                  help="requests of the served Poisson stream (the headline value; "
                       "0 reports the plain leg instead)")
  ap.add_argument("--served-group", nargs="*", default=["fvm", "spectral"], choices=["fvm", "spectral"],
                  help="solvers whose pools the headline server groups into one "
                       "mixed-size group step per chunk")
  ap.add_argument("--served-runs", type=int, default=3,
                  help="passes of the served stream (fresh server each); the median is reported")
  ap.add_argument("--served-warmup-passes", type=int, default=1,
                  help="untimed passes of the served stream before the timed ones")
  ap.add_argument("--served-variants", type=int, default=1,
                  help="also time the stream under the other groupings (0 skips)")
  ap.add_argument("--group-steps", type=int, default=256,
                  help="steps of the mixed-size group line (0 skips it)")
  ap.add_argument("--learned-interpolation", type=int, default=1,
                  help="time the learned-interpolation step (0 skips)")
  ap.add_argument("--li-sizes", type=int, nargs="+", default=[256],
                  help="grid sizes of the learned-interpolation leg")
  ap.add_argument("--adjoint", type=int, default=1, help="time the adjoint gradient of a 256^2 64-step rollout (0 skips)")
  ap.add_argument("--cg-steps", type=int, default=100,
                  help="steps of the solve_cg variant of the 256^2 run (0 = skip)")
  ap.add_argument("--copy-mib", type=int, default=0,
                  help="first copy this many MiB device-to-device once: a kernel of "
                       "known traffic that calibrates rocprofv3's FETCH_SIZE/WRITE_SIZE")
  ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                  help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; "
                       "both made with --copy-mib) into HBM bytes per dispatch, then exit")
  ap.add_argument("--trace-summary", metavar="DIR",
                  help="per-launch-configuration kernel statistics of a rocprofv3 "
                       "--kernel-trace run under DIR, then exit")
  ap.add_argument("--trace-filter", metavar="REGEX",
                  help="with --trace-summary: every launch configuration matching REGEX "
                       "instead of the 40 largest")
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


def _trace_summary(root, top=40, pattern=None):
  """Markdown table of kernel durations per launch configuration (kernel
  name with template arguments, grid size in work-items) from rocprofv3
  --kernel-trace CSV files under `root`, largest total time first: the
  `top` largest, or every configuration whose key matches the regular
  expression `pattern`."""
  import csv
  import re
  per = {}
  for dirpath, _, files in os.walk(root):
    for fn in files:
      if not fn.endswith("kernel_trace.csv"):
        continue
      with open(os.path.join(dirpath, fn), newline="") as f:
        for row in csv.DictReader(f):
          grid = row.get("Grid_Size")
          if grid is None:
            grid = (int(row["Grid_Size_X"]) * int(row["Grid_Size_Y"])
                    * int(row["Grid_Size_Z"]))
          key = _kernel_key(row["Kernel_Name"], grid)
          dur = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
          per.setdefault(key, []).append(dur)
  total = sum(sum(v) for v in per.values()) or 1
  lines = ["| kernel (launch) | calls | avg ns | min ns | max ns | % of GPU time |",
           "|---|---|---|---|---|---|"]
  ranked = sorted(per.items(), key=lambda kv: -sum(kv[1]))
  ranked = [kv for kv in ranked if re.search(pattern, kv[0])] if pattern else ranked[:top]
  for key, d in ranked:
    lines.append(f"| `{key}` | {len(d)} | {sum(d) / len(d):.0f} | {min(d)} | {max(d)} | "
                 f"{100.0 * sum(d) / total:.2f} |")
  return "\n".join(lines)


def _cpu_cores():
  """Host cores the CPU baseline uses: this process's CPU affinity, capped
  at 16 (the host share that comes with one GPU)."""
  try:
    n = len(os.sched_getaffinity(0))
  except AttributeError:
    n = os.cpu_count() or 1
  return max(1, min(n, 16))


def _cpu_model():
  try:
    with open("/proc/cpuinfo") as f:
      for line in f:
        if line.startswith("model name"):
          return line.split(":", 1)[1].strip()
  except OSError:
    pass
  return "unknown"


def _oracle_case(case, n, seed, dtype=None):
  """(step function, initial state) of the oracle for one workload:
  'fvm' = the forced FVM physics (viscosity 1e-3, sin(4y) forcing, drag
  0.1, Van Leer, fast-diagonal projection, dt = stable_time_step(7, 0.5),
  filtered_velocity_field(seed, v_max 7, peak 4)); 'fvm0' = the unforced
  64^2 case (v_max 2, dt = stable_time_step(2, 0.5)); 'spectral' =
  ForcedNavierStokes2D(1e-3, smooth) with CN-RK4 at dt 1e-3 from
  rfft2(curl(filtered_velocity_field(seed, v_max 1, peak 4))).  With seed
  None the initial state is None; `dtype` float64 gives the fp64 oracle."""
  import numpy as np
  from oracle import fvm as ofvm
  from oracle import grid_ops as go
  from oracle import initial_conditions as oic
  from oracle import spectral as osp
  dtype = np.float32 if dtype is None else dtype
  grid = go.Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  if case == "spectral":
    init = None
    if seed is not None:
      v = oic.filtered_velocity_field(seed, grid, 1.0, PEAK)
      ct = np.complex64 if dtype == np.float32 else np.complex128
      init = oic.vorticity_hat_from_velocity(v, grid).astype(np.complex64).astype(ct)
    return osp.crank_nicolson_rk4(osp.ForcedNavierStokes2D(NU, grid, True, dtype=dtype),
                                  1e-3), init
  vmax = 2.0 if case == "fvm0" else VMAX
  dt = ofvm.stable_time_step(vmax, CFL, NU, grid)
  forcing = None if case == "fvm0" else go.Forcing(1.0, 4, -0.1)
  step = ofvm.FVMStep(grid, dt=dt, viscosity=NU, forcing=forcing)
  init = None
  if seed is not None:
    v = oic.filtered_velocity_field(seed, grid, vmax, PEAK)
    init = tuple(go.A(u.data.astype(dtype), u.offset) for u in v)
  return step, init


def _cpu_job(job):
  """Per-step seconds of `runs` timed runs of the oracle's step (run in a
  worker process).  Each run is `steps` steps from the initial field, or,
  when steps is 0, as many steps as take about `run_s` seconds, continuing
  the rollout."""
  case, n, runs, run_s, steps, fft_workers = job
  from oracle import fftlib
  fftlib.set_workers(fft_workers)
  step, init = _oracle_case(case, n, seed=0 if case == "fvm0" else 1)
  state = step(init)  # warm-up (plan constants, allocator)
  fixed = steps > 0
  if not fixed:
    t0 = time.perf_counter()
    state = step(state)
    steps = max(1, int(round(run_s / max(time.perf_counter() - t0, 1e-6))))
  out = []
  for _ in range(runs):
    if fixed:
      state = init
    t0 = time.perf_counter()
    for _ in range(steps):
      state = step(state)
    out.append((time.perf_counter() - t0) / steps)
  return out


def _cpu_baseline(requests, run_s):
  """The oracle (the NumPy restatement of the reference's steps) on this
  host's cores, before the GPU is touched.  Per (solver, n) pool of the
  served stream, one instance per core runs concurrently (the way a CPU
  server would spend its cores), 5 timed runs each; the served stream's
  CPU rate is its cell-steps over the time its requests would take at those
  per-step times spread over the cores.  Also the per-config rows: the
  64^2 unforced 100-step case on one core, one forced 256^2 instance with
  scipy.fft threaded over every core, and the all-core 256^2 FVM and 512^2
  spectral rates.  Medians everywhere."""
  import multiprocessing
  import statistics
  cores = _cpu_cores()
  ctx = multiprocessing.get_context("spawn")  # fresh interpreters, no GPU state
  t_start = time.perf_counter()
  per_step = {}
  with ctx.Pool(cores) as pool:
    for solver in ("fvm", "spectral"):
      for n in (64, 128, 256, 512):
        runs = pool.map(_cpu_job, [(solver, n, 5, run_s, 0, 1)] * cores, chunksize=1)
        per_step[(solver, n)] = statistics.median(t for r in runs for t in r)
    c0 = pool.apply(_cpu_job, (("fvm0", 64, 5, 0.0, 100, 1),))
  c1_one = _cpu_job(("fvm", N, 5, run_s, 0, cores))
  work = sum(r.n * r.n * r.steps for r in requests)
  cpu_time = sum(r.steps * per_step[(r.solver, r.n)] for r in requests) / cores
  rows = {
      "C0_fvm64_unforced_100_steps_one_core": {
          "s_per_run": statistics.median(c0) * 100,
          "cell_steps_per_s": 64 * 64 / statistics.median(c0)},
      "C1_fvm256_forced_one_instance_fft_on_all_cores": {
          "ms_per_step": statistics.median(c1_one) * 1e3,
          "cell_steps_per_s": N * N / statistics.median(c1_one)},
      "C1_fvm256_forced_all_cores": {
          "cell_steps_per_s": cores * N * N / per_step[("fvm", 256)]},
      "C2_spectral512_forced_all_cores": {
          "cell_steps_per_s": cores * 512 * 512 / per_step[("spectral", 512)]},
  }
  return {"value": work / cpu_time, "unit": "cell-steps/s", "cores": cores,
          "kind": "port",
          "sample": f"oracle (NumPy fp32 restatement of the reference steps) on "
                    f"{cores} cores: per (solver, n) of the served stream, {cores} "
                    f"concurrent instances x 5 runs of ~{run_s:g} s, median per-step "
                    f"time; served-stream rate = its cell-steps / (sum of steps x "
                    f"per-step time / {cores} cores)",
          "cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(),
          "seconds": round(time.perf_counter() - t_start, 1),
          "ms_per_step": {f"{s}{n}": round(t * 1e3, 3) for (s, n), t in per_step.items()},
          "rows": rows}


def _rel_l2(a, b):
  import numpy as np
  return float(np.sqrt(sum(float(np.linalg.norm(np.asarray(x) - np.asarray(y)) ** 2)
                           for x, y in zip(a, b))
                       / sum(float(np.linalg.norm(np.asarray(y)) ** 2) for y in b)))


def _oracle_distance(solver, n, init, steps, got):
  """Relative L2 distance of a GPU result `got` (host arrays) after `steps`
  steps from `init` (host arrays) to the fp32 oracle's and to the fp64
  oracle's run of the same steps from the same field, and the fp32
  oracle's own distance to the fp64 run."""
  import numpy as np
  from oracle import grid_ops as go
  out = {}
  for label, dtype in (("fp32", np.float32), ("fp64", np.float64)):
    step, _ = _oracle_case(solver, n, None, dtype)
    if solver == "spectral":
      ct = np.complex64 if dtype == np.float32 else np.complex128
      state = np.asarray(init[0]).astype(ct)
    else:
      grid = go.Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
      state = tuple(go.A(np.asarray(x).astype(dtype), o)
                    for x, o in zip(init, grid.cell_faces))
    for _ in range(steps):
      state = step(state)
    out[label] = [state] if solver == "spectral" else [u.data for u in state]
  return {"steps": steps,
          "rel_l2_vs_oracle_fp32": _rel_l2(got, out["fp32"]),
          "rel_l2_vs_oracle_fp64": _rel_l2(got, out["fp64"]),
          "oracle_fp32_vs_fp64": _rel_l2(out["fp32"], out["fp64"])}


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

  def run(persistent):
    plan.set_cg_persistent(persistent)
    u, v = v0[0].data.clone(), v0[1].data.clone()
    plan.step(u, v, 2, out=(u, v))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(u, v, steps, out=(u, v))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, plan.cg_iterations
  wall1, _ = run(True)  # one cooperative launch per solve
  wall, iters = run(False)  # two launches per iteration, flag read every 64 (default)
  return {"workload": f"fvm {N}x{N} forced, solve_cg rtol=atol=1e-6", "steps": steps,
          "value": N * N * steps / wall, "unit": "cell-steps/s",
          "ms_per_step": wall / steps * 1e3,
          "us_per_iteration": wall / steps / (iters + 1) * 1e6,
          "cg_iterations_last_step": iters,
          "solve": "two launches per iteration",
          "cooperative_form_ms_per_step": wall1 / steps * 1e3}


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


def _per_solver(res):
  """Cell-steps per second of each solver's requests over their own span
  (first arrival to last result)."""
  out = {}
  for s in ("fvm", "spectral"):
    rs = [r for r in res if r.solver == s and r.status == "done"]
    if rs:
      span = max(r.finished for r in rs) - min(r.arrival for r in rs)
      out[s] = sum(r.n * r.n * r.steps for r in rs) / span
  return out


def _copy_rate(dev, nbytes, reps=20):
  """GB/s of a device tensor copy moving nbytes (half read, half written):
  what a plain streaming kernel attains on this GPU, reported beside a
  kernel's algorithmic rate."""
  import torch
  n = max(1, int(nbytes) // 8)
  src = torch.ones(n, dtype=torch.float32, device=dev)
  dst = torch.empty_like(src)
  for _ in range(3):
    dst.copy_(src)
  e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    dst.copy_(src)
  e1.record()
  e1.synchronize()
  ms = e0.elapsed_time(e1) / reps
  del src, dst
  return 2 * n * 4 / (ms * 1e-3) / 1e9


def _served_run(dev, requests, group=()):
  """One timed pass of the served stream on a warmed server."""
  import torch
  from jax_cfd_amd import serving
  srv = serving.RolloutServer(hbm_budget_bytes=200e9, tune=True, group=group)
  warm = [serving.RolloutRequest(10 ** 6 + i, s, n, 40, i)
          for i, (s, n) in enumerate((s, n) for s in ("fvm", "spectral")
                                     for n in (64, 128, 256, 512))]
  srv.run(warm, realtime=False)  # creates the pools (plans, tuned tiles, states)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  res = srv.run(requests, realtime=True)
  wall = time.perf_counter() - t0
  summ = serving.summarize(res, wall)
  summ["per_solver_cell_steps_per_s"] = _per_solver(res)
  summ["grouped"] = list(group)
  return srv, res, wall, summ


def _served(dev, requests, kernel_reps, group=(), runs=3, warmup_passes=1):
  """The served Poisson stream: continuous batching over one pool per
  (solver, n) (tuned tiles, radix-8 schedule), `runs` passes on fresh
  servers, the median pass reported (a pass's rate varies by a few percent
  from run to run), after `warmup_passes` untimed passes (the first pass
  of a process is the one most often in the stream's slow mode).  The plain figure replays the same mix one instance at
  a time on the plain path (tuned single-instance plans, no graphs).  The
  roofline kernel is the one with the largest estimated share of the
  stream's GPU time (launches per step x HIP-event time per launch at the
  pool's full capacity x the pool's steps at full capacity), timed on its
  pool's own plan and state."""
  import numpy as np
  import torch
  from jax_cfd_amd import serving
  for _ in range(max(0, warmup_passes)):
    _served_run(dev, requests, group)
  passes = [_served_run(dev, requests, group) for _ in range(max(1, runs))]
  passes.sort(key=lambda p: p[3]["cell_steps_per_s"])
  srv, res, wall, summ = passes[len(passes) // 2]
  summ["passes_cell_steps_per_s"] = [p[3]["cell_steps_per_s"] for p in passes]
  summ["warmup_passes"] = max(0, warmup_passes)
  print(f"[bench] served {summ['done']} of {len(requests)} requests in {wall:.3f} s",
        file=sys.stderr, flush=True)
  done = [r for r in res if r.status == "done"]
  by_id = {q.request_id: q for q in requests}
  # per-kernel times of every pool at full capacity; the stream's dominant kernel
  est, pools = {}, {}
  for (solver, n), pool in srv.pools.items():
    plan = pool.plan
    kms = (plan.time_kernels(*pool.state, kernel_reps) if solver == "fvm"
           else plan.time_kernels(pool.state[0], kernel_reps))
    kb = plan.kernel_bytes()
    lps = plan.LAUNCHES_PER_STEP if solver == "spectral" else {k: 1 for k in kms}
    steps = sum(r.steps for r in done if (r.solver, r.n) == (solver, n))
    pools[f"{solver}{n}"] = {"capacity": pool.capacity, "tiles": plan.tiles,
                             "ms_per_launch": {k: round(v, 5) for k, v in kms.items()}}
    for k in kms:
      est[(solver, n, k)] = (steps / pool.capacity * lps[k] * kms[k], kms[k], kb[k], plan)
  (solver, n, kname), (_, ms, nbytes, plan) = max(est.items(), key=lambda kv: kv[1][0])
  achieved = nbytes / (ms * 1e-3) / 1e9
  tiles = plan.tiles
  key = (_ns2d_kernel_key(kname, n, plan.batch, tiles) if solver == "spectral"
         else _fvm_plan_kernel_key(kname, n, plan.batch, tiles))
  roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
              "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
              "traffic": _traffic(key), "kernel": kname, "launch": key,
              "pool": f"{solver} {n}^2 x {plan.batch} slots",
              "bytes_per_launch": nbytes, "ms_per_launch": round(ms, 5)}
  copy_gbps = _copy_rate(dev, nbytes)
  roofline["copy_GBps"] = round(copy_gbps, 1)
  roofline["frac_of_copy"] = round(achieved / copy_gbps, 4)
  plain_s = {}
  for s in ("fvm", "spectral"):
    for n in (64, 128, 256, 512):
      plain_s[(s, n)] = _plain_step_seconds(dev, s, n)
  plain_wall = sum(r.steps * plain_s[(r.solver, r.n)] for r in done)
  plain_rate = summ["cell_steps"] / plain_wall
  # served == plain (bitwise in the tests): the shortest request of each pool
  vs_plain, shortest = {}, {}
  for r in done:
    k = (r.solver, r.n)
    if k not in shortest or r.steps < shortest[k].steps:
      shortest[k] = r
  inits = {}
  for (s, n), r in sorted(shortest.items()):
    fam = serving.FAMILIES[s](n, dev)
    p1 = fam.make_plan(1)
    init = fam.initial_state(by_id[r.request_id].seed)
    inits[r.request_id] = [t.cpu().numpy() for t in init]
    out = ([t.cpu().numpy() for t in p1.step(init[0], init[1], r.steps)] if s == "fvm"
           else [p1.step(init[0], r.steps).cpu().numpy()])
    vs_plain[f"{s}{n}"] = _rel_l2(r.state, out)
    p1.close()
  # the metric's qualifier: the smallest request of each solver against the oracle
  vs_oracle = {}
  for s in ("fvm", "spectral"):
    cand = [r for k, r in shortest.items() if k[0] == s]
    if not cand:
      continue
    r = min(cand, key=lambda q: q.n * q.n * q.steps)
    d = _oracle_distance(s, r.n, inits[r.request_id], r.steps, r.state)
    vs_oracle[f"{s}{r.n}_request{r.request_id}"] = d
  summ.update({
      "plain_cell_steps_per_s": plain_rate,
      "served_vs_plain": summ["cell_steps_per_s"] / plain_rate,
      "plain_ms_per_step": {f"{s}{n}": round(v * 1e3, 4) for (s, n), v in plain_s.items()},
      "rel_l2_vs_plain": vs_plain,
      "vs_oracle": vs_oracle,
      "max_rel_l2_vs_oracle_fp32": max(d["rel_l2_vs_oracle_fp32"] for d in vs_oracle.values()),
      "pools": pools,
  })
  return summ, roofline


def _ns2d_kernel_key(kname, n, batch, tiles):
  """_kernel_key of a spectral plan's step kernel for its tiles."""
  w, rb, radix = tiles["cols"], tiles["rows"], tiles["radix"]
  cols_grid = batch * (n // 2) // w * 256
  name = {"ns_cols_inverse": f"ns_cols_inverse<{n}, {w}, 4, {radix}>",
          "ns_rows": f"ns_rows<{n}, {rb}, {radix}>",
          "ns_cols_update_inverse": f"ns_cols_forward<{n}, {w}, 2, {radix}>",
          "ns_cols_update": f"ns_cols_forward<{n}, {w}, 1, {radix}>"}[kname]
  grid = batch * n // rb * 256 if kname == "ns_rows" else cols_grid
  return f"{name} grid={grid}"


def _fvm_plan_kernel_key(kname, n, batch, tiles):
  """_kernel_key of an FVM plan's step kernel for its tiles."""
  radix = tiles["radix"]
  if kname == "fvm_rows_forward":
    rb = tiles["rows_forward"]
    name = (f"fvm_rows_forward_march<{n}, {rb}, {radix}>" if rb >= 16
            else f"fvm_rows_forward<{n}, {rb}, 0, {radix}>")
    return f"{name} grid={batch * n // rb * 256}"
  if kname == "fvm_poisson_cols":
    w = tiles["poisson_cols"]
    return f"fvm_poisson_cols<{n}, {w}, {radix}> grid={batch * (n // 2) // w * 256}"
  rb = tiles["rows_project"]
  return f"fvm_rows_project<{n}, {rb}, 0, {radix}> grid={batch * n // rb * 256}"


def _plain_fvm(dev, steps, warmup, kernel_reps, min_seconds=0.2):
  """The plain forced 256^2 FVM path (one instance, 3 launches per step):
  `steps`-step calls repeated until at least `min_seconds` have passed (and
  5 calls), median seconds per call / steps."""
  import statistics
  import numpy as np
  import torch
  import jax_cfd_amd.base as cfd
  from jax_cfd_amd import autotune
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
  # tile choices are cached per (device, size, batch): a second run of the same
  # command (e.g. under a profiler) steps with the same kernels
  tuned = autotune.tune_fvm(plan, u.clone(), v.clone(), radices=(8, 4))
  tiles = {k: tuned[k] for k in ("rows_forward", "poisson_cols", "rows_project", "radix")}
  plan.step(u, v, warmup, out=(u, v))
  torch.cuda.synchronize()
  # the fp64 figure: 100 steps from the initial field against both oracles
  got = [t.cpu().numpy() for t in plan.step(v0[0].data, v0[1].data, 100)]
  fp64 = _oracle_distance("fvm", N, [v0[0].data.cpu().numpy(), v0[1].data.cpu().numpy()],
                          100, got)
  calls = []
  t_start = time.perf_counter()
  while len(calls) < 5 or time.perf_counter() - t_start < min_seconds:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(u, v, steps, out=(u, v))
    torch.cuda.synchronize()
    calls.append(time.perf_counter() - t0)
  assert torch.isfinite(u).all() and torch.isfinite(v).all()
  sec = statistics.median(calls) / steps
  kms = plan.time_kernels(u, v, kernel_reps)
  kbytes = plan.kernel_bytes()
  dominant = max(kms, key=kms.get)
  achieved = kbytes[dominant] / (kms[dominant] * 1e-3) / 1e9
  key = _fvm_plan_kernel_key(dominant, N, 1, tiles)
  return {
      "workload": f"fvm {N}x{N} forced (Kolmogorov sin(4y), drag 0.1), Van Leer, "
                  f"fast-diagonal projection, dt {dt:.6g}",
      "path": "plain (one instance, 3 launches per step, no graph)",
      "tiles": tiles, "steps_per_call": steps, "calls": len(calls),
      "ms_per_step": sec * 1e3, "value": N * N / sec, "unit": "cell-steps/s",
      "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                   "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                   "traffic": _traffic(key), "kernel": dominant, "launch": key,
                   "bytes_per_launch": kbytes[dominant],
                   "ms_per_launch": round(kms[dominant], 6)},
      "kernels": {k: {"ms_per_launch": round(kms[k], 6), "algorithmic_bytes": kbytes[k],
                      "GBps": round(kbytes[k] / (kms[k] * 1e-3) / 1e9, 2)} for k in kms},
      "fp64": fp64,
  }


def _group_mixed(dev, mix, steps, solver="fvm"):
  """Uneven sizes in one launch per kernel class, one workgroup per work
  item (jcfd_fvm_group / jcfd_ns2d_group), against the same instances as
  one batched launch sequence per size on concurrent streams (tuned
  tiles).  Physics of the
  served families (FVM: the forced 256^2 run; spectral: the forced 512^2
  run) at every size, seeds 0.. per size."""
  import torch
  from jax_cfd_amd import autotune, serving
  from jax_cfd_amd.base import _fvm
  from jax_cfd_amd.spectral import _ns2d
  fvm = solver == "fvm"
  members, pools = [], []
  for n, b in mix.items():
    fam = serving.FAMILIES[solver](n, dev)
    fields = [torch.stack([fam.initial_state(s)[k] for s in range(b)])
              for k in range(2 if fvm else 1)]
    members.append((fam.make_plan(b), *fields, list(range(b))))
    pplan = fam.make_plan(b)
    (autotune.tune_fvm if fvm else autotune.tune_ns2d)(pplan, radices=(8,))
    pools.append((pplan, [f.clone() for f in fields],
                  torch.arange(b, dtype=torch.int32, device=dev), torch.cuda.Stream(device=dev)))
  cells = sum(n * n * b for n, b in mix.items())
  group = _fvm.FvmGroup(dev) if fvm else _ns2d.Ns2dGroup(dev)
  group.step(members, 32)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  group.step(members, steps)
  torch.cuda.synchronize()
  g_wall = time.perf_counter() - t0
  shape = group.launch_shape()

  def run_pools(k):
    for plan, fields, idx, st in pools:
      st.wait_stream(torch.cuda.current_stream())
      with torch.cuda.stream(st):
        plan.step_slots(*fields, idx, k)
    for *_, st in pools:
      torch.cuda.current_stream().wait_stream(st)
  run_pools(32)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  run_pools(steps)
  torch.cuda.synchronize()
  p_wall = time.perf_counter() - t0
  same = all(all(torch.equal(a, b) for a, b in zip(m[1:-1], p[1]))
             for m, p in zip(members, pools))
  tk = group.time_kernels(10)  # advances the group's instances further
  return {"workload": f"{solver} mixed sizes " + ", ".join(f"{b}x{n}^2" for n, b in mix.items()),
          "steps": steps, "group_value": cells * steps / g_wall,
          "pools_value": cells * steps / p_wall, "unit": "cell-steps/s",
          "group_ms_per_step": g_wall / steps * 1e3,
          "bitwise_equal_to_pools": bool(same),
          "phases": {k: {"grid": shape[k]["grid"], "items": shape[k]["items"],
                         "ms_per_launch": round(tk[k], 5),
                         "GBps": round(shape[k]["bytes"] / (tk[k] * 1e-3) / 1e9, 1)}
                     for k in tk}}


def _learned_interpolation(dev, n=256, steps=20, reps=10):
  """The learned-interpolation FVM step with the tower of the reference's
  inference demo (7 periodic 3x3 convolutions: 2 -> 64, five 64 -> 64,
  64 -> 120 logits for 8 interpolations of 4 x 4 stencils), seeded
  haiku-default weights (output layer scaled by 0.1), forced 256^2 physics,
  implicit diffusion (modular_navier_stokes_model's default solver).  The
  tower runs on the matrix cores in fp32: its roofline is fp32 MFMA."""
  import functools
  import numpy as np
  import torch
  import jax_cfd_amd.base as cfd
  from jax_cfd_amd import ml
  grid = cfd.grids.Grid((n, n), domain=((0, 2 * np.pi), (0, 2 * np.pi)))
  dt = cfd.equations.stable_time_step(VMAX, CFL, NU, grid, implicit_diffusion=True)
  tower = functools.partial(ml.towers.forward_tower_factory, num_hidden_channels=64,
                            num_hidden_layers=6)
  params = tower(120, 2).init_params(0)
  params[-1] = (params[-1][0] * 0.1, params[-1][1])
  conv = functools.partial(
      ml.advections.self_advection, advection_module=functools.partial(
          ml.advections.modular_self_advection, interpolation_module=functools.partial(
              ml.interpolations.FusedLearnedInterpolation, tags=("u", "c"),
              tower_factory=tower, params=params)))
  forcing = functools.partial(cfd.forcings.simple_turbulence_forcing, constant_magnitude=1.0,
                              constant_wavenumber=4, linear_coefficient=-0.1,
                              forcing_type="kolmogorov")
  specs = ml.physics_specifications.NavierStokesPhysicsSpecs(1.0, NU, forcing)
  step = ml.equations.modular_navier_stokes_model(grid, dt, specs, convection_module=conv)
  v0 = cfd.initial_conditions.filtered_velocity_field(SEED, grid, VMAX, PEAK)
  plan = step.plan(dev)
  u, v = v0[0].data.clone(), v0[1].data.clone()
  plan.step(u, v, 2, out=(u, v))
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  plan.step(u, v, steps, out=(u, v))
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  assert torch.isfinite(u).all() and torch.isfinite(v).all()
  ms = plan.time_kernels(u, v, reps)
  flops = plan.kernel_flops()
  layers = [{"ms_per_launch": round(ms[k], 5), "flops": flops[k],
             "TFLOPs": round(flops[k] / (ms[k] * 1e-3) / 1e12, 2)}
            for k in range(len(flops))]
  dom = max(range(len(flops)), key=lambda k: ms[k])
  ach = flops[dom] / (ms[dom] * 1e-3) / 1e12
  tower_ms = sum(ms[:len(flops)])
  return {"workload": f"learned interpolation {n}^2, tower 2-64x6-120 (7 convolutions), "
                      "forced physics, implicit diffusion",
          "steps": steps, "ms_per_step": wall / steps * 1e3,
          "value": n * n * steps / wall, "unit": "cell-steps/s",
          "tower_TFLOPs": round(sum(flops) / (tower_ms * 1e-3) / 1e12, 2),
          "layers": layers, "explicit_ms": round(ms[-2], 5), "projection_ms": round(ms[-1], 5),
          "roofline": {"bound": "mfma", "kernel": f"li_conv3x3 layer {dom}",
                       "achieved": round(ach, 2), "peak": 157.3, "unit": "TFLOP/s",
                       "frac": round(ach / 157.3, 4)}}


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
  if args.trace_summary:
    print(_trace_summary(args.trace_summary, pattern=args.trace_filter))
    return
  if args.gpus != 1:
    print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here "
          "(independent instances shard across GPUs with no communication)",
          file=sys.stderr)
    sys.exit(2)
  sys.path.insert(0, os.path.join(ROOT, "jax-cfd_amd"))
  from jax_cfd_amd import serving
  requests = serving.poisson_requests(args.served_requests, rate=200.0, seed=0)
  # the CPU baseline runs first, in worker processes started before the GPU is used
  cpu = _cpu_baseline(requests, args.cpu_run_seconds) if args.cpu_run_seconds > 0 else None
  if cpu:
    print(f"[bench] cpu baseline {json.dumps(cpu)}", file=sys.stderr, flush=True)
  import torch
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  if args.copy_mib > 0:
    src = torch.ones(args.copy_mib * 2 ** 18, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.mul(src, 2.0, out=dst)  # one elementwise kernel: reads and writes copy_mib MiB
    torch.cuda.synchronize()
    del src, dst

  plain = _plain_fvm(dev, args.steps, args.warmup, args.kernel_reps)
  print(f"[bench] plain {json.dumps(plain)}", file=sys.stderr, flush=True)
  headline_group = tuple(sorted(set(args.served_group)))
  served, roofline = (_served(dev, requests, 20, headline_group, args.served_runs,
                              args.served_warmup_passes) if requests
                      else (None, None))
  variants = None
  if served:
    print(f"[bench] served {json.dumps(served)}", file=sys.stderr, flush=True)
    if args.served_variants:
      variants = {}
      for grp in ((), ("fvm",), ("spectral",), ("fvm", "spectral")):
        if grp != headline_group:
          v = _served_run(dev, requests, grp)[3]
          variants["+".join(grp) or "none"] = v
          print(f"[bench] served, grouped {grp}: {json.dumps(v)}", file=sys.stderr, flush=True)
  cg = _cg_variant(dev, args.cg_steps) if args.cg_steps > 0 else None
  if cg:
    print(f"[bench] cg variant {json.dumps(cg)}", file=sys.stderr, flush=True)
  batched = None
  if args.batched_steps > 0:
    batched = _batched_spectral(dev, 64, 512, args.batched_steps, 20)
    print(f"[bench] batched {json.dumps(batched)}", file=sys.stderr, flush=True)
  group = group_spectral = None
  if args.group_steps > 0:
    mix = {64: 64, 128: 32, 256: 16, 512: 8}
    group = _group_mixed(dev, mix, args.group_steps)
    print(f"[bench] group {json.dumps(group)}", file=sys.stderr, flush=True)
    group_spectral = _group_mixed(dev, mix, args.group_steps, "spectral")
    print(f"[bench] spectral group {json.dumps(group_spectral)}", file=sys.stderr, flush=True)
  li = ([_learned_interpolation(dev, n) for n in args.li_sizes]
        if args.learned_interpolation else None)
  if li and len(li) == 1:
    li = li[0]
  if li:
    print(f"[bench] learned interpolation {json.dumps(li)}", file=sys.stderr, flush=True)
  adj = _adjoint(dev) if args.adjoint else None
  if adj:
    print(f"[bench] adjoint {json.dumps(adj)}", file=sys.stderr, flush=True)

  if served:
    value, unit = served["cell_steps_per_s"], "cell-steps/s"
    workload = (f"served Poisson stream: {len(requests)} rollout requests at 200/s "
                "(seed 0), solver in {fvm (forced 256^2 physics), spectral (forced "
                "512^2 physics)}, n in {64,128,256,512}, steps U[100,2000], "
                "200 GB HBM budget, same-size steps batched per launch")
  else:
    value, unit = plain["value"], "cell-steps/s"
    workload = plain["workload"]
  out = {
      "metric": METRIC, "value": value, "unit": unit, "n_gpus": 1,
      "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": plain["ms_per_step"], "higher_is_better": True,
      "scaling": "weak", "vs_baseline": None, "dtype": "f32",
      "data": "synthetic (seeded filtered_velocity_field initial fields)",
      "config": {"workload": workload,
                 "value": "served cell-steps/s (arrival to result on the host)",
                 "step": "steps, warmup and ms_per_step are the plain forced 256^2 "
                         "FVM leg's (one instance, --steps steps per call)"},
      "served_vs_plain": served["served_vs_plain"] if served else None,
      "p50_latency_s": served["p50_latency_s"] if served else None,
      "p99_latency_s": served["p99_latency_s"] if served else None,
      "max_rel_l2_vs_oracle": served["max_rel_l2_vs_oracle_fp32"] if served else None,
      "roofline": roofline if roofline else plain["roofline"],
      "cpu_baseline": cpu,
      "served": served, "served_variants": variants, "plain": plain, "batched": batched, "cg_variant": cg,
      "group": group, "group_spectral": group_spectral, "adjoint": adj, "learned_interpolation": li,
  }
  print(json.dumps(out), flush=True)


if __name__ == "__main__":
  main()
