"""Continuous-batching rollout server for mixed FVM / spectral requests.

Each request asks for `steps` steps of one solver on one grid size, from the
seeded initial field of that solver (FVM: filtered_velocity_field; spectral:
rfft2 of its curl).  The server keeps every admitted request's state resident
on the GPU in a slot of a *pool* -- one pool per (solver, n), holding one
batched plan and a (capacity, ...) state tensor -- and advances all active
slots of a pool together with one slot-indexed launch sequence
(jcfd_fvm_step_slots / jcfd_ns2d_step_slots).  Pools run on their own HIP
streams, so their kernels overlap on the device; the host loop polls each
pool's completion event and, as soon as a pool's chunk is done, retires the
requests that reached their step count, admits waiting requests into the
freed slots and relaunches it.  A chunk advances every active slot by the
smallest remaining step count, so no request overshoots, capped at
`max_chunk` steps (default 16).  Short chunks matter twice: pools share the
device's few hardware queues, so one pool's long chunk holds back the pools
queued behind it, and a pool's active-slot count changes from chunk to
chunk, which would make a plan re-capture its step graph (a chunk below the
plan's graph length runs as plain launches).  On the mixed 400-request
stream, a cap of 16 instead of 128 raised throughput 1.05e10 -> 1.27e10
cell-steps/s and cut p99 latency 2.8 s -> 1.6 s.

Admission is bounded by a device-memory budget: a pool's bytes (plan scratch
+ states) are reserved when it is created; a pool that does not fit waits
until other pools drain, and a request that can never fit is rejected.

Every slot gets exactly the arithmetic of a lone instance (the kernels never
mix instances), so a request's result is bit-identical alone or in a mix.

Grouping (`group`, default both solvers): the pools of a grouped solver
share one stream and advance together, each chunk ONE mixed-size group step
(jcfd_fvm_group_step / jcfd_ns2d_group_step) over the active slots of all of
them -- one launch per kernel class instead of one launch sequence per pool;
the chunk is then the smallest over the solver's pools.  Pools above the
family's GROUP_MAX_N (spectral: 256) keep their own stream.  Results are
unchanged (bit for bit).  On the mixed 400-request stream (three runs each):
per-pool FVM and spectral launches with FVM grouped 1.48-1.51e10
cell-steps/s, p50 0.30-0.34 s, p99 0.74-1.0 s; with the spectral pools up
to 256^2 grouped too 1.54e10, p50 0.20-0.22 s, p99 0.63-0.70 s, spectral
share 6.6 -> 6.8e9 cell-steps/s.  Grouping the 512^2 spectral pool as well
lowered the spectral share (its long launches set the grouped pools' pace).

Trajectories: a request with `save_every = k > 0` also gets the frames of
funcutils.trajectory(repeated(step_fn, k), steps // k) -- the state after
k, 2k, ... steps.  Its pool's chunks then also stop at the request's frame
boundaries, and each frame is copied from the slot to pinned host memory on
the pool stream right behind the chunk that produced it.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Deque, Dict, List, Optional, Tuple

import numpy as np
import torch

from jax_cfd_amd import _lib
from jax_cfd_amd.base import _fvm, equations, grids, initial_conditions
from jax_cfd_amd.base import finite_differences
from jax_cfd_amd.spectral import _ns2d
from jax_cfd_amd.spectral import equations as spectral_equations

TWO_PI = 2 * np.pi


@dataclasses.dataclass
class RolloutRequest:
  request_id: int
  solver: str          # 'fvm' or 'spectral'
  n: int
  steps: int
  seed: int
  arrival: float = 0.0  # seconds after the server starts
  save_every: int = 0   # > 0: also return the state every save_every steps


@dataclasses.dataclass
class RolloutResult:
  request_id: int
  solver: str
  n: int
  steps: int
  status: str                       # 'done' or 'rejected'
  state: Optional[Tuple[np.ndarray, ...]] = None
  arrival: float = 0.0
  admitted: float = 0.0
  finished: float = 0.0
  # save_every > 0: per state component, the (steps // save_every, ...) frames
  frames: Optional[Tuple[np.ndarray, ...]] = None

  @property
  def latency(self) -> float:
    return self.finished - self.arrival


# --------------------------------------------------------------- solvers
class FvmFamily:
  """The forced FVM configuration: viscosity 1e-3, density 1, Kolmogorov
  sin(4y) forcing with linear drag 0.1, Van Leer, fast-diagonalisation
  projection, dt = stable_time_step(7, 0.5); IC filtered_velocity_field(seed,
  v_max 7, peak wavenumber 4)."""
  name = 'fvm'
  V_MAX, PEAK, NU, CFL = 7.0, 4, 1e-3, 0.5

  def __init__(self, n: int, device: torch.device):
    self.n = n
    self.device = device
    self.grid = grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))
    dt = equations.stable_time_step(self.V_MAX, self.CFL, self.NU, self.grid)
    from jax_cfd_amd.base import advection, forcings
    step = equations.semi_implicit_navier_stokes(
        density=1.0, viscosity=self.NU, dt=dt, grid=self.grid,
        convect=lambda v: tuple(advection.advect_van_leer(u, v, dt) for u in v),
        forcing=forcings.simple_turbulence_forcing(
            self.grid, constant_magnitude=1.0, constant_wavenumber=4,
            linear_coefficient=-0.1, forcing_type='kolmogorov'))
    self.spec = step.spec
    self.dt = dt

  def state_bytes(self, capacity: int) -> int:
    return 2 * capacity * self.n * self.n * 4

  def make_plan(self, capacity: int):
    return _fvm.FvmPlan(self.spec, self.device, batch=capacity)

  def alloc_state(self, capacity: int):
    n = self.n
    return (torch.zeros((capacity, n, n), dtype=torch.float32, device=self.device),
            torch.zeros((capacity, n, n), dtype=torch.float32, device=self.device))

  def initial_state(self, seed: int):
    v = initial_conditions.filtered_velocity_field(seed, self.grid, self.V_MAX,
                                                   self.PEAK, device=self.device)
    return (v[0].data, v[1].data)

  def prefetch(self, seed: int):
    """Start drawing the request's initial-field noise on a worker thread."""
    initial_conditions.prefetch_noise(seed, self.grid.shape, self.grid.ndim)

  def write_slot(self, state, slot, init):
    state[0][slot].copy_(init[0])
    state[1][slot].copy_(init[1])

  def advance(self, plan, state, slots, nsteps):
    plan.step_slots(state[0], state[1], slots, nsteps)

  GROUP_MAX_N = 512  # largest grid jcfd_fvm_group_step runs

  def make_group(self):
    return _fvm.FvmGroup(self.device)

  @staticmethod
  def group_member(plan, state, slots):
    return (plan, state[0], state[1], slots)

  def read_slot(self, state, slot):
    """Views of the slot's state (valid until the slot is reused)."""
    return (state[0][slot], state[1][slot])


class SpectralFamily:
  """ForcedNavierStokes2D(viscosity 1e-3, smooth) with crank_nicolson_rk4,
  dt 1e-3; IC rfft2(curl_2d(filtered_velocity_field(seed, v_max 1, peak 4)))
  (v_max 1 is the reference's default for filtered_velocity_field)."""
  name = 'spectral'
  V_MAX, PEAK, NU, DT = 1.0, 4, 1e-3, 1e-3

  def __init__(self, n: int, device: torch.device):
    self.n = n
    self.device = device
    self.grid = grids.Grid((n, n), domain=((0, TWO_PI), (0, TWO_PI)))
    self.equation = spectral_equations.ForcedNavierStokes2D(self.NU, self.grid, True)
    self.spec = self.equation.spec(self.DT)
    self._fft = None

  def state_bytes(self, capacity: int) -> int:
    return capacity * self.n * (self.n // 2 + 1) * 8

  def make_plan(self, capacity: int):
    return _ns2d.Ns2dPlan(self.spec, capacity, self.device)

  def alloc_state(self, capacity: int):
    n = self.n
    return (torch.zeros((capacity, n, n // 2 + 1), dtype=torch.complex64,
                        device=self.device),)

  def initial_state(self, seed: int):
    v = initial_conditions.filtered_velocity_field(seed, self.grid, self.V_MAX,
                                                   self.PEAK, device=self.device)
    w = finite_differences.curl_2d(v).data
    if self._fft is None:
      self._fft = _ns2d.Ns2dPlan(self.spec, 1, self.device)
    return (self._fft.rfft2(w),)

  def prefetch(self, seed: int):
    """Start drawing the request's initial-field noise on a worker thread."""
    initial_conditions.prefetch_noise(seed, self.grid.shape, self.grid.ndim)

  def write_slot(self, state, slot, init):
    state[0][slot].copy_(init[0])

  def advance(self, plan, state, slots, nsteps):
    plan.step_slots(state[0], slots, nsteps)

  # Largest pool a server groups (jcfd_ns2d_group_step runs up to 1024^2):
  # the 512^2 pool's long launches would set every grouped spectral pool's
  # pace, so it keeps its own stream.
  GROUP_MAX_N = 256

  def make_group(self):
    return _ns2d.Ns2dGroup(self.device)

  @staticmethod
  def group_member(plan, state, slots):
    return (plan, state[0], slots)

  def read_slot(self, state, slot):
    """A view of the slot's state (valid until the slot is reused)."""
    return (state[0][slot],)


FAMILIES = {'fvm': FvmFamily, 'spectral': SpectralFamily}


# ------------------------------------------------------------------ pools
class Pool:
  """Device-resident slots of one (solver, n) with one batched plan.  With
  `tune`, the plan's tile shapes are autotuned for its capacity (radix-8
  schedule kept, so results stay bit-identical to untuned plans)."""

  def __init__(self, family, capacity: int, tune: bool = False,
               stream: Optional[torch.cuda.Stream] = None):
    self.family = family
    self.capacity = capacity
    self.plan = family.make_plan(capacity)
    if tune:
      from jax_cfd_amd import autotune
      if family.name == 'fvm':
        autotune.tune_fvm(self.plan, radices=(8,))
      else:
        autotune.tune_ns2d(self.plan, radices=(8,))
    self.state = family.alloc_state(capacity)
    # default priority; pools of a PoolGroup share the group's stream
    self.stream = stream if stream is not None else torch.cuda.Stream(device=family.device)
    # active slot list; a fixed buffer so captured graphs stay valid, filled
    # from a pinned host copy without blocking the host (a copy from pageable
    # memory would wait for the pool stream, which may be waiting for
    # initial fields being built on the server-wide stream)
    self.slot_buf = torch.zeros(capacity, dtype=torch.int32, device=family.device)
    self.slot_host = torch.zeros(capacity, dtype=torch.int32, pin_memory=True)
    # pinned staging for finished states, one region per slot, allocated
    # once with the pool, so the serving loop never page-locks host memory
    # per result; collect() copies a result out into ordinary host memory
    # before its slot is reused.
    staging = [torch.empty(tuple(d.shape), dtype=d.dtype, pin_memory=True)
               for d in self.state]
    self.staging = [tuple(b[slot] for b in staging) for slot in range(capacity)]
    self.free: List[int] = list(range(capacity))[::-1]
    # slot -> [request, remaining, admitted, pinned host frames or None]
    self.active: Dict[int, List] = {}
    self.event: Optional[torch.cuda.Event] = None
    self.chunk_done: List[int] = []
    # finished requests whose states are being copied to the host:
    # (slot, request, admitted, pinned host tensors, host frames, copy-done event)
    self.retiring: List[tuple] = []
    self.group: Optional['PoolGroup'] = None  # set by the server for grouped pools
    self.bytes = (self.plan.device_bytes + family.state_bytes(capacity)
                  + 4 * capacity)

  @property
  def busy(self) -> bool:
    return self.event is not None and not self.event.query()

  def admit(self, req: RolloutRequest, now: float, ic_stream: torch.cuda.Stream):
    """Build the request's initial field into a free slot.  All initial
    fields are built on one server-wide stream (their helper plans are
    shared); the pool stream waits for it before its next launch."""
    slot = self.free.pop()
    with torch.cuda.stream(ic_stream):
      init = self.family.initial_state(req.seed)
      self.family.write_slot(self.state, slot, init)
      ready = torch.cuda.Event()
      ready.record(ic_stream)
    self.stream.wait_event(ready)
    frames = None
    if req.save_every > 0:
      count = req.steps // req.save_every
      frames = tuple(torch.empty((count,) + tuple(d.shape), dtype=d.dtype, pin_memory=True)
                     for d in self.family.read_slot(self.state, slot))
    self.active[slot] = [req, req.steps, now, frames]
    return slot

  @staticmethod
  def _until_stop(entry) -> int:
    """Steps the slot may advance before it must stop: its end, or its
    next frame."""
    req, remaining = entry[0], entry[1]
    if req.save_every <= 0:
      return remaining
    done = req.steps - remaining
    return min(remaining, req.save_every - done % req.save_every)

  def next_chunk(self, max_chunk: int) -> int:
    """Steps every active slot can advance next: the smallest remaining
    count or frame distance, capped at max_chunk."""
    return min(max_chunk, min(self._until_stop(a) for a in self.active.values()))

  def launch(self, max_chunk: int):
    """Advance every active slot by the smallest remaining count (capped)."""
    if not self.active:
      return
    chunk = self.next_chunk(max_chunk)
    slots = sorted(self.active)
    # the previous chunk's copy out of slot_host has run: the pool is idle
    self.slot_host[:len(slots)] = torch.as_tensor(slots, dtype=torch.int32)
    with torch.cuda.stream(self.stream):
      idx = self.slot_buf[:len(slots)]
      idx.copy_(self.slot_host[:len(slots)], non_blocking=True)
      self.family.advance(self.plan, self.state, idx, chunk)
      self.copy_frames(slots, chunk)
      self.event = torch.cuda.Event()
      self.event.record(self.stream)
    self.account(slots, chunk)

  def copy_frames(self, slots, chunk: int):
    """Queue, on the current stream, the frame copies of the slots whose
    frame boundary the chunk just reached."""
    for s in slots:
      req, remaining, _, frames = self.active[s]
      done = req.steps - remaining + chunk
      if frames is not None and chunk > 0 and done % req.save_every == 0:
        t = done // req.save_every - 1
        for h, d in zip(frames, self.family.read_slot(self.state, s)):
          h[t].copy_(d, non_blocking=True)

  def account(self, slots, chunk: int):
    for s in slots:
      self.active[s][1] -= chunk
    self.chunk_done = [s for s in slots if self.active[s][1] == 0]

  def harvest(self):
    """Start copying the finished slots' states to pinned host memory on the
    pool stream.  The host does not wait: a copy queues behind whatever
    shares the stream's hardware queue, and the server keeps serving the
    other pools meanwhile (collect() hands the results out)."""
    for slot in self.chunk_done:
      req, _, admitted, frames = self.active.pop(slot)
      with torch.cuda.stream(self.stream):
        data = self.family.read_slot(self.state, slot)
        host = self.staging[slot]
        for h, d in zip(host, data):
          h.copy_(d, non_blocking=True)
        done = torch.cuda.Event()
        done.record(self.stream)
      self.retiring.append((slot, req, admitted, host, frames, done))
    self.chunk_done = []

  def collect(self, now_fn) -> List[RolloutResult]:
    """Results whose host copies have completed; their slots become free
    only now, so an admission cannot overwrite a state still being read."""
    out, keep = [], []
    for item in self.retiring:
      slot, req, admitted, host, frames, done = item
      if not done.query():
        keep.append(item)
        continue
      out.append(RolloutResult(
          req.request_id, req.solver, req.n, req.steps, 'done',
          tuple(h.numpy().copy() for h in host),
          req.arrival, admitted, now_fn(),
          tuple(f.numpy() for f in frames) if frames is not None else None))
      self.free.append(slot)
    self.retiring = keep
    return out


class PoolGroup:
  """The pools of one solver advanced together: each chunk is ONE group step
  (jcfd_fvm_group_step / jcfd_ns2d_group_step) over the active slots of
  every member pool, whatever their grid sizes -- one launch per kernel
  class instead of one launch sequence per pool.  The chunk is the
  smallest next_chunk of the member pools, so all of them stop together;
  the pools share the group's stream and completion event.  Each slot still
  gets exactly the arithmetic of a lone instance."""

  def __init__(self, family, stream: Optional[torch.cuda.Stream] = None):
    self.group = family.make_group()
    # default priority
    self.stream = stream if stream is not None else torch.cuda.Stream(device=family.device)
    self.pools: List[Pool] = []
    self.event: Optional[torch.cuda.Event] = None

  @property
  def busy(self) -> bool:
    return self.event is not None and not self.event.query()

  def launch(self, max_chunk: int) -> bool:
    live = [p for p in self.pools if p.active]
    if not live or any(p.chunk_done for p in self.pools):
      return False
    chunk = min(p.next_chunk(max_chunk) for p in live)
    slots = [sorted(p.active) for p in live]
    with torch.cuda.stream(self.stream):
      self.group.step([p.family.group_member(p.plan, p.state, s) for p, s in zip(live, slots)],
                      chunk)
      for p, s in zip(live, slots):
        p.copy_frames(s, chunk)
      self.event = torch.cuda.Event()
      self.event.record(self.stream)
    for p, s in zip(live, slots):
      p.event = self.event
      p.account(s, chunk)
    return True


DEFAULT_CAPACITY = {64: 512, 128: 256, 256: 128, 512: 64}


class RolloutServer:
  """Continuous-batching server (see the module docstring)."""

  def __init__(self, hbm_budget_bytes: float = 200e9,
               capacity: Optional[Dict[int, int]] = None, max_chunk: int = 16,
               device: Optional[torch.device] = None, tune: bool = False,
               group: Tuple[str, ...] = ('fvm', 'spectral')):
    self.tune = bool(tune)
    unknown = set(group) - set(FAMILIES)
    if unknown:
      raise ValueError(f'unknown solver(s) to group: {sorted(unknown)}')
    self.group_solvers = tuple(group)
    self.groups: Dict[str, PoolGroup] = {}
    self.device = device if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    self.budget = float(hbm_budget_bytes)
    self.capacity = dict(DEFAULT_CAPACITY if capacity is None else capacity)
    self.max_chunk = int(max_chunk)
    self.pools: Dict[Tuple[str, int], Pool] = {}
    self.families: Dict[Tuple[str, int], object] = {}
    self.reserved = 0
    self.ic_stream = torch.cuda.Stream(device=self.device)
    # the streams of the groups and of the ungrouped default-size pools,
    # taken together here rather than as pools appear between plan set-ups
    self._streams = {s: torch.cuda.Stream(device=self.device) for s in self.group_solvers}
    for s in FAMILIES:
      for n in DEFAULT_CAPACITY:
        if not (s in self.group_solvers and n <= FAMILIES[s].GROUP_MAX_N):
          self._streams[(s, n)] = torch.cuda.Stream(device=self.device)

  def _family(self, solver, n):
    key = (solver, n)
    if key not in self.families:
      if solver not in FAMILIES:
        raise ValueError(f'unknown solver {solver!r}')
      self.families[key] = FAMILIES[solver](n, self.device)
    return self.families[key]

  def _pool(self, solver, n) -> Optional[Pool]:
    """The pool for (solver, n), created if the budget allows; None if it
    must wait.  Raises BudgetExceededError if it can never fit."""
    key = (solver, n)
    if key in self.pools:
      return self.pools[key]
    fam = self._family(solver, n)
    cap = self.capacity.get(n, 16)
    # shrink the pool until it fits the remaining budget (never below 1)
    per_slot = fam.state_bytes(1) + self._scratch_per_slot(fam)
    if per_slot > self.budget:
      raise _lib.BudgetExceededError(
          f'one {solver} {n}x{n} slot needs {per_slot} bytes > budget {self.budget:.0f}')
    cap = int(max(1, min(cap, (self.budget - self.reserved) // per_slot)))
    if self.reserved + cap * per_slot > self.budget:
      return None
    grp = None
    if solver in self.group_solvers and n <= fam.GROUP_MAX_N:
      if solver not in self.groups:
        self.groups[solver] = PoolGroup(fam, self._streams.get(solver))
      grp = self.groups[solver]
    pool = Pool(fam, cap, tune=self.tune,
                stream=grp.stream if grp is not None else self._streams.get(key))
    pool.group = grp
    if grp is not None:
      grp.pools.append(pool)
    self.pools[key] = pool
    self.reserved += pool.bytes
    return pool

  @staticmethod
  def _scratch_per_slot(fam) -> int:
    n = fam.n
    if fam.name == 'fvm':
      return (2 * n * n + n * n // 2 * 2) * 4  # u*, v*, packed spectrum
    return (n * (n // 2 + 1) + 5 * n * (n // 2)) * 8  # h, 4 spectra, product

  def _release_idle_pools(self, waiting_keys):
    for key in list(self.pools):
      pool = self.pools[key]
      if not pool.active and not pool.busy and not pool.retiring and key not in waiting_keys:
        self.reserved -= pool.bytes
        if pool.group is not None:
          pool.group.pools.remove(pool)
        pool.plan.close()
        del self.pools[key]

  def run(self, requests: List[RolloutRequest], realtime: bool = True) -> List[RolloutResult]:
    """Serve `requests` (arrival times relative to now); returns results in
    completion order.  realtime=False admits everything at t = 0.  Frames
    (save_every > 0) are held in pinned host memory, (steps // save_every)
    states per request."""
    for req in requests:
      if req.steps < 0 or req.save_every < 0:
        raise ValueError(f'request {req.request_id}: steps and save_every must be >= 0')
    pending: Deque[RolloutRequest] = collections.deque(
        sorted(requests, key=lambda r: (r.arrival, r.request_id)))
    waiting: Deque[RolloutRequest] = collections.deque()
    results: List[RolloutResult] = []
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    while pending or waiting or any(p.active or p.retiring for p in self.pools.values()):
      t = now()
      while pending and (not realtime or pending[0].arrival <= t):
        req = pending.popleft()
        if req.solver in FAMILIES:
          self._family(req.solver, req.n).prefetch(req.seed)  # overlaps the loop
        waiting.append(req)
      progressed = False
      # 1) hand out copied results, retire finished chunks, admit into free
      #    slots, relaunch
      for pool in list(self.pools.values()):
        done = pool.collect(now)
        if done:
          results.extend(done)
          progressed = True
        if pool.busy:
          continue
        if pool.chunk_done:
          pool.harvest()
          progressed = True
      still: Deque[RolloutRequest] = collections.deque()
      for req in waiting:
        try:
          pool = self._pool(req.solver, req.n)
        except _lib.BudgetExceededError:
          results.append(RolloutResult(req.request_id, req.solver, req.n, req.steps,
                                       'rejected', None, req.arrival, now(), now()))
          progressed = True
          continue
        if pool is None or pool.busy or not pool.free:
          still.append(req)
          continue
        pool.admit(req, now(), self.ic_stream)
        progressed = True
      waiting = still
      for pool in self.pools.values():
        if pool.group is None and not pool.busy and pool.active and not pool.chunk_done:
          pool.launch(self.max_chunk)
          progressed = True
      for grp in self.groups.values():
        if not grp.busy and grp.launch(self.max_chunk):
          progressed = True
      if waiting and not progressed:
        self._release_idle_pools({(r.solver, r.n) for r in waiting})
      if not progressed:
        if pending and realtime and not waiting and not any(
            p.active or p.retiring for p in self.pools.values()):
          time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
        else:
          time.sleep(50e-6)
    torch.cuda.synchronize(self.device)
    return results


def poisson_requests(count: int, rate: float = 200.0, seed: int = 0,
                     sizes=(64, 128, 256, 512), steps=(100, 2000),
                     solvers=('fvm', 'spectral')) -> List[RolloutRequest]:
  """Seeded Poisson arrivals with uniformly drawn solver, size and steps."""
  if count <= 0:
    return []
  rng = np.random.default_rng(seed)
  gaps = rng.exponential(1.0 / rate, count)
  arrivals = np.cumsum(gaps) - gaps[0]
  out = []
  for i in range(count):
    out.append(RolloutRequest(
        request_id=i, solver=str(rng.choice(solvers)), n=int(rng.choice(sizes)),
        steps=int(rng.integers(steps[0], steps[1] + 1)),
        seed=int(rng.integers(0, 2 ** 31 - 1)), arrival=float(arrivals[i])))
  return out


def summarize(results: List[RolloutResult], wall: float) -> dict:
  done = [r for r in results if r.status == 'done']
  lat = np.array([r.latency for r in done]) if done else np.zeros(1)
  cell_steps = float(sum(r.n * r.n * r.steps for r in done))
  return {'requests': len(results), 'done': len(done),
          'rejected': sum(r.status == 'rejected' for r in results),
          'wall_s': wall, 'cell_steps': cell_steps,
          'cell_steps_per_s': cell_steps / wall if wall > 0 else 0.0,
          'p50_latency_s': float(np.percentile(lat, 50)),
          'p99_latency_s': float(np.percentile(lat, 99))}
