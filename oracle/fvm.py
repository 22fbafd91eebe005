"""FVM step of jax-cfd restated on the CPU (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``FVMStep`` mirrors ``semi_implicit_navier_stokes`` (jax_cfd/base/
equations.py:120-151) with ``time_stepper`` forward_euler, midpoint_rk2,
heun_rk2 or classic_rk4 (navier_stokes_rk, time_stepping.py:59-158), an ``advect_van_leer`` or
``advect_van_leer_using_limiters`` convection, explicit diffusion, optional
Kolmogorov + linear forcing, and ``projection`` (pressure.py:181-198) with
either ``solve_fast_diag`` (pressure.py:115-157 via fast_diagonalization.py:
199-266) or ``solve_cg`` (pressure.py:67-112).  ``ImplicitDiffusionStep``
mirrors ``implicit_diffusion_navier_stokes`` (equations.py:154-195) with
``diffusion.solve_fast_diag`` (diffusion.py:170-212) or ``diffusion.solve_cg``
(diffusion.py:137-167).

Pinning (tests/test_oracle_fvm.py): the 4x4 periodic Poisson KAT
``x = b - mean(b)`` (pressure_test.py:197-213); circulant pseudo-inverse vs a
dense ``numpy.linalg.lstsq`` brute force (fast_diagonalization_test.py:
58-130); projection divergence (pressure_test.py:91-110); divergence and
momentum invariants over many steps (equations_test.py:84-163); the
Taylor-Green closed form (validation_test.py:34-169); for the implicit
diffusion solve, constant fields unchanged (diffusion_test.py:58-76), the
residual of (1 - nu dt L) x = b (diffusion_test.py:31-42), a dense
``numpy.linalg.solve`` brute force and the single-Fourier-mode closed form.
The Runge-Kutta driver is pinned by the reference's harmonic-oscillator test
(time_stepping_test.py:27-90, its per-stepper tolerances).

``cg`` restates jax.scipy.sparse.linalg.cg, the third-party solver that both
``pressure.solve_cg`` (pressure.py:105-111) and ``diffusion.solve_cg``
(diffusion.py:156-162) call (jax is unpinned in setup.py:18 and absent here;
the algorithm is the public unpreconditioned CG recurrence with SciPy's
non-legacy stopping rule, as JAX publishes it in
jax/_src/scipy/sparse/linalg.py `_cg_solve`).  It is pinned by what the
mathematics fixes (tests/test_oracle_fvm.py): in fp64, every iterate x_k
equals the Galerkin solution over the Krylov space K_k(A, r0) computed
independently (orthonormal Krylov basis, projected solve), for a random SPD
matrix, a negative-definite periodic Laplacian and the Helmholtz operator of
the diffusion solve from x0 = b; a right-hand side that is one eigenvector
converges in exactly one iteration with x_1 = (r.r / r.Ar) r; one with k
distinct eigenvalues in exactly k; and the stopping rule -- continue while
|r_k|^2 > max(tol^2 |b|^2, atol^2) and k < maxiter -- is pinned by placing
tol, then atol, strictly between consecutive Galerkin residuals |r_{k-1}|^2
and |r_k|^2 and requiring exactly k iterations, by ties (|r_k|^2 equal to
the bound stops), and by maxiter.  The fp32 iterates are the same recurrence
in float32; their rounding (the order of the dot-product sums, which XLA
chooses) is the part only JAX could pin.
"""
from __future__ import annotations

import dataclasses
import functools
from typing import Callable, Optional, Sequence, Tuple

import numpy as np

from oracle import fftlib
from oracle import grid_ops as go
from oracle.grid_ops import A, Grid


# ------------------------------------------------------------ pressure solve
def laplacian_eigenvalues(grid: Grid):
  """Eigenvalues of the periodic 1-D Laplacians, fp64.

  array_utils.py:168-173 builds the circulant with first column
  [-2, 1, 0, ..., 0, 1] / h^2; fast_diagonalization.py:211-214 takes
  ``fft`` of it for every axis but the last and ``rfft`` for the last.
  """
  out = []
  for axis, (n, h) in enumerate(zip(grid.shape, grid.step)):
    column = np.zeros(n)
    column[0] = -2 / h ** 2
    column[1] = column[-1] = 1 / h ** 2
    if axis == grid.ndim - 1:
      out.append(np.fft.rfft(column))
    else:
      out.append(np.fft.fft(column))
  return out


def poisson_pinv_diagonal(grid: Grid, dtype=np.float32):
  """Lambda^+ of fast_diagonalization.pseudoinverse (fast_diagonalization.py:228-266).

  cutoff = 10 * eps(dtype); values are formed in fp64/complex128 and cast to
  the complex type of `dtype` (jnp.asarray with x64 disabled).
  """
  eig = laplacian_eigenvalues(grid)
  summed = functools.reduce(np.add.outer, eig)
  cutoff = 10 * np.finfo(dtype).eps
  with np.errstate(divide='ignore', invalid='ignore'):
    diag = np.where(abs(summed) > cutoff, 1 / summed, 0)
  ctype = np.complex64 if np.dtype(dtype) == np.float32 else np.complex128
  return diag.astype(ctype)


def solve_fast_diag(v: Sequence[A], grid: Grid) -> A:
  """pressure.py:115-157 (periodic): q = irfftn(Lambda^+ * rfftn(div v))."""
  rhs = go.divergence(v, grid)
  dtype = rhs.data.dtype.type
  diag = poisson_pinv_diagonal(grid, dtype)
  q = fftlib.irfftn(diag * fftlib.rfftn(rhs.data), s=rhs.data.shape,
                    axes=tuple(range(rhs.data.ndim)))
  return A(q.real.astype(dtype), rhs.offset)


def _periodic_laplacian(x: np.ndarray, grid: Grid) -> np.ndarray:
  return go.laplacian(A(x, grid.cell_center), grid).data


def _vdot(a: np.ndarray, b: np.ndarray, dtype) -> np.ndarray:
  """_vdot_real_tree of one real leaf: sum(a * b) in the leaves' dtype."""
  return np.vdot(a, b).real.astype(dtype)


def cg(op: Callable[[np.ndarray], np.ndarray], b: np.ndarray, x0: np.ndarray,
       tol: float, atol: float, maxiter: Optional[int] = None,
       callback: Optional[Callable[[int, np.ndarray, np.ndarray], None]] = None):
  """jax.scipy.sparse.linalg.cg(op, b, x0, tol=tol, atol=atol, maxiter=maxiter)
  with the identity preconditioner, restated; returns (x, iterations).

  bs = <b, b>; bound = max(tol^2 bs, atol^2); r0 = b - op(x0), p0 = r0,
  gamma0 = <r0, r0>; while gamma > bound and k < maxiter:
  alpha = gamma / <p, Ap>, x += alpha p, r -= alpha Ap, gamma' = <r, r>,
  p = r + (gamma' / gamma) p.  maxiter defaults to 10 * b.size.  Scalars
  and vectors stay in b's dtype (tol, atol are weakly typed Python floats).
  `callback(k, x_k, r_k)` sees every iterate, k = 0 included."""
  dtype = b.dtype.type
  if maxiter is None:
    maxiter = 10 * b.size
  bs = _vdot(b, b, dtype)
  bound = max(np.square(dtype(tol)) * bs, np.square(dtype(atol)))
  x = x0.astype(dtype)
  r = (b - op(x)).astype(dtype)
  p = r
  gamma = _vdot(r, r, dtype)
  k = 0
  if callback is not None:
    callback(0, x, r)
  while gamma > bound and k < maxiter:
    ap = op(p).astype(dtype)
    alpha = gamma / _vdot(p, ap, dtype)
    x = (x + alpha * p).astype(dtype)
    r = (r - alpha * ap).astype(dtype)
    gamma_new = _vdot(r, r, dtype)
    beta = gamma_new / gamma
    p = (r + beta * p).astype(dtype)
    gamma = gamma_new
    k += 1
    if callback is not None:
      callback(k, x, r)
  return x, k


def solve_cg(v: Sequence[A], grid: Grid, rtol: float = 1e-6,
             atol: float = 1e-6, maxiter: Optional[int] = None,
             return_iterations: bool = False):
  """pressure.py:67-112: cg on the periodic Laplacian (laplacian_with_bcs,
  periodic impose_bc a no-op) with b = div v and x0 = q0 = 0
  (pressure.py:189-192)."""
  rhs = go.divergence(v, grid)
  b = rhs.data
  x, k = cg(lambda x: _periodic_laplacian(x, grid), b, np.zeros_like(b), rtol, atol,
            maxiter)
  out = A(x.astype(b.dtype), rhs.offset)
  return (out, k) if return_iterations else out


def projection(v: Sequence[A], grid: Grid,
               solve: Callable = solve_fast_diag) -> Tuple[A, ...]:
  """pressure.py:181-198: v - forward_difference(q)."""
  q = solve(v, grid)
  out = []
  for axis, u in enumerate(v):
    qg = go.forward_difference(q, axis, grid.step[axis])
    out.append(A(u.data - qg.data, u.offset))
  return tuple(out)


# --------------------------------------------------------------- time step
def stable_time_step(max_velocity: float, max_courant_number: float,
                     viscosity: float, grid: Grid,
                     implicit_diffusion: bool = False) -> float:
  """equations.py:45-59."""
  dt = go.stable_time_step_advection(max_velocity, max_courant_number, grid)
  if not implicit_diffusion:
    diffusion_dt = go.stable_time_step_diffusion(viscosity, grid)
    if diffusion_dt < dt:
      raise ValueError('stable time step for diffusion is smaller than '
                       f'the chosen timestep: {diffusion_dt} vs {dt}')
  return dt


# ------------------------------------------------------- Runge-Kutta steppers
# time_stepping.py:109-158: (a, b) of forward_euler, midpoint_rk2, heun_rk2,
# classic_rk4.
TABLEAUS = {
    'forward_euler': ((), (1,)),
    'midpoint_rk2': (((1 / 2,),), (0, 1)),
    'heun_rk2': (((1,),), (1 / 2, 1 / 2)),
    'classic_rk4': (((1 / 2,), (0, 1 / 2), (0, 0, 1)), (1 / 6, 1 / 3, 1 / 3, 1 / 6)),
}


def navier_stokes_rk(a, b, explicit_terms: Callable, projection_fn: Callable,
                     dt: float) -> Callable:
  """time_stepping.py:59-106 on tuples of arrays: stage states are projected,
  sums skip zero coefficients and start from the integer 0, as the
  reference's ``sum(...)`` does."""
  def combine(u0, coefs, k):
    terms = [(c, kj) for c, kj in zip(coefs, k) if c]
    out = []
    for comp in range(len(u0)):
      acc = 0
      for c, kj in terms:
        acc = acc + c * kj[comp]
      out.append(u0[comp] + dt * acc)
    return tuple(out)

  def step(u0):
    k = [explicit_terms(u0)]
    for i in range(1, len(b)):
      k.append(explicit_terms(projection_fn(combine(u0, a[i - 1], k))))
    return projection_fn(combine(u0, b, k))
  return step


@dataclasses.dataclass
class FVMStep:
  """semi_implicit_navier_stokes + forward_euler, restated.

  convect: 'van_leer' (advect_van_leer for every component, as the
  channel-flow notebook wires it), 'van_leer_limiters' (the factory default,
  advect_van_leer_using_limiters), 'linear' (convect_linear, the
  validation tests' scheme) or 'upwind' (advect_upwind).  pressure: 'fast_diag' or 'cg'.
  """
  grid: Grid
  dt: float
  density: float = 1.0
  viscosity: Optional[float] = 1e-3
  forcing: Optional[go.Forcing] = None
  convect: str = 'van_leer'
  pressure: str = 'fast_diag'
  cg_rtol: float = 1e-6
  cg_atol: float = 1e-6
  time_stepper: str = 'forward_euler'
  # iterations of the last solve_cg call (pressure == 'cg'); a diagnostic
  cg_iterations: int = dataclasses.field(default=0, compare=False)

  def convection(self, v: Sequence[A]):
    if self.convect == 'van_leer':
      fn = go.advect_van_leer
    elif self.convect == 'van_leer_limiters':
      fn = go.advect_van_leer_using_limiters
    elif self.convect == 'linear':  # convect_linear (advection.py:184-207)
      return tuple(go.advect_linear(u, v, self.grid).data for u in v)
    elif self.convect == 'upwind':  # advect_upwind (advection.py:126-130)
      return tuple(go.advect_upwind(u, v, self.grid).data for u in v)
    else:
      raise ValueError(self.convect)
    return tuple(fn(u, v, self.dt, self.grid).data for u in v)

  def explicit_terms(self, v: Sequence[A]):
    """equations.py:102-114: (convection + diffusion) + forcing / density."""
    dv = list(self.convection(v))
    if self.viscosity is not None:
      nu = self.viscosity / self.density
      dv = [d + go.diffuse(u, nu, self.grid).data for d, u in zip(dv, v)]
    if self.forcing is not None:
      f = self.forcing(v, self.grid)
      dv = [d + fi / self.density for d, fi in zip(dv, f)]
    return tuple(A(d, u.offset) for d, u in zip(dv, v))

  def solve(self, v, grid):
    if self.pressure == 'fast_diag':
      return solve_fast_diag(v, grid)
    if self.pressure == 'cg':
      q, self.cg_iterations = solve_cg(v, grid, self.cg_rtol, self.cg_atol,
                                       return_iterations=True)
      return q
    raise ValueError(self.pressure)

  def intermediate(self, v: Sequence[A]):
    """u* = u0 + dt * (0 + 1 * k0)  (time_stepping.py:88-101, b = [1])."""
    k0 = self.explicit_terms(v)
    return tuple(A(u.data + self.dt * (0 + 1 * k.data), u.offset)
                 for u, k in zip(v, k0))

  def __call__(self, v: Sequence[A]) -> Tuple[A, ...]:
    if self.time_stepper == 'forward_euler':
      return projection(self.intermediate(v), self.grid, self.solve)
    a, b = TABLEAUS[self.time_stepper]
    offsets = [u.offset for u in v]
    wrap = lambda xs: tuple(A(x, o) for x, o in zip(xs, offsets))  # noqa: E731
    F = lambda xs: tuple(k.data for k in self.explicit_terms(wrap(xs)))  # noqa: E731
    P = lambda xs: tuple(u.data for u in projection(wrap(xs), self.grid, self.solve))  # noqa: E731
    return wrap(navier_stokes_rk(a, b, F, P, self.dt)(tuple(u.data for u in v)))

  def rollout(self, v: Sequence[A], steps: int) -> Tuple[A, ...]:
    for _ in range(steps):
      v = self(v)
    return v


# ------------------------------------------------------- implicit diffusion
def implicit_diffusion_diagonal(grid: Grid, nu: float, dt: float, dtype=np.float32):
  """The diagonal of (1 - nu dt L)^-1 - 1 in the Laplacian's Fourier basis.

  diffusion.py:180-185: func(x) = (dt nu) x / (1 - (dt nu) x), applied by
  fast_diagonalization.transform (fast_diagonalization.py:196-225, 'rfft')
  to the summed eigenvalues of the periodic 1-D Laplacians (fft per axis,
  rfft on the last), formed in fp64/complex128 and cast to the complex type
  of `dtype` (jnp.asarray with x64 disabled).
  """
  eig = laplacian_eigenvalues(grid)
  summed = functools.reduce(np.add.outer, eig)
  dt_nu_x = (dt * nu) * summed
  diag = dt_nu_x / (1 - dt_nu_x)
  ctype = np.complex64 if np.dtype(dtype) == np.float32 else np.complex128
  return diag.astype(ctype)


def solve_fast_diag_diffusion(v: Sequence[A], nu: float, dt: float,
                              grid: Grid) -> Tuple[A, ...]:
  """diffusion.solve_fast_diag (diffusion.py:170-212), periodic: each
  component becomes u + irfftn(D rfftn(u)) = (1 - nu dt L)^-1 u."""
  out = []
  for u in v:
    dtype = u.data.dtype.type
    diag = implicit_diffusion_diagonal(grid, nu, dt, dtype)
    t = fftlib.irfftn(diag * fftlib.rfftn(u.data), s=u.data.shape,
                      axes=tuple(range(u.data.ndim)))
    out.append(A(u.data + t.real.astype(dtype), u.offset))
  return tuple(out)


def helmholtz_op(grid: Grid, nu: float, dt: float, dtype):
  """linear_op of diffusion.solve_cg (diffusion.py:150-153): x - dt * nu *
  laplacian(x), the Python product dt * nu rounded to the field's dtype."""
  coef = dtype(dt * nu)

  def op(x):
    return x - coef * _periodic_laplacian(x, grid).astype(dtype)
  return op


def solve_cg_diffusion(v: Sequence[A], nu: float, dt: float, grid: Grid,
                       rtol: float = 1e-6, atol: float = 1e-6,
                       maxiter: Optional[int] = None, return_iterations: bool = False):
  """diffusion.solve_cg (diffusion.py:137-167), periodic: each component u
  becomes cg(1 - nu dt L, b = u, x0 = u) (diffusion.py:156-164)."""
  out, iters = [], []
  for u in v:
    b = u.data
    x, k = cg(helmholtz_op(grid, nu, dt, b.dtype.type), b, b, rtol, atol, maxiter)
    out.append(A(x.astype(b.dtype), u.offset))
    iters.append(k)
  return (tuple(out), iters) if return_iterations else tuple(out)


@dataclasses.dataclass
class ImplicitDiffusionStep:
  """implicit_diffusion_navier_stokes (equations.py:154-195), restated:
  dv/dt = convection (+ forcing / density) (sum_fields, equations.py:41-42),
  v <- v + dv/dt * dt, projection, then the implicit diffusion solve with
  the (kinematic) `viscosity` as given (equations.py:191-193)."""
  grid: Grid
  dt: float
  density: float = 1.0
  viscosity: float = 1e-3
  forcing: Optional[go.Forcing] = None
  convect: str = 'van_leer_limiters'  # the factory default (equations.py:165-168)
  pressure: str = 'fast_diag'
  cg_rtol: float = 1e-6
  cg_atol: float = 1e-6
  diffusion: str = 'fast_diag'  # or 'cg' (diffusion.solve_cg)
  diff_cg_rtol: float = 1e-6
  diff_cg_atol: float = 1e-6
  diff_cg_maxiter: Optional[int] = None

  def accelerations(self, v: Sequence[A]):
    conv = FVMStep(self.grid, self.dt, convect=self.convect).convection(v)
    dv = [0 + c for c in conv]
    if self.forcing is not None:
      f = self.forcing(v, self.grid)
      dv = [d + fi / self.density for d, fi in zip(dv, f)]
    return dv

  def __call__(self, v: Sequence[A]) -> Tuple[A, ...]:
    dvdt = self.accelerations(v)
    v = tuple(A(u.data + d * self.dt, u.offset) for u, d in zip(v, dvdt))
    solve = FVMStep(self.grid, self.dt, pressure=self.pressure, cg_rtol=self.cg_rtol,
                    cg_atol=self.cg_atol).solve
    v = projection(v, self.grid, solve)
    if self.diffusion == 'cg':
      return solve_cg_diffusion(v, self.viscosity, self.dt, self.grid, self.diff_cg_rtol,
                                self.diff_cg_atol, self.diff_cg_maxiter)
    return solve_fast_diag_diffusion(v, self.viscosity, self.dt, self.grid)

  def rollout(self, v: Sequence[A], steps: int) -> Tuple[A, ...]:
    for _ in range(steps):
      v = self(v)
    return v


def velocity(u: np.ndarray, v: np.ndarray, grid: Grid) -> Tuple[A, A]:
  """Wrap two arrays at the staggered cell-face offsets (grids.py:567-577)."""
  f = grid.cell_faces
  return (A(u, f[0]), A(v, f[1]))
