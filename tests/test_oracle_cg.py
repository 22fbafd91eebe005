"""Pins the oracle's restatement of jax.scipy.sparse.linalg.cg
(oracle/fvm.py `cg`), which pressure.solve_cg (pressure.py:67-112) and
diffusion.solve_cg (diffusion.py:137-167) call, by what the mathematics
fixes rather than by retyping the recurrence:

* CG's k-th iterate is the Galerkin solution over the Krylov space
  K_k(A, r0): x_k = x0 + V y with V an orthonormal basis of
  span{r0, A r0, ..., A^(k-1) r0} and (V^T A V) y = V^T r0.  Computed here
  in fp64 from dense matrices built independently of the oracle's operators
  (Kronecker sums of circulants), for a random SPD matrix, the singular
  negative-semidefinite periodic Laplacian of the pressure solve, and the
  Helmholtz operator 1 - c L of the diffusion solve from x0 = b.  A dropped
  term, a wrong sign or a swapped alpha/beta fails this at k = 2 or 3.
* A right-hand side that is one eigenvector converges in exactly one
  iteration with x_1 = (r.r / r.Ar) r = b / lambda; one spanning k distinct
  eigenvalues (any number of eigenvectors) in exactly k.
* The stopping rule, continue while |r_k|^2 > max(tol^2 |b|^2, atol^2) and
  k < maxiter: tol, then atol, placed strictly between consecutive Galerkin
  residuals |r_{k-1}|^2 and |r_k|^2 must give exactly k iterations; the
  larger of the two bounds wins; an exact tie |r_0|^2 == atol^2 stops
  (strict inequality); maxiter caps the count.
"""
import functools

import numpy as np
import pytest

from oracle import fvm
from oracle.grid_ops import Grid


def _circulant_laplacian_1d(n, h):
  d = np.zeros((n, n))
  for i in range(n):
    d[i, i] = -2.0
    d[i, (i + 1) % n] += 1.0
    d[i, (i - 1) % n] += 1.0
  return d / (h * h)


def _laplacian_matrix(nx, ny, dx, dy):
  """Dense periodic 5-point Laplacian on row-major (nx, ny) fields."""
  return (np.kron(_circulant_laplacian_1d(nx, dx), np.eye(ny)) +
          np.kron(np.eye(nx), _circulant_laplacian_1d(ny, dy)))


def _galerkin_iterates(a, b, x0, kmax):
  """x_k, k = 0..kmax, by the Galerkin condition on K_k(A, r0), fp64."""
  r0 = b - a @ x0
  basis = []
  w = r0 / np.linalg.norm(r0)
  out = [x0.copy()]
  for _ in range(kmax):
    basis.append(w)
    v = np.stack(basis, axis=1)
    y = np.linalg.solve(v.T @ a @ v, v.T @ r0)
    out.append(x0 + v @ y)
    w = a @ w
    for _ in range(2):  # Gram-Schmidt, twice for orthogonality
      w = w - v @ (v.T @ w)
    w = w / np.linalg.norm(w)
  return out


def _run_cg(op, b, x0, tol, atol, maxiter=None):
  iterates = []
  x, k = fvm.cg(op, b, x0, tol, atol, maxiter,
                callback=lambda j, xj, rj: iterates.append((xj.copy(), rj.copy())))
  return x, k, iterates


def _cases():
  rng = np.random.default_rng(0)
  m = rng.standard_normal((40, 40))
  spd = m @ m.T + 40 * np.eye(40)
  nx, ny = 8, 6
  grid = Grid.from_domain((nx, ny), ((0, 2 * np.pi), (0, 3.0)))
  lap = _laplacian_matrix(nx, ny, *grid.step)
  b_lap = rng.standard_normal(nx * ny)
  b_lap -= b_lap.mean()  # consistent right-hand side (range of L)
  coef = 0.05 * 0.3
  helm = np.eye(nx * ny) - coef * lap
  b_helm = rng.standard_normal(nx * ny)

  def lap_op(x):  # the oracle's operator on the field, flattened back
    return fvm._periodic_laplacian(x.reshape(nx, ny), grid).reshape(-1)

  helm_grid_op = fvm.helmholtz_op(grid, 0.3, 0.05, np.float64)
  return {
      'spd': (lambda x: spd @ x, spd, rng.standard_normal(40), np.zeros(40)),
      'pressure-laplacian': (lap_op, lap, b_lap, np.zeros(nx * ny)),
      'diffusion-helmholtz': (lambda x: helm_grid_op(x.reshape(nx, ny)).reshape(-1), helm,
                              b_helm, b_helm.copy()),
  }


@pytest.mark.parametrize("name", ['spd', 'pressure-laplacian', 'diffusion-helmholtz'])
def test_cg_iterates_are_krylov_galerkin_solutions(name):
  op, a, b, x0 = _cases()[name]
  # the operator the oracle applies is the independently built matrix
  probe = np.random.default_rng(1).standard_normal(b.size)
  np.testing.assert_allclose(op(probe), a @ probe, rtol=1e-12, atol=1e-9)
  kmax = 6
  _, k, iterates = _run_cg(op, b, x0, 0.0, 0.0, maxiter=kmax)
  assert k == kmax
  ref = _galerkin_iterates(a, b, x0, kmax)
  scale = np.linalg.norm(ref[kmax])  # x0 may be zero
  for j in range(kmax + 1):
    xj, rj = iterates[j]
    err = np.linalg.norm(xj - ref[j]) / scale
    assert err <= 1e-9, (name, j, err)
    np.testing.assert_allclose(rj, b - a @ ref[j], rtol=0, atol=1e-8 * np.linalg.norm(b))
  # maxiter caps the count and the returned x is the last iterate
  x3, k3, _ = _run_cg(op, b, x0, 0.0, 0.0, maxiter=3)
  assert k3 == 3 and np.array_equal(x3, iterates[3][0])


def _eigen_grid(n=16):
  grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  i = np.arange(n)[:, None]
  j = np.arange(n)[None, :]

  def lam(p, q):
    h = grid.step[0]
    return (2 * np.cos(2 * np.pi * p / n) - 2) / h ** 2 + (2 * np.cos(2 * np.pi * q / n) - 2) / h ** 2

  def mode(p, q, phase=0.0):
    return np.cos(2 * np.pi * (p * i + q * j) / n + phase) * np.ones((n, n))
  return grid, lam, mode


def test_cg_one_eigenvector_converges_in_one_iteration():
  grid, lam, mode = _eigen_grid()
  b = 3.0 * mode(2, 1)
  op = functools.partial(fvm._periodic_laplacian, grid=grid)
  x, k = fvm.cg(op, b, np.zeros_like(b), 1e-10, 0.0)
  assert k == 1
  r = b
  step = np.vdot(r, r) / np.vdot(r, op(r))
  np.testing.assert_allclose(x, step * r, rtol=1e-13, atol=1e-14)
  np.testing.assert_allclose(x, b / lam(2, 1), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_cg_k_distinct_eigenvalues_converge_in_k_iterations(k):
  grid, lam, mode = _eigen_grid()
  # eigenvalue classes (p, q): each class gets two eigenvectors of the same
  # eigenvalue (a cos and a shifted one, and the mirrored mode), which CG
  # must count once
  classes = [(1, 0), (2, 0), (3, 1), (5, 2), (7, 0)][:k]
  assert len({round(lam(p, q), 9) for p, q in classes}) == k
  rng = np.random.default_rng(k)
  b = sum(rng.uniform(0.5, 2) * mode(p, q, rng.uniform(0, 6)) +
          rng.uniform(0.5, 2) * mode(q, p, rng.uniform(0, 6)) for p, q in classes)
  op = functools.partial(fvm._periodic_laplacian, grid=grid)
  _, iters = fvm.cg(op, b, np.zeros_like(b), 1e-10, 0.0)
  assert iters == k


def _spd_residuals(kmax=10):
  op, a, b, x0 = _cases()['spd']
  ref = _galerkin_iterates(a, b, x0, kmax)
  res2 = [float(np.sum((b - a @ x) ** 2)) for x in ref]
  return op, b, x0, res2


@pytest.mark.parametrize("k", [1, 3, 6])
def test_cg_stopping_rule_relative_and_absolute(k):
  op, b, x0, res2 = _spd_residuals()
  bs = float(np.sum(b * b))
  assert all(res2[j] > res2[j + 1] for j in range(len(res2) - 1))
  between = np.sqrt(res2[k - 1] * res2[k])  # strictly inside (|r_k|^2, |r_{k-1}|^2)
  # relative bound alone
  _, it = fvm.cg(op, b, x0, np.sqrt(between / bs), 0.0)
  assert it == k
  # absolute bound alone
  _, it = fvm.cg(op, b, x0, 0.0, np.sqrt(between))
  assert it == k
  # both: the larger bound stops first
  later = np.sqrt(res2[k + 1] * res2[k + 2])
  _, it = fvm.cg(op, b, x0, np.sqrt(later / bs), np.sqrt(between))
  assert it == k
  _, it = fvm.cg(op, b, x0, np.sqrt(between / bs), np.sqrt(later))
  assert it == k
  # maxiter below the converged count wins
  _, it = fvm.cg(op, b, x0, np.sqrt(between / bs), 0.0, maxiter=k - 1)
  assert it == k - 1


def test_cg_stopping_rule_is_strict():
  # |r_0|^2 = |b|^2 = 16 exactly; atol^2 = 16 exactly: the loop must not run
  grid, _, mode = _eigen_grid(4)
  b = np.where(np.arange(16).reshape(4, 4) % 2 == 0, 1.0, -1.0)
  op = functools.partial(fvm._periodic_laplacian, grid=grid)
  _, it = fvm.cg(op, b, np.zeros_like(b), 0.0, 4.0)
  assert it == 0
  _, it = fvm.cg(op, b, np.zeros_like(b), 1.0, 0.0)  # tol^2 |b|^2 == |b|^2
  assert it == 0
  _, it = fvm.cg(op, b, np.zeros_like(b), 0.0, np.nextafter(4.0, 0.0))
  assert it >= 1
  # and in float32, the dtype the pressure solve runs in
  b32 = b.astype(np.float32)
  _, it = fvm.cg(op, b32, np.zeros_like(b32), 0.0, 4.0)
  assert it == 0


def test_solve_cg_wrappers_call_cg_with_the_reference_arguments():
  """pressure.solve_cg: b = div v, x0 = 0; diffusion.solve_cg: b = x0 = u,
  operator u - fp32(dt nu) L u.  Checked against cg called directly."""
  from oracle import grid_ops as go
  from oracle.grid_ops import A
  n = 16
  grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  rng = np.random.default_rng(3)
  v = tuple(A(rng.standard_normal((n, n)).astype(np.float32), o) for o in grid.cell_faces)
  q, it = fvm.solve_cg(v, grid, 1e-6, 1e-6, return_iterations=True)
  b = go.divergence(v, grid).data
  x, k = fvm.cg(lambda z: fvm._periodic_laplacian(z, grid), b, np.zeros_like(b), 1e-6, 1e-6)
  assert it == k and np.array_equal(q.data, x)
  (ud,), (itd,) = fvm.solve_cg_diffusion(v[:1], 0.01, 0.3, grid, 1e-6, 1e-6,
                                         return_iterations=True)
  u = v[0].data
  coef = np.float32(0.3 * 0.01)
  x, k = fvm.cg(lambda z: z - coef * fvm._periodic_laplacian(z, grid), u, u, 1e-6, 1e-6)
  assert itd == k and np.array_equal(ud.data, x)
