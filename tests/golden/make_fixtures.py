"""Writes the golden fixtures of this directory from the oracle.

The reference cannot run here (JAX is not installed), so no vector comes
from it; these fixtures are the oracle's fp32 outputs on seeded 32x32 (and
one 64x64) inputs, committed so that (a) any later change to the oracle that
moves its results is caught (tests/test_golden_cpu.py) and (b) the GPU path
is checked against fixed vectors that do not need the oracle at run time
(tests/test_golden_gpu.py).  The oracle itself is pinned by the reference's
known-answer tests and closed forms (tests/test_oracle_*.py).

Every fixture is an .npz of plain float32 / complex64 arrays (no pickles).
Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fvm as ofvm  # noqa: E402
from oracle import grid_ops as go  # noqa: E402
from oracle import initial_conditions as oic  # noqa: E402
from oracle import spectral as osp  # noqa: E402
from oracle.grid_ops import A, Grid  # noqa: E402

TWO_PI = 2 * np.pi


def _grid(n):
  return Grid.from_domain((n, n), ((0, TWO_PI), (0, TWO_PI)))


def fixtures():
  """name -> dict of arrays."""
  out = {}
  forcing = go.Forcing(scale=1.0, wavenumber=4, linear_coefficient=-0.1)
  for n, seed, vmax, forced, convect in ((32, 3, 2.0, True, 'van_leer'),
                                         (32, 4, 1.0, False, 'van_leer_limiters'),
                                         (64, 5, 7.0, True, 'van_leer')):
    g = _grid(n)
    v0 = oic.filtered_velocity_field(seed, g, vmax, 4)
    dt = ofvm.stable_time_step(vmax, 0.5, 1e-3, g)
    step = ofvm.FVMStep(g, dt=dt, viscosity=1e-3, forcing=forcing if forced else None,
                        convect=convect)
    v1 = step(v0)
    v5 = step.rollout(v0, 5)
    k = step.explicit_terms(v0)
    p = ofvm.projection(tuple(A(u.data + 0.1 * kk.data, u.offset) for u, kk in zip(v0, k)), g)
    out[f'fvm_{convect}_{n}_{"forced" if forced else "unforced"}'] = {
        'n': np.int32(n), 'dt': np.float64(dt), 'forced': np.int32(forced),
        'u0': v0[0].data, 'v0': v0[1].data,
        'explicit_u': k[0].data, 'explicit_v': k[1].data,
        'projected_u': p[0].data, 'projected_v': p[1].data,
        'u1': v1[0].data, 'v1': v1[1].data, 'u5': v5[0].data, 'v5': v5[1].data}
  for n, seed, forced in ((32, 6, True), (64, 7, False)):
    g = _grid(n)
    v = oic.filtered_velocity_field(seed, g, 1.0, 4)
    w0 = oic.vorticity_hat_from_velocity(v, g).astype(np.complex64)
    eq = osp.ForcedNavierStokes2D(1e-3, g, True) if forced else osp.NavierStokes2D(
        1e-3, g, drag=0.1)
    step = osp.crank_nicolson_rk4(eq, 1e-3)
    out[f'spectral_{n}_{"forced" if forced else "drag"}'] = {
        'n': np.int32(n), 'forced': np.int32(forced), 'w0': w0,
        'explicit': np.asarray(eq.explicit_terms(w0), np.complex64),
        'w1': np.asarray(step(w0), np.complex64),
        'w10': np.asarray(osp.rollout(step, w0, 10), np.complex64)}
  g = _grid(32)
  v = oic.filtered_velocity_field(8, g, 2.0, 4)
  out['initial_velocity_32'] = {'seed': np.int32(8), 'vmax': np.float64(2.0),
                                'peak': np.int32(4), 'u': v[0].data, 'v': v[1].data}
  return out


def main():
  for name, arrays in fixtures().items():
    path = os.path.join(HERE, name + '.npz')
    np.savez(path, **arrays)
    print('wrote', path, sum(a.nbytes for a in arrays.values()), 'bytes')


if __name__ == '__main__':
  main()
