"""CPU restatement of jax-cfd's time-step hot path -- TEST INFRASTRUCTURE ONLY.

This package is the oracle: a NumPy restatement of the reference algorithms,
written to follow the reference's operation order and dtype behaviour (fp32 by
default, fp64 for the side-by-side accuracy runs).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline.  The product
(``jax_cfd_amd``) never imports it and has no CPU fallback.

Modules:
  grid_ops   -- offset-tracking arrays, periodic shift, interpolation, finite
                differences, advection schemes, diffusion, forcing
                (jax_cfd/base/{grids,boundaries,interpolation,finite_differences,
                advection,diffusion,forcings}.py)
  fvm        -- explicit terms, forward Euler, fast-diagonalisation and CG
                pressure solves, projection, stable time step, the FVM step
                (jax_cfd/base/{equations,time_stepping,pressure,
                fast_diagonalization,array_utils}.py)
  spectral   -- NavierStokes2D / ForcedNavierStokes2D and crank_nicolson_rk4
                (jax_cfd/spectral/{equations,time_stepping,utils}.py)
  initial_conditions -- filtered_velocity_field with this project's seeded
                generator (jax_cfd/base/{initial_conditions,filter_utils}.py)
  dft        -- O(n^2) DFT used to pin the FFT-based pieces.

Pinning status of each function is listed in the module docstrings; the tests
in tests/test_oracle_*.py hold the checks.
"""
