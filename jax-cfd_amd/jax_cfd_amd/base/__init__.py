"""MI355X-native counterpart of ``jax_cfd.base`` for the FVM hot path."""
from jax_cfd_amd.base import advection
from jax_cfd_amd.base import boundaries
from jax_cfd_amd.base import diffusion
from jax_cfd_amd.base import equations
from jax_cfd_amd.base import finite_differences
from jax_cfd_amd.base import forcings
from jax_cfd_amd.base import funcutils
from jax_cfd_amd.base import grids
from jax_cfd_amd.base import initial_conditions
from jax_cfd_amd.base import pressure
from jax_cfd_amd.base import time_stepping
