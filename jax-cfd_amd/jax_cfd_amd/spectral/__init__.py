"""MI355X-native counterpart of ``jax_cfd.spectral`` for the vorticity step."""
from jax_cfd_amd.spectral import equations
from jax_cfd_amd.spectral import time_stepping
from jax_cfd_amd.spectral import utils
