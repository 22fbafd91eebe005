"""Counterpart of ``jax_cfd.data`` for rollout output: trajectory datasets."""
from jax_cfd_amd.data import xarray_utils
