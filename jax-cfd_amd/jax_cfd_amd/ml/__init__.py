"""MI355X-native counterpart of ``jax_cfd.ml`` for the learned-interpolation
FVM step: FusedLearnedInterpolation (its CNN tower on the matrix cores, its
polynomially constrained stencils folded into the tower's output layer) under
modular_self_advection / self_advection in modular_navier_stokes_model."""
from jax_cfd_amd.ml import advections
from jax_cfd_amd.ml import equations
from jax_cfd_amd.ml import interpolations
from jax_cfd_amd.ml import layers
from jax_cfd_amd.ml import layers_util
from jax_cfd_amd.ml import physics_specifications
from jax_cfd_amd.ml import pressures
from jax_cfd_amd.ml import towers
