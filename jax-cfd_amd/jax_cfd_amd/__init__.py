"""jax_cfd_amd: jax-cfd's time-step hot path as hand-written HIP for MI355X.

``import jax_cfd_amd.base as cfd`` mirrors ``import jax_cfd.base as cfd`` for
the FVM step; ``jax_cfd_amd.spectral`` mirrors ``jax_cfd.spectral`` for the
pseudospectral vorticity step.  All compute runs in libjcfd_amd.so.
"""
__version__ = '0.1.0'
