"""FFT entry points of the oracle's time steps (test infrastructure).

numpy.fft by default, which every parity test and golden fixture uses.
`set_workers(k)` with k > 1 routes the same calls through scipy.fft with k
threads: the CPU baseline's one-instance legs use it so that the host's
cores share each transform (the reference's jnp.fft runs the XLA CPU FFT,
which threads large transforms the same way).
"""
import numpy as np

_WORKERS = 1


def set_workers(k: int) -> None:
  global _WORKERS
  _WORKERS = max(1, int(k))


def workers() -> int:
  return _WORKERS


def _lib():
  if _WORKERS > 1:
    import scipy.fft
    return scipy.fft, {"workers": _WORKERS}
  return np.fft, {}


def rfftn(x, s=None, axes=None):
  lib, kw = _lib()
  return lib.rfftn(x, s=s, axes=axes, **kw)


def irfftn(x, s=None, axes=None):
  lib, kw = _lib()
  return lib.irfftn(x, s=s, axes=axes, **kw)


def rfft2(x):
  lib, kw = _lib()
  return lib.rfft2(x, **kw)
