"""O(n^2) discrete Fourier transforms used to pin FFT-based code (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Forward transforms are unnormalised and inverse transforms carry 1/N, the
NumPy/JAX convention used by jnp.fft.rfftn/irfftn (the reference's
fast_diagonalization.py:223 and spectral/equations.py:163-180).  Computed in
fp64 from the definition, so they are independent of any FFT library.
"""
import numpy as np


def dft_matrix(n: int, sign: int = -1) -> np.ndarray:
  k = np.arange(n)
  return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def rfft2(x: np.ndarray) -> np.ndarray:
  """Half-spectrum 2-D DFT of a real (n0, n1) array: (n0, n1//2 + 1)."""
  n0, n1 = x.shape
  full = dft_matrix(n0) @ x.astype(np.float64) @ dft_matrix(n1).T
  return full[:, : n1 // 2 + 1]


def irfft2(xh: np.ndarray, n1: int) -> np.ndarray:
  """Inverse of rfft2: Hermitian-extend along the last axis, then 2-D IDFT.

  Like pocketfft's c2r, the imaginary parts of the DC and (even n1) Nyquist
  columns are ignored.
  """
  n0 = xh.shape[0]
  xh = xh.astype(np.complex128)
  # ifft along axis 0 first (complex), then c2r along axis 1.
  col = dft_matrix(n0, +1) @ xh / n0
  col[:, 0] = col[:, 0].real
  if n1 % 2 == 0:
    col[:, n1 // 2] = col[:, n1 // 2].real
  full = np.zeros((n0, n1), np.complex128)
  full[:, : n1 // 2 + 1] = col
  for k in range(1, (n1 + 1) // 2):
    full[:, n1 - k] = np.conj(col[:, k])
  return (full @ dft_matrix(n1, +1).T / n1).real
