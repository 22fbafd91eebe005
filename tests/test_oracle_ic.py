"""Pins the oracle's filtered_velocity_field (no GPU needed)."""
import numpy as np
import pytest

from oracle import grid_ops as go
from oracle import initial_conditions as ic
from oracle.grid_ops import Grid


@pytest.mark.parametrize("n,vmax,peak", [(64, 2.0, 4), (128, 7.0, 4)])
def test_filtered_velocity_field_properties(n, vmax, peak):
  """initial_conditions_test.py:39-60: divergence free, exact max speed."""
  grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  v = ic.filtered_velocity_field(0, grid, vmax, peak)
  assert all(u.data.dtype == np.float32 for u in v)
  div = go.divergence(v, grid).data
  assert np.abs(div).max() < 5e-4 * vmax / 2
  assert ic.max_speed(v) == pytest.approx(vmax, abs=1e-5 * vmax)


def test_noise_is_seeded_and_distinct():
  a = ic.noise(1, (8, 8))
  b = ic.noise(1, (8, 8))
  c = ic.noise(2, (8, 8))
  np.testing.assert_array_equal(a, b)
  assert not np.array_equal(a, c)
  assert not np.array_equal(a[0], a[1])
  assert a.dtype == np.float32


def test_spectrum_peaks_near_peak_wavenumber():
  """The log-normal filter (initial_conditions.py:60-64, 95-96) puts the
  kinetic-energy maximum near |k| = peak."""
  n, peak = 128, 4
  grid = Grid.from_domain((n, n), ((0, 2 * np.pi), (0, 2 * np.pi)))
  v = ic.filtered_velocity_field(3, grid, 1.0, peak)
  k = ic.angular_frequency_magnitude(grid, np.float64)
  e = sum(np.abs(np.fft.fftn(u.data)) ** 2 for u in v)
  shells = np.bincount(np.rint(k).astype(int).ravel(), e.ravel())
  assert abs(int(np.argmax(shells[1:20])) + 1 - peak) <= 1


def test_filter_is_zero_at_k0_and_positive_elsewhere():
  grid = Grid.from_domain((16, 16), ((0, 2 * np.pi), (0, 2 * np.pi)))
  f = ic.spectral_filter(grid, 3)
  assert f[0, 0] == 0
  assert (f.ravel()[1:] > 0).all()
