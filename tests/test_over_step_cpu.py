"""CPU check of the kernels' x / h quotient (jcfd_internal.h over_step):
q = RN(x r), RN(q + RN(x - q h) r) with r = RN(1/h), evaluated here with the
C library's correctly rounded fmaf, against IEEE float32 division x / h.

The claim under test: bitwise equal for every |x| >= 2^-100.  Below that
the remainder x - q h can fall under the subnormal grid (it is a multiple of
about 2^(e_x - 47)), the correction is lost and the result is RN(x r), at
most one ulp of the quotient off; x = -0 gives +0.  Velocity differences of
a flow never come that close to zero without being exactly zero.  The grid
steps are the ones the plans use: 2 pi / n and pi / n (a domain of length
pi), 1 / n and 1."""
import ctypes
import ctypes.util

import numpy as np
import pytest

_libm = ctypes.CDLL(ctypes.util.find_library('m') or 'libm.so.6')
_libm.fmaf.restype = ctypes.c_float
_libm.fmaf.argtypes = (ctypes.c_float, ctypes.c_float, ctypes.c_float)
_fmaf = np.frompyfunc(lambda a, b, c: _libm.fmaf(a, b, c), 3, 1)

SAFE = np.float32(2.0 ** -100)  # |x| from which the remainder is exact


def _over_step(x, h, r):
  q = (x * r).astype(np.float32)
  rem = _fmaf(-q, h, x).astype(np.float32)
  return _fmaf(rem, r, q).astype(np.float32)


def _steps():
  out = []
  for n in (32, 64, 128, 256, 512, 1024):
    out += [np.float32(2 * np.pi / n), np.float32(np.pi / n), np.float32(1.0 / n)]
  out.append(np.float32(1.0))
  return out


def _samples(h, rng):
  xs = []
  # random values over 60 binades of quotient, both signs
  e = rng.uniform(-40, 20, 4000)
  xs.append((rng.choice([-1, 1], 4000) * 2.0 ** e).astype(np.float32))
  # quotients next to a rounding midpoint: x = RN(h * (q0 + ulp(q0) / 2))
  q0 = (rng.uniform(1, 2, 2000) * 2.0 ** rng.integers(-30, 30, 2000)).astype(np.float32)
  mid = (q0.astype(np.float64) + np.spacing(q0).astype(np.float64) / 2)
  xs.append((mid * np.float64(h)).astype(np.float32))
  # subnormal and near-underflow x, zeros, exact multiples of h
  xs.append((rng.integers(1, 2 ** 23, 1000) * 2.0 ** -149).astype(np.float32))
  xs.append((rng.uniform(1, 4, 500) * 2.0 ** -126).astype(np.float32))
  xs.append(np.array([0.0, -0.0], np.float32))
  xs.append((np.arange(-50, 51) * np.float64(h)).astype(np.float32))
  return np.concatenate(xs)


@pytest.mark.parametrize("h", _steps())
def test_over_step_equals_ieee_division(h):
  rng = np.random.default_rng(int(h.view(np.uint32)))
  r = np.float32(1.0) / h
  x = _samples(h, rng)
  with np.errstate(under='ignore'):
    exact = (x / h).astype(np.float32)
    got = _over_step(x, h, r)
  safe = np.abs(x) >= SAFE
  bad = safe & (got.view(np.uint32) != exact.view(np.uint32))
  assert not bad.any(), (h, x[bad][:5], got[bad][:5], exact[bad][:5])
  # the listed exception: tiny x, within one ulp of the quotient
  tiny = ~safe & (x != 0)
  diff = np.abs(got[tiny].astype(np.float64) - exact[tiny].astype(np.float64))
  assert (diff <= np.spacing(np.abs(exact[tiny])).astype(np.float64)).all()
  n_off = int((diff > 0).sum())
  print(f"h={h:.9g}: {int(safe.sum())} values bitwise equal; "
        f"{n_off} of {int(tiny.sum())} values below 2^-100 one ulp off")
  # zeros: +0 -> +0 and -0 -> +0 (IEEE gives -0)
  z = _over_step(np.array([0.0, -0.0], np.float32), h, r)
  assert (z == 0).all() and not np.signbit(z).any()
