"""C-ABI checks that need no GPU: the library loads, exports every function
include/jcfd.h declares, and rejects invalid configurations before touching
the device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in
           sorted(os.listdir(os.path.join(ROOT, "include"))) if h.endswith(".h")]


def _declared_functions():
  names = set()
  for h in HEADERS:
    text = open(h).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names.update(re.findall(r"\b(jcfd_\w+)\s*\(", text))
  return sorted(names)


def test_header_declares_functions():
  assert "jcfd_fvm_step" in _declared_functions()


def test_library_exports_every_declared_symbol():
  from jax_cfd_amd import _lib
  lib = _lib.load()
  missing = [n for n in _declared_functions() if not hasattr(lib, n)]
  assert not missing, missing
  # and the Python binding table covers them all
  assert set(_lib.exported_symbols()) == set(_declared_functions())


def test_version_and_last_error_empty():
  from jax_cfd_amd import _lib
  lib = _lib.load()
  assert b"gfx950" in lib.jcfd_version()


def _cfg(**kw):
  from jax_cfd_amd import _lib
  c = _lib.FvmConfig()
  c.n, c.dx, c.dy, c.dt, c.density = 64, 0.1, 0.1, 0.01, 1.0
  for k, v in kw.items():
    setattr(c, k, v)
  return c


@pytest.mark.parametrize("kw,code,msg", [
    (dict(n=48), -2, "grid size"),
    (dict(n=2048), -2, "grid size"),
    (dict(dx=0.0), -1, "grid steps"),
    (dict(density=0.0), -1, "density"),
    (dict(convect=7), -2, "advect_van_leer"),
    (dict(pressure=5), -2, "solve_fast_diag"),
    (dict(pressure=1, batch=4), -2, "one instance"),
])
def test_plan_create_rejects_before_device(kw, code, msg):
  from jax_cfd_amd import _lib
  lib = _lib.load()
  handle = ctypes.c_void_p()
  rc = lib.jcfd_fvm_plan_create(ctypes.byref(_cfg(**kw)), ctypes.byref(handle))
  assert rc == code
  assert msg in _lib.last_error()
  assert not handle.value


def test_error_mapping():
  from jax_cfd_amd import _lib
  with pytest.raises(NotImplementedError):
    _lib.check(-2, "x")
  with pytest.raises(ValueError):
    _lib.check(-1, "x")
  with pytest.raises(_lib.BudgetExceededError):
    _lib.check(-5, "x")
