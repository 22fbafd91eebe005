"""Host-side pieces of bench.py and the oracle's threaded FFT switch."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fftlib_threads_match_numpy():
  from oracle import fftlib
  rng = np.random.default_rng(0)
  x = rng.standard_normal((64, 64)).astype(np.float32)
  want = np.fft.rfftn(x)
  try:
    fftlib.set_workers(4)
    got = fftlib.rfftn(x)
    back = fftlib.irfftn(got, s=x.shape, axes=(0, 1))
  finally:
    fftlib.set_workers(1)
  assert got.dtype == want.dtype
  assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
  assert np.abs(back - x).max() <= 1e-5
  assert fftlib.workers() == 1


def test_trace_summary_groups_by_launch(tmp_path):
  import bench
  d = tmp_path / "run"
  d.mkdir()
  rows = [
      ("void jcfd::(anonymous namespace)::ns_rows<512, 2, 8>(jcfd::NsParams)", 100, 300, 1024),
      ("void jcfd::(anonymous namespace)::ns_rows<512, 2, 8>(jcfd::NsParams)", 400, 500, 1024),
      ("void jcfd::(anonymous namespace)::ns_rows<512, 2, 8>(jcfd::NsParams)", 600, 1000, 512),
  ]
  with open(d / "x_kernel_trace.csv", "w") as f:
    f.write("Kernel_Name,Start_Timestamp,End_Timestamp,Grid_Size_X,Grid_Size_Y,Grid_Size_Z\n")
    for name, a, b, g in rows:
      f.write(f"\"{name}\",{a},{b},{g},1,1\n")
  text = bench._trace_summary(str(d))
  lines = [ln for ln in text.splitlines() if ln.startswith("| `")]
  assert lines[0].startswith("| `ns_rows<512, 2, 8> grid=512` | 1 | 400 |")
  assert lines[1].startswith("| `ns_rows<512, 2, 8> grid=1024` | 2 | 150 |")


def test_bench_refuses_more_than_one_gpu():
  res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
  assert res.returncode == 2
  assert "one GPU" in res.stderr
