"""Builds libjcfd_amd.so (all HIP sources, gfx950) in-tree.

Usage: python jax-cfd_amd/build.py   (or __graft_entry__.build()).
The library lands next to the Python package so it ships with the tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "jax_cfd_amd", "libjcfd_amd.so")
SOURCES = ["capi.hip", "fvm.hip", "spectral.hip", "cg.hip", "adjoint.hip", "ml.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wno-unused-result",
    # no packed-fp32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no multiply-add contraction anywhere (fixed fp32 operation order)
    "-ffp-contract=off",
]


def build(verbose: bool = False) -> str:
  """Compile every source to an object in parallel, then link the library."""
  objdir = os.path.join(HERE, "_build")
  os.makedirs(objdir, exist_ok=True)
  jobs = []
  for src in SOURCES:
    obj = os.path.join(objdir, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", "-I", INCLUDE, "-I", CSRC, "-o", obj, os.path.join(CSRC, src)]
    cmd.remove("-shared")
    if verbose:
      print(" ".join(cmd), flush=True)
    jobs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                       text=True)))
  failed = False
  for obj, proc in jobs:
    out, err = proc.communicate()
    if proc.returncode != 0:
      sys.stderr.write(out + err)
      failed = True
    elif verbose and err.strip():
      sys.stderr.write(err)
  if failed:
    raise RuntimeError("hipcc failed")
  tmp = OUT + ".tmp"
  link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in jobs]]
  res = subprocess.run(link, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError(f"link failed with exit code {res.returncode}")
  os.replace(tmp, OUT)
  return OUT


if __name__ == "__main__":
  print(build(verbose=True))
