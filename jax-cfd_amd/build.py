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
SOURCES = ["capi.hip", "fvm.hip", "spectral.hip", "cg.hip", "adjoint.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wno-unused-result",
    # no packed-fp32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
  srcs = [os.path.join(CSRC, s) for s in SOURCES]
  tmp = OUT + ".tmp"
  cmd = [HIPCC, *FLAGS, "-I", INCLUDE, "-I", CSRC, "-o", tmp, *srcs]
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
  if verbose and res.stderr.strip():
    sys.stderr.write(res.stderr)
  os.replace(tmp, OUT)
  return OUT


if __name__ == "__main__":
  print(build(verbose=True))
