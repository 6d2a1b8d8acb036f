"""Build recipe of liblolahip.so (gfx950).

  lh_device.hip  -> hipcc --offload-arch=gfx950 (kernels + device layer)
  lh_phantom.hip -> hipcc (synthetic projections)
  lh_unet.hip    -> hipcc (3D U-Net: MFMA convolutions, backward, Adam)
  lh_learned.hip -> hipcc (patch gather, forward projector, residual)
  lh_metric.hip  -> hipcc (the -c HU metric on the GPU)
  lh_roof.hip    -> hipcc (gather roof microbenchmark)
  lh_module.c    -> gcc -fopenmp (RabbitCT module, clip files, CPU ranges)
  lh_hosthalf.c  -> gcc -fopenmp (half-precision host views)
  link           -> hipcc -shared, in-tree: rabbitct_amd/liblolahip.so
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblolahip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are switched off for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-I" + INCLUDE,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall"]
HIP_SOURCES = ["lh_device.hip", "lh_phantom.hip", "lh_unet.hip", "lh_learned.hip",
               "lh_metric.hip", "lh_roof.hip"]
# Per-source flags.  The backprojection kernels branch only on wave-uniform
# values inside their view loop; left unstructurized, those branches keep the
# accumulators in place across the tap paths (structurized, each path's join
# carried undefined values that forced a register copy of every accumulator
# per view).
HIP_EXTRA = {"lh_device.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions"]}
C_SOURCES = ["lh_module.c", "lh_hosthalf.c"]
C_FLAGS = ["-O3", "-fPIC", "-std=c11", "-fopenmp", "-Wall", "-Wextra", "-I" + INCLUDE]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return p


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, *HIP_EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in C_SOURCES:
        o = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp, *objs,
          "-Wl,-soname,liblolahip.so"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
