"""Builds ``hemelb_amd/libhemelb_amd.so`` in-tree.

* ``csrc/lbm_kernels.hip`` and ``csrc/unet_kernels.hip`` -> hipcc for gfx950
  (device code; no FMA contraction so binary64 results match the
  reference's evaluation order).
* ``csrc/solver.cpp``, ``csrc/geometry.cpp``, ``csrc/extraction.cpp`` and
  ``csrc/unet.cpp`` -> g++ (host code; the same compiler runtime and libm as the reference's own
  g++ build).
* linked with g++ against libamdhip64 and zlib.
* ``driver/hemelb_driver.cpp`` -> ``hemelb_amd/hemelb_driver``: the C++
  driver written against the header-only layer ``include/hemelb_amd.hpp``.

Usage: ``python -m hemelb_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LIB = os.path.join(PKG, "libhemelb_amd.so")
DRIVER = os.path.join(PKG, "hemelb_driver")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]
GXX_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
    "-I", INCLUDE, "-I", CSRC,
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def build() -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for hip in ("lbm_kernels.hip", "unet_kernels.hip"):
        k_obj = os.path.join(BUILD, hip.replace(".hip", ".o"))
        _run([os.path.join(ROCM, "bin", "hipcc")] + HIPCC_FLAGS +
             ["-c", os.path.join(CSRC, hip), "-o", k_obj])
        objs.append(k_obj)
    for src in ("solver.cpp", "geometry.cpp", "extraction.cpp", "unet.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++"] + GXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs +
         ["-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-lz",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    # the C++ driver written against include/hemelb_amd.hpp
    tmpd = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE,
          os.path.join(PKG, "driver", "hemelb_driver.cpp"), "-o", tmpd, "-L", PKG,
          "-lhemelb_amd", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath-link," + os.path.join(ROCM, "lib")])
    os.replace(tmpd, DRIVER)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
