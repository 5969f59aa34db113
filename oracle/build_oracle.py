"""Build recipe for the CPU oracle (test infrastructure).

* ``oracle/liboracle.so`` -- the C restatement in ``hlb_oracle.c`` (OpenMP over
  sites in the step); always built.
* ``oracle/_ref/libhlbref.so`` -- ``ref_glue.cpp`` compiled together with the
  reference's own ``Code/util/Bessel.cc``, ``Code/util/UnitConverter.cc``, ``Code/util/Matrix3D.cc`` and
  ``Code/lb/kernels/DHumieresD3Q{15,19}MRTBasis.cc`` where they lie under
  ``/root/reference`` (same flags as the reference's default x86_64 build:
  ``-std=c++20 -msse3 -DHEMELB_USE_SSE3``).  Built only when that tree exists;
  nothing from it is copied into this repository.

Run ``python oracle/build_oracle.py``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference"
REF_CODE = os.path.join(REF_ROOT, "Code")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_oracle() -> str:
    out = os.path.join(HERE, "liboracle.so")
    _run(["gcc", "-std=c11", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
          "-fno-fast-math", "-fopenmp", "-Wall", "-o", out, os.path.join(HERE, "hlb_oracle.c"),
          "-lm"])
    return out


def build_ref() -> str | None:
    if not os.path.isdir(REF_CODE):
        return None
    outdir = os.path.join(HERE, "_ref")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libhlbref.so")
    srcs = [os.path.join(HERE, "ref_glue.cpp")] + [
        os.path.join(REF_CODE, p) for p in (
            "util/Bessel.cc", "util/UnitConverter.cc", "util/Matrix3D.cc",
            "lb/kernels/DHumieresD3Q15MRTBasis.cc",
            "lb/kernels/DHumieresD3Q19MRTBasis.cc")]
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-msse3",
          "-DHEMELB_USE_SSE3", "-I", REF_CODE, "-o", out] + srcs)
    return out


def main() -> int:
    build_oracle()
    build_ref()
    return 0


if __name__ == "__main__":
    sys.exit(main())
