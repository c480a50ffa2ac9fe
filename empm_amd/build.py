"""Build libempm.so for gfx950 with hipcc (in-tree, so the library travels with
the source tree).  Run: python -m empm_amd.build"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, "csrc", f) for f in ("empm_kernels.hip", "empm_lbs.hip")]
OUT = os.path.join(HERE, "libempm.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-shared",
    "-std=c++17",
    # hardware float atomics (global_atomic_add_f32) instead of CAS loops
    "-munsafe-fp-atomics",
    # FP32 division and sqrt as v_rcp / v_sqrt sequences (~2.5 ulp) instead
    # of the correctly rounded expansions: 6% off the SVD-adjoint kernel; the
    # parity bars (1e-4 on states, 1e-3 on gradients) hold
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # FP32 denormals flushed to zero: division and square root without the
    # denormal range scaling (7-10% fewer VALU instructions in the SVD
    # kernels); the path's values are far from the denormal range
    "-fgpu-flush-denormals-to-zero",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", OUT + ".tmp", *SRCS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}): {' '.join(cmd)}")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
