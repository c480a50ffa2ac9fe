"""empm_amd: the differentiable MLS-MPM substep of rai-opensource/empm and its
adjoint as hand-written HIP kernels for MI355X (gfx950).

Layout:
  csrc/              HIP kernels and the C-ABI implementation (libempm.so)
  _lib.py            ctypes binding of include/empm.h
  runtime.py         device context, particle states, checkpointed rollouts
  scenes.py          seeded synthetic workloads
  diff_simulator/    drop-in mirror of phys_sim/diff_simulator (WarpMPMWrapper,
                     MPMWARPDiff, state/model structs, autograd Functions)
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PACKAGE_DIR, "libempm.so")

__all__ = ["PACKAGE_DIR", "LIB_PATH"]
