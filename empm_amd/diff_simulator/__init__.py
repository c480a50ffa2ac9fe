"""Drop-in mirror of phys_sim.diff_simulator (reference
phys_sim/diff_simulator/__init__.py:1) on libempm."""
from .warp_mpm.warp_mpm_wrapper import WarpMPMWrapper

__all__ = ["WarpMPMWrapper"]
