"""The MPM-facing caller step of the reference's engine (phys_sim/engine):
the trainer's simulator construction and its CMA-ES evaluation, with a
batched evaluation of a whole generation."""
from .trainer_warp_mpm import InvPhyTrainerWarpMPM  # noqa: F401
