"""Mirror of phys_sim/diff_simulator/warp_mpm on libempm."""
