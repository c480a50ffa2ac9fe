"""Writes the golden rollout fixtures of tests/golden/ with the FP64 CPU
oracle.  The reference ships no golden vectors for this path and its Warp
dependency is absent, so these fixtures are oracle-generated: they pin the
HIP kernels (and later refactors of either side) to the oracle, which is
itself pinned by tests/test_oracle.py.

One file per material id 0-7: a 128-particle ellipsoid with velocity and
spin on a 16^3 grid inside 3-cell sticky walls, 20 substeps of dt = 1e-4,
all inputs as float32 (what the GPU path consumes), outputs x, v, C, F_trial
after the last substep as float64.

Run from the repository root:  python -m tests.golden.make_golden
"""
from __future__ import annotations

import os

import numpy as np

from empm_amd import scenes
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = 20
N_PARTICLES = 128


def make(material: int) -> dict:
    sc = scenes.small_scene(n_particles=N_PARTICLES, n_grid=16, material=material, seed=100 + material, E=3e4,
                            ys=200.0, dt=1e-4)
    N = sc.n
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    E, nu = sc.E.astype(np.float64), sc.nu.astype(np.float64)
    mu = f32(E / (2 * (1 + nu)))
    lam = f32(E * nu / ((1 + nu) * (1 - 2 * nu)))
    mass = f32(sc.vol * sc.rho)
    v0 = f32(sc.params["v0"])
    P = orc.make_params(N, sc.n_grid, material=material, gravity=sc.gravity, friction_angle=sc.friction_angle)
    cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
    st = orc.OracleState(sc.x, v0, np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam, sc.yield_stress,
                         mass, sc.vol, precision="f64")
    orc.rollout(P, cols, 0.0, sc.dt, STEPS, st)
    return dict(material=np.int32(material), n_grid=np.int32(sc.n_grid), dt=np.float64(sc.dt),
                steps=np.int32(STEPS), gravity=np.asarray(sc.gravity, np.float32), alpha=np.float32(P.alpha),
                col_point=np.asarray([c["point"] for c in sc.colliders], np.float32),
                col_normal=np.asarray([c["normal"] for c in sc.colliders], np.float32),
                x0=f32(sc.x), v0=v0, mu=mu, lam=lam, ys=f32(sc.yield_stress), mass=mass, vol=f32(sc.vol),
                x=st.x.copy(), v=st.v.copy(), C=st.C.copy(), F=st.F.copy(),
                ys_out=np.asarray(st.ys, np.float64).copy())


def main():
    orc.set_threads(1)
    for m in range(8):
        d = make(m)
        path = os.path.join(HERE, f"rollout_material{m}.npz")
        np.savez_compressed(path, **d)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
