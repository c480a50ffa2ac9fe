"""Seeded synthetic scenes for the benchmark workloads and parity tests.

Every generator uses numpy's PCG64 with an explicit seed and returns host
(numpy) arrays, so the same inputs can be fed to the GPU path and to the CPU
checker.  Axis convention: z is up, gravity (0, 0, -9.8), as the reference's
WarpMPMWrapper (warp_mpm_wrapper.py:110).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scene:
    name: str
    n_grid: int
    grid_lim: float
    dt: float
    n_steps: int
    material: int
    x: np.ndarray            # (N,3) float32
    vol: np.ndarray          # (N,) float32
    rho: np.ndarray          # (N,) float32
    E: np.ndarray            # (N,) float32
    nu: np.ndarray           # (N,) float32
    yield_stress: np.ndarray  # (N,) float32
    gravity: tuple = (0.0, 0.0, -9.8)
    friction_angle: float = 25.0
    softening: float = 0.1
    grid_v_damping_scale: float = 1.1
    rpic_damping: float = 0.0
    # colliders as reference-style dicts: {"kind": "surface"|"cuboid"|"bbox", ...}
    colliders: list = field(default_factory=list)
    params: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.x.shape[0])

    @property
    def dx(self):
        return self.grid_lim / self.n_grid


def sticky_walls(n_grid, cells=3, grid_lim=1.0):
    """'k-cell sticky ground/walls': v = 0 on nodes with any index < k or
    >= n-k, as six sticky surface colliders (strict `offset . n < 0` test,
    mpm_solver_diff.py:489-491).  The planes sit half a cell away from the
    nodes ((k-0.5) dx and (n-k-0.5) dx) so the node sets do not depend on
    float rounding."""
    dx = grid_lim / n_grid
    lo = (cells - 0.5) * dx
    hi = (n_grid - cells - 0.5) * dx
    cols = []
    for d in range(3):
        nrm = [0.0, 0.0, 0.0]
        nrm[d] = 1.0
        p = [0.0, 0.0, 0.0]
        p[d] = lo
        cols.append({"kind": "surface", "point": tuple(p), "normal": tuple(nrm)})
        nrm2 = [0.0, 0.0, 0.0]
        nrm2[d] = -1.0
        p2 = [0.0, 0.0, 0.0]
        p2[d] = hi
        cols.append({"kind": "surface", "point": tuple(p2), "normal": tuple(nrm2)})
    return cols


def jittered_cube(center, side, n_grid, ppc, rng, grid_lim=1.0):
    """ppc particles per covered cell, uniformly jittered inside the cell."""
    dx = grid_lim / n_grid
    lo = np.asarray(center) - side / 2
    hi = np.asarray(center) + side / 2
    i0 = np.floor(lo / dx + 1e-9).astype(int)
    i1 = np.ceil(hi / dx - 1e-9).astype(int)
    pts = []
    for i in range(i0[0], i1[0]):
        for j in range(i0[1], i1[1]):
            for k in range(i0[2], i1[2]):
                cell_lo = np.array([i, j, k]) * dx
                p = cell_lo + rng.random((ppc, 3)) * dx
                m = np.all((p >= lo) & (p <= hi), axis=1)
                pts.append(p[m])
    return np.concatenate(pts, axis=0)


def ellipsoid_points(center, radii, n, rng):
    out = np.zeros((0, 3))
    radii = np.asarray(radii, np.float64)
    while out.shape[0] < n:
        p = (rng.random((2 * n, 3)) * 2 - 1)
        m = np.sum(p * p, axis=1) <= 1.0
        out = np.concatenate([out, p[m] * radii + np.asarray(center)], axis=0)
    return out[:n]


def cube_c1(seed=0):
    """Plain forward, reference CPU case: 32^3 grid on the unit cube, cube of
    side 0.3 centred at (0.5, 0.4, 0.5) with 8 jittered particles per covered
    cell, V = dx^3/8, rho = 1000, FCR E=2e5 nu=0.3, g = -9.8 (z), dt=1e-4,
    200 substeps, 3-cell sticky walls."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 32
    dx = 1.0 / n
    x = jittered_cube((0.5, 0.4, 0.5), 0.3, n, 8, rng)
    N = x.shape[0]
    f = lambda v: np.full(N, v, np.float32)
    return Scene(name="cube_32^3_fcr", n_grid=n, grid_lim=1.0, dt=1e-4, n_steps=200, material=0,
                 x=x.astype(np.float32), vol=f(dx ** 3 / 8), rho=f(1000.0), E=f(2e5), nu=f(0.3),
                 yield_stress=f(0.0), colliders=sticky_walls(n, 3))


def sample_params_c2(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    logE = rng.uniform(np.log(1e4), np.log(1e6))
    nu = rng.uniform(0.2, 0.45)
    logY = rng.uniform(np.log(1e2), np.log(1e4))
    return {"E": float(np.exp(logE)), "nu": float(nu), "yield_stress": float(np.exp(logY))}


def ellipsoid_c2(n_particles=200_000, n_grid=64, n_steps=400, param_seed=3, material=7, seed=1, drop_height=0.3):
    """Single rollout forward + adjoint: 200k particles in an ellipsoid with
    radii (0.2, 0.1, 0.15) on a 64^3 grid, Neo-Hookean + von Mises
    (material 7), dt = 5e-5, 400 substeps, rho = 1000, 3-cell sticky
    ground/walls.  "Dropped from height 0.3": a 400-substep rollout lasts
    0.02 s, too short for a body released at rest 0.3 m up to reach the
    ground, so the rollout starts at the moment of impact, bottom just above
    the sticky band with the impact velocity sqrt(2 g h) downwards.
    Parameters drawn from `param_seed` (log-uniform E in [1e4,1e6], uniform
    nu in [0.2,0.45], log-uniform yield stress in [1e2,1e4])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    radii = (0.2, 0.1, 0.15)
    dx = 1.0 / n_grid
    cz = 3 * dx + radii[2] + 0.25 * dx
    x = ellipsoid_points((0.5, 0.5, cz), radii, n_particles, rng)
    N = x.shape[0]
    vol_total = 4.0 / 3.0 * np.pi * radii[0] * radii[1] * radii[2]
    prm = sample_params_c2(param_seed)
    f = lambda v: np.full(N, v, np.float32)
    sc = Scene(name=f"ellipsoid_{N}_{n_grid}^3_nh_vm", n_grid=n_grid, grid_lim=1.0, dt=5e-5, n_steps=n_steps,
               material=material, x=x.astype(np.float32), vol=f(vol_total / N), rho=f(1000.0), E=f(prm["E"]),
               nu=f(prm["nu"]), yield_stress=f(prm["yield_stress"]), colliders=sticky_walls(n_grid, 3), params=prm)
    v0 = np.zeros((N, 3), np.float32)
    v0[:, 2] = -np.sqrt(2 * 9.8 * drop_height)
    sc.params["v0"] = v0
    return sc


def small_scene(n_particles=512, n_grid=16, material=0, seed=0, E=2e4, nu=0.3, ys=300.0, center=(0.5, 0.5, 0.45),
                radii=(0.15, 0.12, 0.1), vel=(0.3, -0.2, 0.1), dt=1e-4, spin=2.0):
    """A small ellipsoid with an initial velocity and spin about z, for
    parity and finite-difference tests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = ellipsoid_points(center, radii, n_particles, rng)
    N = x.shape[0]
    vol_total = 4.0 / 3.0 * np.pi * radii[0] * radii[1] * radii[2]
    f = lambda v: np.full(N, v, np.float32)
    sc = Scene(name=f"small_{N}_{n_grid}", n_grid=n_grid, grid_lim=1.0, dt=dt, n_steps=20, material=material,
               x=x.astype(np.float32), vol=f(vol_total / N), rho=f(1000.0), E=f(E), nu=f(nu), yield_stress=f(ys),
               colliders=sticky_walls(n_grid, 3))
    c = np.asarray(center)
    w = np.array([0.0, 0.0, spin])
    v0 = np.asarray(vel)[None, :] + np.cross(w[None, :], x - c[None, :])
    sc.params["v0"] = v0.astype(np.float32)
    return sc


C3_SHAPES = ("cube", "rod", "ellipsoid")
C3_HALF_Z = {"cube": 0.1, "rod": 0.03, "ellipsoid": 0.15}
C3_NOMINAL = {"E": 1e5, "nu": 0.3, "friction_angle": 30.0, "yield_stress": 5000.0}


def shape_points(shape, n, rng, center):
    """n points uniform in a cube of side 0.2, a rod 0.4 x 0.06 x 0.06 along
    x, or an ellipsoid with radii (0.2, 0.1, 0.15) (half heights along z in
    C3_HALF_Z); returns (points, volume, half height along z)."""
    c = np.asarray(center, np.float64)
    if shape == "cube":
        ext = np.array([0.2, 0.2, 0.2])
    elif shape == "rod":
        ext = np.array([0.4, 0.06, 0.06])
    elif shape == "ellipsoid":
        radii = (0.2, 0.1, 0.15)
        return ellipsoid_points(c, radii, n, rng), 4.0 / 3.0 * np.pi * np.prod(radii), radii[2]
    else:
        raise ValueError(f"unknown shape {shape}")
    return c + (rng.random((n, 3)) - 0.5) * ext, float(np.prod(ext)), ext[2] / 2


def population_c3(n_cand=32, seed=4, n_grid=64, n_steps=300, dt=1e-4, second_material=5, nominal=False,
                  drop_height=0.3, n_range=(50_000, 150_000)):
    """A CMA-ES population: n_cand candidates, each an object of N ~ U{50k..150k}
    particles shaped as a cube (side 0.2), a rod (0.4 x 0.06 x 0.06) or an
    ellipsoid (radii 0.2, 0.1, 0.15), drawn from the seed, on its own 64^3
    grid with 3-cell sticky walls.  Materials alternate Drucker-Prager sand
    (id 2) and von Mises plasticine (id 5, yield stress 5000, softening 0.1,
    the reference's gallery preset, warp_mpm_wrapper.py:43-47), or another
    second material (second_material, e.g. 7: Neo-Hookean + von Mises with
    yield stress 5000).  Parameters are log-normal
    around the nominal E = 1e5, nu = 0.3, friction angle 30 degrees:
    E = 1e5 exp(0.3 z0), nu = min(0.3 exp(0.1 z1), 0.489),
    angle = 30 exp(0.1 z2); nominal=True gives the same objects with the
    nominal parameters (the fitness target).  As in the single-rollout
    workload, each object starts at the moment of impact after a drop of
    `drop_height`: bottom a quarter cell above the sticky band, moving down
    at sqrt(2 g h).  rho = 1000, dt = 1e-4, 300 substeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 1.0 / n_grid
    out = []
    for c in range(n_cand):
        N = int(rng.integers(n_range[0], n_range[1] + 1))
        shape = C3_SHAPES[int(rng.integers(len(C3_SHAPES)))]
        z = rng.standard_normal(3)
        material = (2, int(second_material))[c % 2]
        E = 1e5 * np.exp(0.3 * z[0])
        nu = min(0.3 * np.exp(0.1 * z[1]), 0.489)
        phi = 30.0 * np.exp(0.1 * z[2])
        if nominal:
            E, nu, phi = C3_NOMINAL["E"], C3_NOMINAL["nu"], C3_NOMINAL["friction_angle"]
        ys = C3_NOMINAL["yield_stress"]
        srng = np.random.Generator(np.random.PCG64([seed, c]))
        x, vol_total, _ = shape_points(shape, N, srng, (0.5, 0.5, 3 * dx + C3_HALF_Z[shape] + 0.25 * dx))
        f = lambda v: np.full(N, v, np.float32)
        sc = Scene(name=f"c{c}_{shape}_{N}_m{material}", n_grid=n_grid, grid_lim=1.0, dt=dt, n_steps=n_steps,
                   material=material, x=x.astype(np.float32), vol=f(vol_total / N), rho=f(1000.0), E=f(E),
                   nu=f(nu), yield_stress=f(ys), friction_angle=float(phi), softening=0.1,
                   colliders=sticky_walls(n_grid, 3),
                   params={"E": float(E), "nu": float(nu), "friction_angle": float(phi), "yield_stress": ys,
                           "shape": shape})
        v0 = np.zeros((N, 3), np.float32)
        v0[:, 2] = -np.sqrt(2 * 9.8 * drop_height)
        sc.params["v0"] = v0
        out.append(sc)
    return out


def serving_requests_c5(n_requests=60, seed=6, grids=(32, 64, 128), n_range=(10_000, 500_000),
                        steps_range=(100, 2000), grad_fraction=0.4, drop_height=0.3):
    """A serving request mix: per request, N log-uniform over n_range, a grid
    from `grids`, a horizon uniform over steps_range, forward only or (with
    probability grad_fraction) forward + gradient, a material from {jelly
    (0), Neo-Hookean + von Mises (7), sand (2)} with the single-rollout
    workload's parameter distribution (log-uniform E in [1e4, 1e6],
    uniform nu in [0.2, 0.45], log-uniform yield stress in [1e2, 1e4]), and
    an object shaped as in the population workload (cube, rod or ellipsoid)
    starting at the moment of impact on the 3-cell floor band.  Gradient
    requests carry a target 1 cm below the initial positions.  Returns dicts
    of RolloutRequest fields."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_requests):
        N = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        grid = int(grids[int(rng.integers(len(grids)))])
        T = int(rng.integers(steps_range[0], steps_range[1] + 1))
        grad = bool(rng.random() < grad_fraction)
        material = (0, 7, 2)[int(rng.integers(3))]
        prm = sample_params_c2(int(rng.integers(1 << 30)))
        shape = C3_SHAPES[int(rng.integers(len(C3_SHAPES)))]
        dx = 1.0 / grid
        srng = np.random.Generator(np.random.PCG64([seed, i]))
        x, vol_total, _ = shape_points(shape, N, srng, (0.5, 0.5, 3 * dx + C3_HALF_Z[shape] + 0.25 * dx))
        x = x.astype(np.float32)
        v = np.zeros((N, 3), np.float32)
        v[:, 2] = -np.sqrt(2 * 9.8 * drop_height)
        out.append(dict(x=x, v=v, vol=np.full(N, vol_total / N, np.float32), material=material, E=prm["E"],
                        nu=prm["nu"], yield_stress=prm["yield_stress"], n_grid=grid, n_steps=T, want_grad=grad,
                        target=(x - np.array([0.0, 0.0, 0.01], np.float32)) if grad else None, shape=shape))
    return out


def slab_c4(n_particles=300_000, n_grid=96, seed=5, dims=(0.5, 0.05, 0.2), n_frames=30, frame_steps=100, dt=5e-5,
            vmax=0.5, gripper_half=(0.05, 0.05, 0.03)):
    """Long-horizon embodied workload: a slab of dims (x, y, z) = 0.5 x 0.05 x
    0.2 (standing on the floor, z up) of uniformly sampled particles on an
    n_grid^3 grid with 3-cell sticky walls, Neo-Hookean + von Mises
    (material 7), rho = 1000, dt = 5e-5, n_frames frames of frame_steps
    substeps (3000 substeps).  A kinematic box gripper
    (set_velocity_on_cuboid, half extents gripper_half) starts just above the
    slab's top at its centre and follows a seeded piecewise-linear path:
    pieces of 5 frames, the first pressing down, the others dragging in the
    x-z plane, every speed at most vmax (0.5 m/s).  params["gripper"] holds
    the start point, the half extents and the (n_frames, 3) per-frame
    velocities."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 1.0 / n_grid
    d = np.asarray(dims, np.float64)
    lo = np.array([0.5 - d[0] / 2, 0.5 - d[1] / 2, 3 * dx + 0.25 * dx])
    x = (lo + rng.random((n_particles, 3)) * d).astype(np.float32)
    N = x.shape[0]
    f = lambda v: np.full(N, v, np.float32)
    prm = sample_params_c2(int(rng.integers(1 << 30)))
    vel = np.zeros((n_frames, 3))
    for p0 in range(0, n_frames, 5):
        if p0 == 0:
            v = np.array([0.0, 0.0, -rng.uniform(0.2, vmax)])
        else:
            ang = rng.uniform(-0.25 * np.pi, 0.25 * np.pi) + (np.pi if rng.random() < 0.5 else 0.0)
            s = rng.uniform(0.1, vmax)
            v = np.array([s * np.cos(ang), 0.0, -abs(s * np.sin(ang)) * 0.3])
        vel[p0:p0 + 5] = v
    top = lo[2] + d[2]
    grip = {"point": (0.5, 0.5, top + gripper_half[2] + 0.5 * dx), "half": tuple(gripper_half), "velocity": vel}
    sc = Scene(name=f"slab_{N}_{n_grid}^3_nh_vm", n_grid=n_grid, grid_lim=1.0, dt=dt, n_steps=n_frames * frame_steps,
               material=7, x=x, vol=f(float(np.prod(d)) / N), rho=f(1000.0), E=f(prm["E"]), nu=f(prm["nu"]),
               yield_stress=f(prm["yield_stress"]), colliders=sticky_walls(n_grid, 3), params=prm)
    sc.params["gripper"] = grip
    sc.params["frame_steps"] = frame_steps
    sc.params["n_frames"] = n_frames
    sc.params["v0"] = np.zeros((N, 3), np.float32)
    return sc
