"""Every grid post-process (collider) kind and the particle selection on
the GPU against the FP32 oracle, and the adjoint's masking of the nodes they
overwrite against central differences of the FP64 oracle.

Kinds (mpm_solver_diff.py:437-938,1264-1289): bounding box; geometry
colliders (cube, sphere, cylinder, half-space; the half-space and the
surface collider also with the "cut" rule); the grid geometry mover with an
angular velocity, whose point the rollout advances; the mask; the velocity
cuboid with reset == 1, which zeroes the whole grid from before its start
until 15 substeps after its end; and particles with selection != 0, which
the substep does not simulate (partial_clone defaults,
mpm_data_structure.py:356-387).  Each: a 12-substep rollout (positions and
velocities, 1e-4 of the positions and 1e-3 of the displacement) and one
substep's adjoint in a random direction in exact mode (1e-3).
"""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd import scenes
from oracle import oracle as orc
from tests.helpers import fmt_errs, gpu_collider, oracle_state, state_errors

pytestmark = pytest.mark.gpu

N_GRID = 24
DT = 1e-4
STEPS = 12
CENTER = (0.5, 0.5, 0.4)


def _geo(kind, geo_type, point, size=(0, 0, 0), radius=0.0, half_height=0.0, normal=(0, 0, 1), velocity=(0, 0, 0),
         rot_axis=(0, 0, 1), angular_velocity=0.0, active=1, start_time=0.0, end_time=999.0, reset=0,
         surface_type=0):
    c = orc.OracleCollider()
    c.kind, c.geo_type, c.active, c.reset, c.surface_type = kind, geo_type, active, reset, surface_type
    nrm = np.asarray(normal, np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    for i in range(3):
        c.point[i], c.size[i] = float(np.float32(point[i])), float(np.float32(size[i]))
        c.normal[i] = float(np.float32(nrm[i]))
        c.velocity[i], c.rot_axis[i] = float(np.float32(velocity[i])), float(np.float32(rot_axis[i]))
    c.radius, c.half_height = float(np.float32(radius)), float(np.float32(half_height))
    c.angular_velocity = float(np.float32(angular_velocity))
    c.start_time, c.end_time = float(np.float32(start_time)), float(np.float32(end_time))
    return c


# name -> (collider list builder, particle velocity, uses a mask, uses selection)
CASES = {
    "bounding_box": (lambda: [orc.bounding_box()], (2.5, -2.5, -2.0)),
    "geometry_cube": (lambda: [_geo(3, 1, CENTER, size=(0.05, 0.04, 0.06))], (0.3, -0.2, -1.0)),
    "geometry_sphere": (lambda: [_geo(3, 2, CENTER, radius=0.07, start_time=2e-4)], (0.3, -0.2, -1.0)),
    "geometry_cylinder": (lambda: [_geo(3, 3, CENTER, radius=0.06, half_height=0.05, normal=(1, 1, 0))],
                          (0.3, -0.2, -1.0)),
    "geometry_halfspace": (lambda: [_geo(3, 4, (0.5, 0.5, 0.38), normal=(0.2, 0.1, 1.0))], (0.3, -0.2, -1.0)),
    "geometry_halfspace_cut": (lambda: [_geo(3, 4, (0.5, 0.5, 0.45), normal=(0, 0, 1), surface_type=11)],
                               (0.3, -0.2, -1.0)),
    "surface_cut": (lambda: [orc.surface_collider((0.5, 0.5, 0.5), (0, 0, 1), surface="cut")], (0.3, -0.2, -1.0)),
    "surface_window": (lambda: [orc.surface_collider((0.5, 0.5, 0.4), (0, 0.3, 1), start_time=3e-4,
                                                     end_time=8e-4)], (0.3, -0.2, -1.0)),
    "mover_rotating": (lambda: [_geo(4, 2, CENTER, radius=0.08, velocity=(0.4, 0.0, -0.2), rot_axis=(0, 0.6, 0.8),
                                     angular_velocity=6.0)], (0.3, -0.2, -1.0)),
    "mover_cube_reset": (lambda: [_geo(4, 1, CENTER, size=(0.05, 0.05, 0.05), velocity=(0.5, 0, 0), end_time=5e-4,
                                       reset=1)], (0.3, -0.2, -1.0)),
    "cuboid_reset": (lambda: [orc.cuboid(CENTER, 0.05, (0.5, 0.0, 0.0), start_time=3e-4, end_time=6e-4, reset=1)],
                     (0.3, -0.2, -1.0)),
    "mask": (lambda: [None], (0.3, -0.2, -1.0)),
    "selection": (lambda: [orc.surface_collider((0, 0, 0.3), (0, 0, 1))], (0.3, -0.2, -1.0)),
}


def _scene(name):
    sc = scenes.small_scene(n_particles=700, n_grid=N_GRID, material=0, seed=31, E=2e4, center=CENTER,
                            vel=CASES[name][1], spin=3.0, dt=DT)
    return sc


def _colliders(name, dev):
    """(oracle colliders, GPU colliders, keep-alive)."""
    cols = CASES[name][0]()
    keep = []
    if name == "mask":
        mask = np.random.Generator(np.random.PCG64(5)).integers(-1, 4, (N_GRID,) * 3).astype(np.int32)
        mask[mask == 2] = 0  # about a quarter of the nodes marked
        c = orc.OracleCollider()
        c.kind = 5
        c.mask = mask.ctypes.data
        cols = [c]
        dmask = torch.as_tensor(mask, device=dev).contiguous()
        keep += [mask, dmask]
        g = gpu_collider(c)
        g.mask = dmask.data_ptr()
        return cols, [g], keep
    return cols, [gpu_collider(c) for c in cols], keep


def _selection(name, N):
    if name != "selection":
        return None
    sel = np.zeros(N, np.int32)
    sel[::9] = 1
    return sel


def _model(sc, dev, sel):
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    N = sc.n
    E64, nu64 = sc.E.astype(np.float64), sc.nu.astype(np.float64)
    mu = (E64 / (2 * (1 + nu64))).astype(np.float32)
    lam = (E64 * nu64 / ((1 + nu64) * (1 - 2 * nu64))).astype(np.float32)
    P = orc.make_params(N, N_GRID, material=0, gravity=sc.gravity)
    m = rt.ModelArrays(material=0, n_grid=N_GRID, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(mu), lam=t(lam),
                       yield_stress=t(sc.yield_stress), mass=t(sc.vol * sc.rho), vol=t(sc.vol),
                       selection=None if sel is None else torch.as_tensor(sel, device=dev).contiguous(),
                       gravity=sc.gravity)
    d = dict(x=sc.x, v=sc.params["v0"], C=np.zeros((N, 3, 3), np.float32),
             F=np.tile(np.eye(3, dtype=np.float32), (N, 1, 1)), mu=mu, lam=lam, ys=sc.yield_stress,
             mass=(sc.vol * sc.rho).astype(np.float32), vol=sc.vol)
    return m, P, d


@pytest.mark.parametrize("name", list(CASES))
def test_collider_rollout_matches_oracle(dev, name):
    sc = _scene(name)
    N = sc.n
    sel = _selection(name, N)
    model, P, d = _model(sc, dev, sel)
    cols, gcols, keep = _colliders(name, dev)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    s0 = rt.ParticleState(t(d["x"]), t(d["v"]), t(d["C"]), t(d["F"]))
    ctx = rt.Context(N_GRID, N, dev)
    cols_out = []
    final, t1, _, _ = rt.Rollout(ctx, model, N).forward(s0, list(gcols), 0.0, DT, STEPS, cols_out=cols_out)
    torch.cuda.synchronize()
    o = orc.OracleState(d["x"], d["v"], d["C"], d["F"], d["mu"], d["lam"], d["ys"], d["mass"], d["vol"], sel=sel,
                        precision="f32")
    arr, _ = orc.colliders_array(cols)
    orc.rollout(P, arr, 0.0, DT, STEPS, o)
    g = {"x": final.x.cpu().numpy(), "v": final.v.cpu().numpy(), "F": final.F.cpu().numpy()}
    e = state_errors(g, o, x0=d["x"])
    print(f"[collider {name}] N={N}: {fmt_errs(e)}", flush=True)
    assert e["x"] < 1e-4 and e["disp"] < 1e-3 and e["v"] < 1e-3, e
    # moving colliders: the rollout advanced the point as the host loop does
    for k, c in enumerate(cols):
        if c is not None and c.kind in (1, 4):
            for i in range(3):
                assert abs(cols_out[k].point[i] - arr[k].point[i]) < 1e-7, (k, i)
    if name == "selection":
        unsim = sel != 0
        assert np.all(g["x"][unsim] == 0) and np.all(g["v"][unsim] == 0)
    if name in ("cuboid_reset", "mover_cube_reset"):
        # the whole grid is zeroed in the reset window: nothing moves there
        assert np.abs(g["v"]).max() < 1e-6, np.abs(g["v"]).max()
    del keep


@pytest.mark.parametrize("name", ["bounding_box", "geometry_sphere", "geometry_halfspace_cut", "surface_cut",
                                  "mover_rotating", "mask", "cuboid_reset", "selection"])
def test_collider_adjoint_masks_overwritten_nodes(dev, name):
    """<J^T w, u> of one substep with the collider (exact mode) against
    central differences of the FP64 oracle: overwritten nodes carry no
    sensitivity to the particles that scattered into them."""
    sc = _scene(name)
    N = sc.n
    sel = _selection(name, N)
    model, P, d = _model(sc, dev, sel)
    cols, gcols, keep = _colliders(name, dev)
    time = 4e-4 if name == "cuboid_reset" else 2e-4  # inside the cuboid's window
    rng = np.random.Generator(np.random.PCG64(17))
    w = {k: rng.normal(0, 1, d[k].shape) for k in ("x", "v", "C", "F")}
    u = {k: rng.normal(0, 1, d[k].shape) for k in ("x", "v", "F")}
    scale = {"x": 1e-3 / N_GRID, "v": 1e-2, "F": 1e-3}
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    s_in = rt.ParticleState(t(d["x"]), t(d["v"]), t(d["C"]), t(d["F"]))
    d_out = rt.ParticleState(t(w["x"]), t(w["v"]), t(w["C"]), t(w["F"]))
    d_in = rt.ParticleState.empty(N, dev)
    grads = rt.ParamGrads.zeros(N, dev)
    ctx = rt.Context(N_GRID, N, dev)
    rt.substep_adj(ctx, model, list(gcols), time, DT, s_in, d_out, d_in, grads, 0)
    torch.cuda.synchronize()
    an = sum(float(np.sum(getattr(d_in, k).double().cpu().numpy() * u[k] * scale[k])) for k in u)

    def J(sign):
        dd = {k: np.asarray(v, np.float64).copy() for k, v in d.items()}
        for k in u:
            dd[k] = dd[k] + sign * 1e-4 * u[k] * scale[k]
        st = orc.OracleState(dd["x"], dd["v"], dd["C"], dd["F"], dd["mu"], dd["lam"], dd["ys"], dd["mass"], dd["vol"],
                             sel=sel)
        o = orc.substep(P, cols, time, DT, st)
        return sum(float(np.sum(w[k] * getattr(o, k))) for k in ("x", "v", "C", "F"))

    fd = (J(1) - J(-1)) / 2e-4
    rel = abs(an - fd) / abs(fd)
    print(f"[collider adjoint {name}] adjoint {an:+.6e} fd {fd:+.6e} rel {rel:.1e}", flush=True)
    assert rel < 1e-3, (an, fd, rel)
    del keep
