"""GPU parity of one substep and of its adjoint, through the C-ABI.

Forward: libempm vs the FP32 oracle (bars: 1e-6 relative on x, 1e-4 on
F_trial, 1e-3 on v and C (v' and C' are sums of 27 nodes whose FP32
contributions arrive in another order), and on the displacement x' - x) with
the error against the FP64 oracle reported.  Adjoint: <J^T w, u> from libempm
vs the central difference of <w, step(s)> in FP64 on the oracle (bar: 1e-3
relative), in both adjoint modes: exact, and with the grid-mass adjoint
dropped as the reference Tape does (the FD then holds the normalisation's
grid mass at the unperturbed substep's value).
"""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from oracle import oracle as orc
from tests.helpers import fmt_errs, oracle_state, random_state, rel_err, state_errors

pytestmark = pytest.mark.gpu

N_GRID = 16
N = 600
DT = 1e-4
MATERIALS = [0, 1, 2, 3, 4, 5, 6, 7]


def _model(d, material, dev, gravity=(0.0, 0.0, -9.8), alpha=None, damping=1.1, rpic=0.0, visc=0.0, hardening=0.0,
           xi=0.0):
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    P = orc.make_params(len(d["x"]), N_GRID, material=material, gravity=gravity, grid_v_damping_scale=damping,
                        rpic_damping=rpic, plastic_viscosity=visc, hardening=hardening, xi=xi)
    m = rt.ModelArrays(material=material, n_grid=N_GRID, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(d["mu"]),
                       lam=t(d["lam"]), yield_stress=t(d["ys"]), mass=t(d["mass"]), vol=t(d["vol"]),
                       gravity=gravity, alpha=P.alpha if alpha is None else alpha, grid_v_damping_scale=damping,
                       rpic_damping=rpic, plastic_viscosity=visc, hardening=hardening, xi=xi)
    return m, P


def _pstate(d, dev):
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    return rt.ParticleState(t(d["x"]), t(d["v"]), t(d["C"]), t(d["F"]))


def _deform(material):
    # plastic materials: large enough deformation that a good share yields
    return {1: 0.15, 2: 0.15, 3: 0.15, 5: 0.15, 7: 0.15}.get(material, 0.05)


@pytest.mark.parametrize("material", MATERIALS)
def test_substep_forward_parity(dev, material):
    print(f"[substep fwd] material={material} N={N} n={N_GRID}", flush=True)
    d = random_state(N, N_GRID, seed=10 + material, material=material, deform=_deform(material))
    cols = [orc.surface_collider((0, 0, 0.3), (0, 0, 1))]
    model, P = _model(d, material, dev, visc=0.5 if material == 3 else 0.0)
    ctx = rt.Context(N_GRID, N, dev)
    s_in = _pstate(d, dev)
    s_out = rt.ParticleState.empty(N, dev)
    gcols = [_gpu_collider(c) for c in cols]
    rt.substep_fwd(ctx, model, gcols, 0.0, DT, s_in, s_out)
    torch.cuda.synchronize()
    o32 = orc.substep(P, cols, 0.0, DT, oracle_state(d, "f32"))
    o64 = orc.substep(P, cols, 0.0, DT, oracle_state(d, "f64"))
    g = {k: getattr(s_out, k).cpu().numpy() for k in ("x", "v", "C", "F")}
    errs = state_errors(g, o32, x0=d["x"], F0=d["F"])
    errs64 = state_errors(g, o64, x0=d["x"], F0=d["F"])
    print(f"  vs f32 oracle: {fmt_errs(errs)}\n  vs f64 oracle: {fmt_errs(errs64)}", flush=True)
    assert errs["x"] < 1e-6, errs
    assert errs["F"] < 1e-4, errs
    assert errs["v"] < 1e-3, errs
    assert errs["C"] < 1e-3, errs
    # x' - x = dt v' (about 5e-5 m): its FP32 rounding at |x| ~ 0.5 alone is
    # 3e-8 / 5e-5 ~ 6e-4 of it
    assert errs["disp"] < 2e-3, errs
    if material == 5:
        # material 5 rewrites yield stress / mu / lam in place
        assert rel_err(model.yield_stress.cpu().numpy(), o32.ys) < 1e-4


def _gpu_collider(c):
    from empm_amd._lib import Collider

    g = Collider()
    for f, _ in Collider._fields_:
        if f == "mask":
            continue
        val = getattr(c, f)
        if isinstance(val, (int, float)):
            setattr(g, f, val)
        else:
            for i in range(3):
                getattr(g, f)[i] = val[i]
    return g


def _oracle_loss(d, material, w, P, cols, m_fixed=None):
    """<w, state'> + <w_mu, mu'> + <w_lam, lam'> + <w_ys, Y'>: the substep's
    outputs include the material state that material 5 and hardening rewrite
    (the oracle updates its arrays in place; for the other materials they
    pass through)."""
    o = orc.substep(P, cols, 0.0, DT, oracle_state(d, "f64"), m_fixed=m_fixed)
    return (np.sum(w["x"] * o.x) + np.sum(w["v"] * o.v) + np.sum(w["C"] * o.C) + np.sum(w["F"] * o.F)
            + np.sum(w["mu"] * o.mu) + np.sum(w["lam"] * o.lam) + np.sum(w["ys"] * o.ys))


@pytest.mark.parametrize("material", [0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("mode", ["exact", "drop_grid_mass"])
def test_substep_adjoint_directional(dev, material, mode):
    """<J^T w, u> vs central differences of the FP64 oracle."""
    _adjoint_directional(dev, material, mode)


@pytest.mark.parametrize("material", [1, 5, 7])
def test_substep_adjoint_directional_hardening(dev, material):
    """hardening == 1 (mpm_utils.py:124-127,178-181): the yield stress is
    state, Y' = Y (- softening ...) + 2 mu' xi dgamma; its cotangent runs
    back through the update, as do mu's and lam's for material 5."""
    _adjoint_directional(dev, material, "exact", hardening=1.0)


@pytest.mark.parametrize("material", [0, 7])
def test_substep_adjoint_directional_stencils_leaving_grid(dev, material):
    """As above with 12 particles whose stencils leave the grid (base -1 in
    x, base n - 2 in z): the adjoint kernels test the bounds of every node
    for the wave holding them and load the other waves' nodes unmasked."""
    _adjoint_directional(dev, material, "exact", edge=True)


def _adjoint_directional(dev, material, mode, edge=False, hardening=0.0):
    print(f"[substep adj] material={material} mode={mode} edge={edge} hardening={hardening}", flush=True)
    Ns = 200
    d = random_state(Ns, N_GRID, seed=100 + material, material=material, deform=_deform(material))
    if edge:
        d["x"][:6, 0] = -0.7 / N_GRID  # x inv_dx - 0.5 = -1.2: base -1 (truncation)
        d["x"][6:12, 2] = (N_GRID - 1.3) / N_GRID  # base n - 2: node n is outside
    if material == 5:
        # softening drives these to Y' <= 0 (mu' = lam' = 0), and these are
        # already damaged (Y <= 0: F_e = F_trial)
        d["ys"][20:30] = np.float32(1e-2)
        d["ys"][30:35] = np.float32(-1.0)
        d["mu"][30:35] = 0.0
        d["lam"][30:35] = 0.0
    rng = np.random.Generator(np.random.PCG64(7 + material))
    cols = [orc.surface_collider((0, 0, 0.3), (0, 0, 1))]
    model, P = _model(d, material, dev, visc=0.5 if material == 3 else 0.0, damping=0.98, hardening=hardening,
                      xi=0.3 if hardening else 0.0)
    P.n_particles = Ns
    P.grid_v_damping_scale = 0.98
    ctx = rt.Context(N_GRID, Ns, dev)
    w = {k: rng.normal(0, 1, d[k].shape) for k in ("x", "v", "C", "F")}
    # cotangents of the material state after the substep, on the scale of
    # the state's other terms
    for k in ("mu", "lam", "ys"):
        w[k] = rng.normal(0, 1, d[k].shape) / np.maximum(np.abs(d[k]), 1.0) * 1e-2
    u = {k: rng.normal(0, 1, d[k].shape) for k in ("x", "v", "C", "F", "mu", "lam", "ys")}
    s_in = _pstate(d, dev)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)
    d_out = rt.ParticleState(t(w["x"]), t(w["v"]), t(w["C"]), t(w["F"]))
    d_in = rt.ParticleState.empty(Ns, dev)
    grads = rt.ParamGrads.zeros(Ns, dev)
    grads.mu.copy_(t(w["mu"]))
    grads.lam.copy_(t(w["lam"]))
    grads.ys.copy_(t(w["ys"]))
    flags = 0 if mode == "exact" else 1
    rt.substep_adj(ctx, model, [_gpu_collider(c) for c in cols], 0.0, DT, s_in, d_out, d_in, grads, flags)
    torch.cuda.synchronize()
    g = {"x": d_in.x, "v": d_in.v, "C": d_in.C, "F": d_in.F, "mu": grads.mu, "lam": grads.lam, "ys": grads.ys}
    g = {k: v.cpu().numpy().astype(np.float64) for k, v in g.items()}
    scales = {"x": 1e-3 / N_GRID, "v": 1e-2, "C": 1e-1, "F": 1e-3, "mu": 1e-3 * np.abs(d["mu"]),
              "lam": 1e-3 * np.abs(d["lam"]), "ys": 1e-3 * np.maximum(np.abs(d["ys"]), 1e-2)}
    m_fixed = None
    if mode == "drop_grid_mass":
        # the reference Tape's function: grid mass held at this substep's value
        _, _, _, grid = orc.substep(P, cols, 0.0, DT, oracle_state(d, "f64"), return_fe=True)
        m_fixed = grid[..., 3].copy()
    for key in ("x", "v", "C", "F", "mu", "lam", "ys"):
        du = u[key] * (scales[key] if np.ndim(scales[key]) == 0 else scales[key].reshape((-1,) + (1,) * (u[key].ndim - 1)))
        an = float(np.sum(g[key] * du))
        dd_p = {k: (np.asarray(v, np.float64).copy()) for k, v in d.items()}
        dd_m = {k: (np.asarray(v, np.float64).copy()) for k, v in d.items()}
        eps = 1e-4
        dd_p[key] = dd_p[key] + eps * du
        dd_m[key] = dd_m[key] - eps * du
        fd = (_oracle_loss(dd_p, material, w, P, cols, m_fixed) - _oracle_loss(dd_m, material, w, P, cols, m_fixed)) \
            / (2 * eps)
        rel = abs(an - fd) / max(abs(fd), 1e-12)
        print(f"  d/d{key}: adjoint {an:+.6e}  fd {fd:+.6e}  rel {rel:.2e}", flush=True)
        assert rel < 1e-3, (key, mode, an, fd, rel)
