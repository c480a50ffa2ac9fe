"""Shared helpers for the parity tests: seeded random substep inputs, the GPU
path through the C-ABI (empm_amd.runtime) and the CPU oracle."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc

MATERIAL_NAMES = {0: "jelly", 1: "metal", 2: "sand", 3: "foam", 4: "snow", 5: "plasticine", 6: "neo-hookean",
                  7: "nh-vm"}


def random_state(N, n_grid, seed, material, deform=0.05, vscale=0.5, cscale=5.0):
    """Random particles in the interior of a grid of n_grid cells on [0,1]^3
    (3..n-3 cells away from the walls), with velocities, affine C and
    F_trial = R (I + deform * A) with a random rotation R."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 1.0 / n_grid
    x = rng.uniform(3.5 * dx, 1.0 - 3.5 * dx, (N, 3))
    v = rng.normal(0, vscale, (N, 3))
    Cm = rng.normal(0, cscale, (N, 3, 3))
    A = rng.normal(0, 1.0, (N, 3, 3))
    q = rng.normal(0, 1, (N, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, a, b, c = q.T
    R = np.stack([
        np.stack([1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)], -1),
        np.stack([2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)], -1),
        np.stack([2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)], -1),
    ], 1)
    F = R @ (np.eye(3)[None] + deform * A)
    E = rng.uniform(1e4, 1e5, N)
    nu = rng.uniform(0.2, 0.4, N)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    ys = rng.uniform(50.0, 2000.0, N)
    vol = np.full(N, (dx ** 3) / 8)
    mass = vol * 1000.0
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(x=f32(x), v=f32(v), C=f32(Cm), F=f32(F), mu=f32(mu), lam=f32(lam), ys=f32(ys), mass=f32(mass),
                vol=f32(vol), E=f32(E), nu=f32(nu))


def oracle_state(d, precision):
    return orc.OracleState(d["x"], d["v"], d["C"], d["F"], d["mu"], d["lam"], d["ys"], d["mass"], d["vol"],
                           precision=precision)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def state_errors(g, ref, x0=None, F0=None):
    """Errors of a GPU state against a reference state (oracle).

    g, ref: objects or dicts with x, v, C, F (any present).  Returns
    - "x": max |dx| / max |x_ref| (positions, relative to their magnitude);
    - "disp" (with x0): max |dx| / max |x_ref - x0|, the error relative to
      how far the particles moved.  Positions near 0.5 m hide a displacement
      error of several percent behind a 1e-4 position bar; this does not;
    - "v", "C": max |d| / max |ref|;
    - "F": max |dF| / max |F_ref| and, with F0, "F_strain": max |dF| /
      max |F_ref - F0| (relative to the deformation itself)."""
    get = (lambda o, k: o[k]) if isinstance(g, dict) else getattr
    getr = (lambda o, k: o[k]) if isinstance(ref, dict) else getattr
    out = {}
    for k in ("x", "v", "C", "F"):
        try:
            a = np.asarray(get(g, k), np.float64)
            b = np.asarray(getr(ref, k), np.float64)
        except (KeyError, AttributeError):
            continue
        d = np.max(np.abs(a - b)) if a.size else 0.0
        out[k] = float(d / max(np.max(np.abs(b)) if b.size else 0.0, 1e-30))
        if k == "x" and x0 is not None:
            out["disp"] = float(d / max(np.max(np.abs(b - np.asarray(x0, np.float64))), 1e-30))
        if k == "F" and F0 is not None:
            out["F_strain"] = float(d / max(np.max(np.abs(b - np.asarray(F0, np.float64))), 1e-30))
    return out


def fmt_errs(e):
    return " ".join(f"{k} {v:.2e}" for k, v in e.items())


def gpu_collider(c):
    """The empm_collider of an oracle collider (same fields, no mask)."""
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
