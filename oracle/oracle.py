"""ctypes wrapper of the CPU oracle (liboracle_f64.so / liboracle_f32.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
The C restatement is mpm_oracle.c in this folder; build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OracleParams(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int),
        ("n_grid", C.c_int),
        ("material", C.c_int),
        ("pad0", C.c_int),
        ("grid_lim", C.c_double),
        ("dx", C.c_double),
        ("inv_dx", C.c_double),
        ("rpic_damping", C.c_double),
        ("grid_v_damping_scale", C.c_double),
        ("gravity", C.c_double * 3),
        ("alpha", C.c_double),
        ("hardening", C.c_double),
        ("xi", C.c_double),
        ("plastic_viscosity", C.c_double),
        ("softening", C.c_double),
    ]


class OracleCollider(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("surface_type", C.c_int),
        ("geo_type", C.c_int),
        ("reset", C.c_int),
        ("active", C.c_int),
        ("pad0", C.c_int),
        ("start_time", C.c_double),
        ("end_time", C.c_double),
        ("point", C.c_double * 3),
        ("normal", C.c_double * 3),
        ("size", C.c_double * 3),
        ("velocity", C.c_double * 3),
        ("rot_axis", C.c_double * 3),
        ("angular_velocity", C.c_double),
        ("radius", C.c_double),
        ("half_height", C.c_double),
        ("friction", C.c_double),
        ("mask", C.c_void_p),
    ]


_libs = {}


def load(precision: str = "f64"):
    if precision not in ("f64", "f32"):
        raise ValueError(precision)
    if precision in _libs:
        return _libs[precision]
    path = os.path.join(HERE, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run `make -C oracle`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.oracle_svd3.argtypes = [vp, vp, vp, vp]
    lib.oracle_svd3.restype = C.c_int
    lib.oracle_constitutive.argtypes = [C.c_int, vp, vp, vp, vp, C.POINTER(OracleParams), C.c_double, vp, vp]
    lib.oracle_constitutive.restype = C.c_int
    lib.oracle_mu_lam.argtypes = [C.c_int, vp, vp, vp, vp]
    lib.oracle_mu_lam.restype = None
    lib.oracle_bspline.argtypes = [vp, C.c_double, vp, vp, vp, vp]
    lib.oracle_bspline.restype = None
    lib.oracle_substep.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleCollider), C.c_int, C.c_double,
                                   C.c_double] + [vp] * 16
    lib.oracle_substep.restype = C.c_int
    lib.oracle_rollout.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleCollider), C.c_int, C.c_double,
                                   C.c_double, C.c_int] + [vp] * 10
    lib.oracle_rollout.restype = C.c_double
    lib.oracle_substep_frozen_mass.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleCollider), C.c_int,
                                               C.c_double, C.c_double] + [vp] * 15
    lib.oracle_substep_frozen_mass.restype = C.c_int
    lib.oracle_rollout_mass_hist.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleCollider), C.c_int,
                                             C.c_double, C.c_double, C.c_int] + [vp] * 12
    lib.oracle_rollout_mass_hist.restype = C.c_double
    lib.oracle_grid_update.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleCollider), C.c_int, C.c_double,
                                       C.c_double, vp, vp]
    lib.oracle_grid_update.restype = None
    lib.oracle_advance_colliders.argtypes = [C.POINTER(OracleCollider), C.c_int, C.c_double, C.c_double]
    lib.oracle_advance_colliders.restype = None
    lib.oracle_restart_F.argtypes = [C.POINTER(OracleParams), vp, vp, vp, vp, vp]
    lib.oracle_restart_F.restype = C.c_int
    lib.oracle_num_threads.restype = C.c_int
    lib.oracle_set_num_threads.argtypes = [C.c_int]
    lib.oracle_set_num_threads.restype = None
    _libs[precision] = lib
    return lib


def _dt(precision):
    return np.float64 if precision == "f64" else np.float32


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(n_particles, n_grid, material=0, grid_lim=1.0, rpic_damping=0.0, grid_v_damping_scale=1.1,
                gravity=(0.0, 0.0, -9.8), friction_angle=25.0, hardening=0.0, xi=0.0, plastic_viscosity=0.0,
                softening=0.1, dx=None, inv_dx=None):
    """Model scalars with the defaults of init_other_params
    (mpm_data_structure.py:456-484); dx/inv_dx are float32 there."""
    P = OracleParams()
    P.n_particles = int(n_particles)
    P.n_grid = int(n_grid)
    P.material = int(material)
    P.grid_lim = grid_lim
    P.dx = float(np.float32(grid_lim / n_grid)) if dx is None else dx
    P.inv_dx = float(np.float32(n_grid / grid_lim)) if inv_dx is None else inv_dx
    P.rpic_damping = rpic_damping
    P.grid_v_damping_scale = grid_v_damping_scale
    for i in range(3):
        P.gravity[i] = float(np.float32(gravity[i]))
    sin_phi = np.sin(friction_angle / 180.0 * 3.14159265)
    P.alpha = float(np.float32(np.sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi)))
    P.hardening = hardening
    P.xi = xi
    P.plastic_viscosity = plastic_viscosity
    P.softening = softening
    return P


def colliders_array(cols):
    arr = (OracleCollider * max(1, len(cols)))()
    for i, c in enumerate(cols):
        arr[i] = c
    return arr, len(cols)


class OracleState:
    """Host arrays of one scene in one precision."""

    def __init__(self, x, v, C_, F, mu, lam, ys, mass, vol, sel=None, precision="f64"):
        dt = _dt(precision)
        self.precision = precision
        self.x = np.ascontiguousarray(x, dtype=dt).reshape(-1, 3).copy()
        self.v = np.ascontiguousarray(v, dtype=dt).reshape(-1, 3).copy()
        self.C = np.ascontiguousarray(C_, dtype=dt).reshape(-1, 3, 3).copy()
        self.F = np.ascontiguousarray(F, dtype=dt).reshape(-1, 3, 3).copy()
        self.mu = np.ascontiguousarray(mu, dtype=dt).copy()
        self.lam = np.ascontiguousarray(lam, dtype=dt).copy()
        self.ys = np.ascontiguousarray(ys, dtype=dt).copy()
        self.mass = np.ascontiguousarray(mass, dtype=dt).copy()
        self.vol = np.ascontiguousarray(vol, dtype=dt).copy()
        self.sel = None if sel is None else np.ascontiguousarray(sel, dtype=np.int32).copy()

    def copy(self):
        return OracleState(self.x, self.v, self.C, self.F, self.mu, self.lam, self.ys, self.mass, self.vol,
                           self.sel, self.precision)


def substep(P, cols, time, dt, s: OracleState, return_fe=False, m_fixed=None):
    """One substep; returns the next OracleState (mu/lam/ys updated in place
    in `s` by material 5, then copied).  m_fixed (n^3 grid masses) holds the
    normalisation's grid mass at those values: central differences of that
    function give the gradient with the grid-mass adjoint dropped."""
    lib = load(s.precision)
    N, n = P.n_particles, P.n_grid
    d = _dt(s.precision)
    x1 = np.zeros((N, 3), d)
    v1 = np.zeros((N, 3), d)
    C1 = np.zeros((N, 3, 3), d)
    F1 = np.zeros((N, 3, 3), d)
    grid = np.zeros(n * n * n * 4, d)
    Fe = np.zeros((N, 3, 3), d)
    tau = np.zeros((N, 3, 3), d)
    arr, nc = colliders_array(cols)
    if m_fixed is not None:
        if return_fe:
            raise ValueError("return_fe with m_fixed is not supported")
        mf = np.ascontiguousarray(m_fixed, dtype=d).reshape(-1)
        if mf.size != n * n * n:
            raise ValueError(f"m_fixed has {mf.size} entries, the grid {n ** 3}")
        rc = lib.oracle_substep_frozen_mass(C.byref(P), arr, nc, time, dt, _p(s.x), _p(s.v), _p(s.C), _p(s.F),
                                            _p(s.mu), _p(s.lam), _p(s.ys), _p(s.mass), _p(s.vol), _p(s.sel), _p(x1),
                                            _p(v1), _p(C1), _p(F1), _p(grid), _p(mf))
    else:
        rc = lib.oracle_substep(C.byref(P), arr, nc, time, dt, _p(s.x), _p(s.v), _p(s.C), _p(s.F), _p(s.mu),
                                _p(s.lam), _p(s.ys), _p(s.mass), _p(s.vol), _p(s.sel), _p(x1), _p(v1), _p(C1),
                                _p(F1), _p(grid), _p(Fe), _p(tau))
    if rc != 0:
        raise RuntimeError("oracle_substep failed")
    out = OracleState(x1, v1, C1, F1, s.mu, s.lam, s.ys, s.mass, s.vol, s.sel, s.precision)
    if return_fe:
        return out, Fe, tau, grid.reshape(n, n, n, 4)
    return out


def restart_F(P, x, target, mass, F, sel=None, precision="f64"):
    """restart_and_compute_F_C's deformation gradient (mpm_solver_diff.py:
    1214-1262, mpm_utils.py:843-919): returns a copy of F (N, 3, 3) with
    every selected particle's F_trial = I + the grid gradient of the
    mass-weighted displacement target - x (oracle_restart_F)."""
    lib = load(precision)
    d = _dt(precision)
    x = np.ascontiguousarray(x, dtype=d)
    target = np.ascontiguousarray(target, dtype=d)
    mass = np.ascontiguousarray(mass, dtype=d)
    out = np.ascontiguousarray(F, dtype=d).copy()
    sl = None if sel is None else np.ascontiguousarray(sel, dtype=np.int32)
    if lib.oracle_restart_F(C.byref(P), _p(x), _p(target), _p(mass), _p(sl), _p(out)) != 0:
        raise RuntimeError("oracle_restart_F failed")
    return out


def rollout(P, cols, t0, dt, n_steps, s: OracleState, m_hist_out=None, m_hist_in=None):
    """In-place rollout of n_steps; returns the final time.  `cols` is a
    ctypes array (advanced in place) or a list.  m_hist_out (n_steps, n^3),
    if given, receives every substep's grid mass; m_hist_in (same shape)
    holds every substep's normalisation mass at those values."""
    lib = load(s.precision)
    if not isinstance(cols, C.Array):
        cols, nc = colliders_array(cols)
    else:
        nc = len(cols)
    if m_hist_out is None and m_hist_in is None:
        t = lib.oracle_rollout(C.byref(P), cols, nc, t0, dt, int(n_steps), _p(s.x), _p(s.v), _p(s.C), _p(s.F),
                               _p(s.mu), _p(s.lam), _p(s.ys), _p(s.mass), _p(s.vol), _p(s.sel))
    else:
        G = P.n_grid ** 3
        for h in (m_hist_out, m_hist_in):
            if h is not None and (h.dtype != _dt(s.precision) or h.size != int(n_steps) * G
                                  or not h.flags.c_contiguous):
                raise ValueError(f"mass history must be a contiguous {_dt(s.precision).__name__} array of "
                                 f"{n_steps} x {G}")
        t = lib.oracle_rollout_mass_hist(C.byref(P), cols, nc, t0, dt, int(n_steps), _p(s.x), _p(s.v), _p(s.C),
                                         _p(s.F), _p(s.mu), _p(s.lam), _p(s.ys), _p(s.mass), _p(s.vol), _p(s.sel),
                                         _p(m_hist_out), _p(m_hist_in))
    if t < 0:
        raise RuntimeError("oracle_rollout failed")
    return t


def grid_update(P, cols, time, dt, grid, precision="f64"):
    """v_out (n,n,n,3) from a (n,n,n,4) grid of (mv, m)."""
    lib = load(precision)
    d = _dt(precision)
    n = P.n_grid
    g = np.ascontiguousarray(grid, dtype=d).reshape(n, n, n, 4)
    vout = np.zeros((n, n, n, 3), d)
    arr, nc = colliders_array(cols)
    lib.oracle_grid_update(C.byref(P), arr, nc, time, dt, _p(g), _p(vout))
    return vout


def advance_colliders(cols, time, dt):
    arr, nc = colliders_array(cols)
    load("f64").oracle_advance_colliders(arr, nc, time, dt)
    return [arr[i] for i in range(nc)]


def svd3(F, precision="f64"):
    lib = load(precision)
    d = _dt(precision)
    F = np.ascontiguousarray(F, dtype=d).reshape(3, 3)
    U = np.zeros((3, 3), d)
    s = np.zeros(3, d)
    V = np.zeros((3, 3), d)
    lib.oracle_svd3(_p(F), _p(U), _p(s), _p(V))
    return U, s, V


def constitutive(material, Ftr, mu, lam, ys, P, dt=1e-4, precision="f64"):
    """(Fe, tau_sym, mu, lam, ys) for one particle."""
    lib = load(precision)
    d = _dt(precision)
    F = np.ascontiguousarray(Ftr, dtype=d).reshape(3, 3)
    m = np.array([mu], d)
    l = np.array([lam], d)
    y = np.array([ys], d)
    Fe = np.zeros((3, 3), d)
    tau = np.zeros((3, 3), d)
    lib.oracle_constitutive(int(material), _p(F), _p(m), _p(l), _p(y), C.byref(P), dt, _p(Fe), _p(tau))
    return Fe, tau, m[0], l[0], y[0]


def bspline(xp, inv_dx, precision="f64"):
    lib = load(precision)
    d = _dt(precision)
    xp = np.ascontiguousarray(xp, dtype=d).reshape(3)
    base = np.zeros(3, np.int32)
    f = np.zeros(3, d)
    w = np.zeros((3, 3), d)
    dw = np.zeros((3, 3), d)
    lib.oracle_bspline(_p(xp), inv_dx, _p(base), _p(f), _p(w), _p(dw))
    return base, f, w, dw


def mu_lam(E, nu, precision="f64"):
    lib = load(precision)
    d = _dt(precision)
    E = np.ascontiguousarray(E, dtype=d)
    nu = np.ascontiguousarray(nu, dtype=d)
    mu = np.zeros_like(E)
    lam = np.zeros_like(E)
    lib.oracle_mu_lam(int(E.shape[0]), _p(E), _p(nu), _p(mu), _p(lam))
    return mu, lam


def set_threads(k):
    for p in ("f64", "f32"):
        try:
            load(p).oracle_set_num_threads(int(k))
        except ImportError:
            pass


# ----------------------------------------------------------------------
# Collider constructors with the reference's argument semantics
# ----------------------------------------------------------------------
def surface_collider(point, normal, surface="sticky", friction=0.0, start_time=0.0, end_time=999.0):
    """add_surface_collider (mpm_solver_diff.py:437-531)."""
    c = OracleCollider()
    c.kind = 0
    nrm = np.asarray(normal, np.float64)
    nrm = nrm / np.sqrt(float(np.sum(nrm ** 2)))
    for i in range(3):
        c.point[i] = float(np.float32(point[i]))
        c.normal[i] = float(np.float32(nrm[i]))
    c.surface_type = {"sticky": 0, "slip": 1, "cut": 11}.get(surface, 2)
    c.friction = friction
    c.start_time = float(np.float32(start_time))
    c.end_time = float(np.float32(end_time))
    return c


def cuboid(point, size, velocity, start_time=0.0, end_time=999.0, reset=0):
    """set_velocity_on_cuboid (mpm_solver_diff.py:814-869); scalar size is
    broadcast to the three axes."""
    c = OracleCollider()
    c.kind = 1
    sz = np.broadcast_to(np.asarray(size, np.float64), (3,))
    for i in range(3):
        c.point[i] = float(np.float32(point[i]))
        c.size[i] = float(np.float32(sz[i]))
        c.velocity[i] = float(np.float32(velocity[i]))
    c.start_time = float(np.float32(start_time))
    c.end_time = float(np.float32(end_time))
    c.reset = int(reset)
    return c


def bounding_box(start_time=0.0, end_time=999.0):
    c = OracleCollider()
    c.kind = 2
    c.start_time = float(np.float32(start_time))
    c.end_time = float(np.float32(end_time))
    return c
