"""Pins the CPU oracle (oracle/mpm_oracle.c) against what mathematics fixes,
independently of the oracle's own formulas.  The reference ships no tests,
golden vectors or fixtures for this path, and Warp is absent, so each check
below is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.Generator(np.random.PCG64(1234))


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, a, b, c = q
    return np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
                     [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
                     [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)]])


# ---------------------------------------------------------------- B-spline
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_bspline_moment_identities(prec, tol):
    n = 32
    inv_dx = float(n)
    dx = 1.0 / n
    for _ in range(50):
        xp = RNG.uniform(3 * dx, 1 - 3 * dx, 3)
        base, f, w, dw = orc.bspline(xp, inv_dx, prec)
        W = np.zeros(27)
        off = np.zeros((27, 3))
        grad = np.zeros((27, 3))
        k = 0
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    W[k] = w[0, a] * w[1, b] * w[2, c]
                    grad[k] = inv_dx * np.array([dw[0, a] * w[1, b] * w[2, c], w[0, a] * dw[1, b] * w[2, c],
                                                 w[0, a] * w[1, b] * dw[2, c]])
                    xi = (base + np.array([a, b, c])) * dx
                    off[k] = xi - xp
                    k += 1
        assert abs(W.sum() - 1) < tol                                  # partition of unity
        assert np.abs((W[:, None] * off).sum(0)).max() < tol * dx      # first moment
        assert np.abs(grad.sum(0)).max() < tol * inv_dx                # sum of gradients
        M = (off[:, :, None] * grad[:, None, :]).sum(0)                # sum (x_i - x_p) grad^T = I
        assert np.abs(M - np.eye(3)).max() < 10 * tol
        D = (W[:, None, None] * off[:, :, None] * off[:, None, :]).sum(0)  # quadratic B-spline: dx^2/4 I
        assert np.abs(D - dx * dx / 4 * np.eye(3)).max() < tol * dx * dx
        assert np.all(base == np.trunc(xp * inv_dx - 0.5).astype(int))


# ---------------------------------------------------------------- SVD
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_svd_against_lapack(prec, tol):
    for i in range(200):
        F = RNG.normal(size=(3, 3))
        if i % 3 == 0:
            F = rand_rot(RNG) @ np.diag(RNG.uniform(0.5, 1.5, 3)) @ rand_rot(RNG).T
        U, s, V = orc.svd3(F, prec)
        assert np.abs(U @ np.diag(s) @ V.T - F).max() < tol * 10 * np.abs(F).max()
        assert abs(np.linalg.det(U) - 1) < 10 * tol and abs(np.linalg.det(V) - 1) < 10 * tol
        assert np.abs(U.T @ U - np.eye(3)).max() < 10 * tol
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.abs(np.abs(s) - ref).max() < 10 * tol * ref[0]
        assert s[0] >= 0 and s[1] >= 0
        assert np.sign(s[2]) == np.sign(np.linalg.det(F)) or abs(s[2]) < 1e-12
        assert abs(s[0]) >= abs(s[1]) - 1e-12 and abs(s[1]) >= abs(s[2]) - 1e-12


def test_svd_degenerate_inputs():
    for F in (np.eye(3), np.zeros((3, 3)), np.diag([2.0, 2.0, 2.0]), np.diag([1.0, 1.0, -1.0]),
              np.outer([1, 2, 3], [0.5, -1, 2.0]), rand_rot(RNG)):
        U, s, V = orc.svd3(F)
        assert np.abs(U @ np.diag(s) @ V.T - F).max() < 1e-10
        assert abs(np.linalg.det(U) - 1) < 1e-10 and abs(np.linalg.det(V) - 1) < 1e-10


# ---------------------------------------------------------------- stresses
def P_(material, **kw):
    return orc.make_params(1, 16, material=material, **kw)


@pytest.mark.parametrize("material", [0, 6, 1, 2, 3, 4, 5, 7])
def test_zero_stress_at_identity(material):
    Fe, tau, *_ = orc.constitutive(material, np.eye(3), 3e4, 5e4, 1e3, P_(material))
    assert np.abs(tau).max() < 1e-9
    assert np.abs(Fe - np.eye(3)).max() < 1e-12


def test_fcr_zero_at_rotations_and_scale_closed_form():
    mu, lam = 3e4, 5e4
    for _ in range(20):
        R = rand_rot(RNG)
        _, tau, *_ = orc.constitutive(0, R, mu, lam, 0.0, P_(0))
        assert np.abs(tau).max() < 1e-7
    for s in (0.8, 0.95, 1.1, 1.3):
        _, tau, *_ = orc.constitutive(0, s * np.eye(3), mu, lam, 0.0, P_(0))
        ref = (2 * mu * (s - 1) * s + lam * s ** 3 * (s ** 3 - 1)) * np.eye(3)
        assert np.abs(tau - ref).max() < 1e-8 * np.abs(ref).max()


def test_fcr_rotation_invariance():
    """tau(R F) = R tau(F) R^T (frame indifference) — fails on a transposed R."""
    mu, lam = 3e4, 5e4
    for _ in range(20):
        F = np.eye(3) + 0.2 * RNG.normal(size=(3, 3))
        R = rand_rot(RNG)
        _, t1, *_ = orc.constitutive(0, F, mu, lam, 0.0, P_(0))
        _, t2, *_ = orc.constitutive(0, R @ F, mu, lam, 0.0, P_(0))
        assert np.abs(t2 - R @ t1 @ R.T).max() < 1e-8 * np.abs(t1).max()


def test_nh_closed_forms():
    mu, lam = 3e4, 5e4
    for s in (0.8, 0.95, 1.1, 1.3):
        _, tau, *_ = orc.constitutive(6, s * np.eye(3), mu, lam, 0.0, P_(6))
        ref = s * (lam / 2) * (s ** 6 - 1) * np.eye(3)
        assert np.abs(tau - ref).max() < 1e-8 * np.abs(ref).max()
    for _ in range(10):
        _, tau, *_ = orc.constitutive(6, rand_rot(RNG), mu, lam, 0.0, P_(6))
        assert np.abs(tau).max() < 1e-7
    # diagonal F: tau = diag(t_i * s_i) with t_i = mu J^(-2/3) (b_i - mean b) + lam/2 (J^2 - 1)
    s = np.array([1.2, 0.9, 1.05])
    _, tau, *_ = orc.constitutive(6, np.diag(s), mu, lam, 0.0, P_(6))
    J = s.prod()
    b = s ** 2
    t = mu * J ** (-2 / 3) * (b - b.mean()) + lam / 2 * (J * J - 1)
    assert np.abs(tau - np.diag(t * s)).max() < 1e-8 * np.abs(tau).max()


@pytest.mark.parametrize("material", [1, 2, 3])
def test_log_stress_diagonal_closed_form(material):
    """StVK (1, 3) and DP (2) stresses on a diagonal F below yield, with
    t = 2 mu eps + lam tr eps, eps = log s: DP gives tau = diag(t); StVK keeps
    the reference's U diag(t) V^T F^T = diag(t * s) (mpm_utils.py:61-66)."""
    mu, lam = 3e4, 5e4
    s = np.array([0.99, 0.985, 0.995])  # compression inside the DP cone
    # a high yield stress (and friction making DP inside the cone) keeps F_e = F
    Fe, tau, *_ = orc.constitutive(material, np.diag(s), mu, lam, 1e9, P_(material, friction_angle=60.0))
    e = np.log(s)
    if material == 2:
        # DP with tr(eps) < 0 can project; check it did not, then the closed form
        assert np.abs(Fe - np.diag(s)).max() < 1e-12
    t = 2 * mu * e + lam * e.sum()
    ref = np.diag(t if material == 2 else t * s)
    assert np.abs(tau - ref).max() < 1e-8 * np.abs(ref).max()


# ---------------------------------------------------------------- return maps
def test_von_mises_projection_onto_yield_surface():
    mu, lam = 3e4, 5e4
    P = P_(1)
    for _ in range(30):
        F = rand_rot(RNG) @ np.diag(RNG.uniform(0.7, 1.3, 3)) @ rand_rot(RNG).T
        ys = RNG.uniform(100, 3000)
        Fe, tau, mu_o, lam_o, ys_o = orc.constitutive(1, F, mu, lam, ys, P)
        s_tr = np.linalg.svd(F, compute_uv=False)
        e_tr = np.log(np.maximum(s_tr, 0.01))
        dev_norm = 2 * mu * np.linalg.norm(e_tr - e_tr.mean())
        s_e = np.linalg.svd(Fe, compute_uv=False)
        e_e = np.log(s_e)
        assert abs(e_e.sum() - e_tr.sum()) < 1e-10                    # volume preserved
        if dev_norm > ys:
            # |dev tau| on the yield surface up to the reference's +1e-6 in the norm
            n_hat = np.linalg.norm(e_tr - e_tr.mean())
            expect = ys * (n_hat / (n_hat + 1e-6))
            assert abs(2 * mu * np.linalg.norm(e_e - e_e.mean()) - expect) < 1e-6 * max(ys, 1)
        else:
            assert np.abs(Fe - F).max() < 1e-14
        assert ys_o == ys and mu_o == mu


def test_plasticine_softening_closed_form():
    mu, lam, ys = 3e4, 5e4, 500.0
    P = P_(5, softening=0.1)
    F = np.diag([1.1, 0.95, 0.97])
    Fe, tau, mu_o, lam_o, ys_o = orc.constitutive(5, F, mu, lam, ys, P)
    e = np.log(np.diag(F))
    eh = e - e.mean()
    n = np.linalg.norm(eh) + 1e-6
    dg = n - ys / (2 * mu)
    assert abs(ys_o - (ys - 0.1 * np.linalg.norm(dg / n * eh))) < 1e-10
    assert ys_o < ys


def test_drucker_prager_three_cases():
    mu, lam = 3e4, 5e4
    P = P_(2, friction_angle=30.0)
    alpha = P.alpha
    c = (3 * lam + 2 * mu) / (2 * mu) * alpha
    # expansion: tr > 0 -> F_e is the rotation U V^T (all singular values 1)
    R1, R2 = rand_rot(RNG), rand_rot(RNG)
    F = R1 @ np.diag([1.1, 1.05, 1.02]) @ R2.T
    Fe, *_ = orc.constitutive(2, F, mu, lam, 0, P)
    assert np.abs(np.linalg.svd(Fe, compute_uv=False) - 1).max() < 1e-12
    assert np.abs(Fe - R1 @ R2.T).max() < 1e-12
    # compression inside the cone: identity
    F = np.diag([0.99, 0.99, 0.99])
    Fe, *_ = orc.constitutive(2, F, mu, lam, 0, P)
    assert np.abs(Fe - F).max() < 1e-14
    # compression with shear: projected onto the cone |eps_hat| + c tr = 0, tr kept
    F = np.diag([1.1, 0.9, 0.97])
    e = np.log(np.diag(F))
    assert np.linalg.norm(e - e.mean()) + c * e.sum() > 0 and e.sum() < 0
    Fe, *_ = orc.constitutive(2, F, mu, lam, 0, P)
    ee = np.log(np.diag(Fe))
    assert abs(ee.sum() - e.sum()) < 1e-12
    assert abs(np.linalg.norm(ee - ee.mean()) + c * ee.sum()) < 1e-12


def test_viscoplastic_return_limits():
    """Material 3 (mpm_utils.py:189-234): with zero viscosity the return is
    the radial return onto |dev s| = sqrt(2/3) Y with the log-volume kept;
    with viscosity the deviatoric stress lies strictly between that surface
    and the trial state and tends to the trial state as viscosity grows."""
    mu, lam, dt = 3e4, 5e4, 1e-4
    for _ in range(20):
        F = rand_rot(RNG) @ np.diag(RNG.uniform(0.8, 1.25, 3)) @ rand_rot(RNG).T
        ys = RNG.uniform(100, 2000)
        s_tr = np.linalg.svd(F, compute_uv=False)
        e_tr = np.log(np.maximum(s_tr, 0.01))
        dev_tr = 2 * mu * np.linalg.norm(e_tr - e_tr.mean())
        yv = np.sqrt(2.0 / 3.0) * ys
        devs = []
        for eta in (0.0, 10.0, 1e9):
            Fe, *_ = orc.constitutive(3, F, mu, lam, ys, P_(3, plastic_viscosity=eta), dt=dt)
            e_e = np.log(np.linalg.svd(Fe, compute_uv=False))
            assert abs(e_e.sum() - e_tr.sum()) < 1e-10
            devs.append(2 * mu * np.linalg.norm(e_e - e_e.mean()))
        if dev_tr > yv:
            assert abs(devs[0] - yv) < 1e-8 * yv
            assert yv < devs[1] < dev_tr
            assert abs(devs[2] - dev_tr) < 1e-6 * dev_tr
        else:
            assert max(abs(d - dev_tr) for d in devs) < 1e-9 * max(dev_tr, 1)


def test_mu_lam():
    E = RNG.uniform(1e4, 1e6, 100)
    nu = RNG.uniform(0.1, 0.45, 100)
    mu, lam = orc.mu_lam(E, nu)
    assert np.allclose(mu, E / (2 * (1 + nu)), rtol=1e-14)
    assert np.allclose(lam, E * nu / ((1 + nu) * (1 - 2 * nu)), rtol=1e-14)


# ---------------------------------------------------------------- substep
def _state(N, n, rng, v=None, C=None, material=4):
    dx = 1.0 / n
    x = rng.uniform(4 * dx, 1 - 4 * dx, (N, 3))
    vv = rng.normal(size=(N, 3)) if v is None else v(x)
    CC = np.zeros((N, 3, 3)) if C is None else np.broadcast_to(C, (N, 3, 3)).copy()
    F = np.tile(np.eye(3), (N, 1, 1)) + 0.05 * rng.normal(size=(N, 3, 3))
    mass = rng.uniform(0.5, 2.0, N) * 1e-4
    vol = mass / 1000
    return orc.OracleState(x, vv, CC, F, np.full(N, 3e3), np.full(N, 5e3), np.full(N, 1e9), mass, vol)


@pytest.mark.parametrize("material", [0, 6, 7])
def test_p2g_conserves_mass_and_momentum(material):
    rng = np.random.Generator(np.random.PCG64(5))
    N, n = 300, 16
    s = _state(N, n, rng, C=rng.normal(size=(3, 3)))
    P = orc.make_params(N, n, material=material, gravity=(0, 0, 0))
    _, Fe, tau, grid = orc.substep(P, [], 0.0, 1e-4, s, return_fe=True)
    assert abs(grid[..., 3].sum() - s.mass.sum()) < 1e-12 * s.mass.sum()
    # internal forces cancel (sum of grad w = 0); the APIC term has zero moment
    mom = (s.mass[:, None] * s.v).sum(0)
    assert np.abs(grid[..., :3].reshape(-1, 3).sum(0) - mom).max() < 1e-12 * np.abs(mom).max() * 10
    assert np.abs(tau).max() > 0


def test_apic_reproduces_affine_fields_exactly():
    """v = A x + b with C = A (snow: no stress, no gravity): after one substep
    v' = A x + b, C' = A, F' = (I + dt A) F at every particle."""
    rng = np.random.Generator(np.random.PCG64(6))
    N, n, dt = 400, 16, 1e-4
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    s = _state(N, n, rng, v=lambda x: x @ A.T + b, C=A)
    P = orc.make_params(N, n, material=4, gravity=(0, 0, 0))
    o = orc.substep(P, [], 0.0, dt, s)
    assert np.abs(o.v - (s.x @ A.T + b)).max() < 1e-10
    assert np.abs(o.C - A[None]).max() < 1e-9
    assert np.abs(o.F - np.einsum("ij,njk->nik", np.eye(3) + dt * A, s.F)).max() < 1e-12
    x_exp = s.x + dt * (s.x @ A.T + b)
    assert np.abs(o.x - x_exp).max() < 1e-12


def test_free_fall_closed_form():
    """Uniform velocity, no stress, gravity g: v_n = v0 + n dt g and
    x_n = x0 + n dt v0 + n(n+1)/2 dt^2 g exactly (symplectic update)."""
    rng = np.random.Generator(np.random.PCG64(7))
    N, n, dt, steps = 200, 32, 1e-3, 30
    v0 = np.array([0.1, -0.2, 0.05])
    s = _state(N, n, rng, v=lambda x: np.tile(v0, (len(x), 1)))
    s.x = np.clip(s.x, 0.35, 0.65)
    x0 = s.x.copy()
    g = np.array([0, 0, -9.8])
    P = orc.make_params(N, n, material=4, gravity=tuple(g))
    gf = np.array([P.gravity[0], P.gravity[1], P.gravity[2]])
    orc.rollout(P, [], 0.0, dt, steps, s)
    assert np.abs(s.v - (v0 + steps * dt * gf)).max() < 1e-11
    xe = x0 + steps * dt * v0 + steps * (steps + 1) / 2 * dt * dt * gf
    assert np.abs(s.x - xe).max() < 1e-11


# ---------------------------------------------------------------- colliders
def _uniform_grid(n, v=(1.0, -2.0, 3.0)):
    g = np.zeros((n, n, n, 4))
    g[..., 3] = 2.0
    g[..., :3] = 2.0 * np.asarray(v)
    return g


def test_sticky_walls_node_set():
    from empm_amd import scenes

    n = 16
    cols = [orc.surface_collider(c["point"], c["normal"]) for c in scenes.sticky_walls(n, 3)]
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    vout = orc.grid_update(P, cols, 0.0, 1e-4, _uniform_grid(n))
    idx = np.indices((n, n, n))
    wall = np.any((idx < 3) | (idx >= n - 3), axis=0)
    assert np.all(vout[wall] == 0)
    assert np.all(vout[~wall] == np.array([1.0, -2.0, 3.0]))


def test_surface_collider_time_window_and_cut():
    n = 16
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    c = orc.surface_collider((0, 0, 0.5), (0, 0, 1), start_time=0.1, end_time=0.2)
    v0 = orc.grid_update(P, [c], 0.05, 1e-4, _uniform_grid(n))
    assert np.all(v0 == np.array([1.0, -2.0, 3.0]))           # before the window
    v1 = orc.grid_update(P, [c], 0.15, 1e-4, _uniform_grid(n))
    z = np.arange(n) / n
    assert np.all(v1[:, :, z < 0.5] == 0) and np.all(v1[:, :, z >= 0.5] != 0)
    cut = orc.surface_collider((0, 0, 0.5), (0, 0, 1), surface="cut")
    v2 = orc.grid_update(P, [cut], 0.0, 1e-4, _uniform_grid(n))
    for k in range(n):
        zk = np.float32(k) * np.float32(1.0 / n)
        if zk >= 0.5:
            continue
        if zk < 0.4 or zk > 0.53:
            assert np.all(v2[:, :, k] == 0)
        else:
            assert np.allclose(v2[:, :, k], np.array([0.3, 0.0, 0.9], np.float32), rtol=1e-6)


def test_cuboid_sets_velocity_and_host_advance():
    n = 16
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    c = orc.cuboid((0.5, 0.5, 0.5), 0.1, (0.2, 0.0, -0.1), start_time=0.0, end_time=1.0)
    v = orc.grid_update(P, [c], 0.0, 1e-3, _uniform_grid(n))
    pos = np.indices((n, n, n)).transpose(1, 2, 3, 0) / n
    inside = np.all(np.abs(pos - 0.5) < 0.1, axis=-1)
    assert np.allclose(v[inside], np.array([0.2, 0.0, -0.1], np.float32))
    assert np.all(v[~inside] == np.array([1.0, -2.0, 3.0]))
    moved = orc.advance_colliders([c], 0.0, 1e-3)[0]
    assert abs(moved.point[0] - np.float32(np.float32(0.5) + 1e-3 * np.float32(0.2))) < 1e-7
    # reset: outside the window the whole grid is zeroed for 15 substeps
    r = orc.cuboid((0.5, 0.5, 0.5), 0.1, (0.2, 0, 0), start_time=0.0, end_time=0.01, reset=1)
    assert np.all(orc.grid_update(P, [r], 0.012, 1e-3, _uniform_grid(n)) == 0)
    assert np.all(orc.grid_update(P, [r], 0.03, 1e-3, _uniform_grid(n)) != 0)


def test_bounding_box_components():
    n = 16
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    g = _uniform_grid(n, v=(-1.0, 1.0, -1.0))
    v = orc.grid_update(P, [orc.bounding_box()], 0.0, 1e-4, g)
    assert np.all(v[:3, :, :, 0] == 0) and np.all(v[3:, :, :, 0] == -1)
    assert np.all(v[:, n - 3:, :, 1] == 0) and np.all(v[:, :n - 3, :, 1] == 1)
    assert np.all(v[:, :, :3, 2] == 0) and np.all(v[:, :, 3:, 2] == -1)


def _geo(kind, geo_type, point, size=(0, 0, 0), radius=0.0, half_height=0.0, normal=(0, 0, 1), velocity=(0, 0, 0),
         rot_axis=(0, 0, 1), angular_velocity=0.0, active=1, start_time=0.0, end_time=999.0):
    c = orc.OracleCollider()
    c.kind, c.geo_type, c.active = kind, geo_type, active
    for i in range(3):
        c.point[i], c.size[i], c.normal[i] = point[i], size[i], normal[i]
        c.velocity[i], c.rot_axis[i] = velocity[i], rot_axis[i]
    c.radius, c.half_height, c.angular_velocity = radius, half_height, angular_velocity
    c.start_time, c.end_time = start_time, end_time
    return c


def test_geometry_collider_index_sets():
    """add_geometry_collider (kind 3): nodes strictly inside a cube, sphere or
    cylinder (axis = normal) are zeroed, nothing else changes; inactive or
    out-of-window colliders do nothing."""
    n = 16
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    pos = np.indices((n, n, n)).transpose(1, 2, 3, 0) / n
    c0 = np.array([0.5, 0.5, 0.5])
    off = pos - c0
    cases = [
        (_geo(3, 1, c0, size=(0.2, 0.1, 0.15)), np.all(np.abs(off) < [0.2, 0.1, 0.15], axis=-1)),
        (_geo(3, 2, c0, radius=0.23), np.linalg.norm(off, axis=-1) < 0.23),
        (_geo(3, 3, c0, radius=0.2, half_height=0.13, normal=(1, 0, 0)),
         (np.abs(off[..., 0]) < 0.13) & (np.linalg.norm(off[..., 1:], axis=-1) < 0.2)),
    ]
    for c, inside in cases:
        v = orc.grid_update(P, [c], 0.0, 1e-4, _uniform_grid(n))
        assert inside.any() and (~inside).any()
        assert np.all(v[inside] == 0) and np.all(v[~inside] == np.array([1.0, -2.0, 3.0]))
    for c in (_geo(3, 2, c0, radius=0.3, active=0), _geo(3, 2, c0, radius=0.3, start_time=1.0)):
        assert np.all(orc.grid_update(P, [c], 0.0, 1e-4, _uniform_grid(n)) == np.array([1.0, -2.0, 3.0]))


def test_grid_geometry_mover_is_a_rigid_motion():
    """set_grid_velocity_on_geometry (kind 4): inside the sphere the node
    velocities form a rigid motion (relative velocity orthogonal to relative
    position, angular speed as given), outside nothing changes."""
    n = 16
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    axis = np.array([1.0, 2.0, 2.0]) / 3.0
    c = _geo(4, 2, (0.5, 0.5, 0.5), radius=0.25, velocity=(0.1, -0.2, 0.3), rot_axis=axis, angular_velocity=2.0)
    v = orc.grid_update(P, [c], 0.0, 1e-4, _uniform_grid(n))
    pos = np.indices((n, n, n)).transpose(1, 2, 3, 0) / n
    inside = np.linalg.norm(pos - 0.5, axis=-1) < 0.25
    assert np.all(v[~inside] == np.array([1.0, -2.0, 3.0]))
    xi, vi = pos[inside], v[inside]
    rng = np.random.Generator(np.random.PCG64(7))
    for _ in range(50):
        a, b = rng.integers(0, len(xi), 2)
        dxv, dvv = xi[a] - xi[b], vi[a] - vi[b]
        assert abs(np.dot(dxv, dvv)) < 1e-6
        # |omega x r| = omega |r| sin(angle(axis, r))
        r = dxv
        expect = 2.0 * np.linalg.norm(np.cross(axis, r))
        assert abs(np.linalg.norm(dvv) - expect) < 1e-5
    centre = np.all(np.isclose(pos, 0.5), axis=-1)
    assert np.allclose(v[centre], [0.1, -0.2, 0.3], atol=1e-7)


def test_mask_collider_zeroes_marked_nodes():
    """enforce_grid_velocity_by_mask (kind 5): nodes with mask >= 1 are
    zeroed, all others keep their velocity."""
    n = 16
    P = orc.make_params(1, n, gravity=(0, 0, 0))
    mask = np.random.Generator(np.random.PCG64(3)).integers(-1, 3, (n, n, n)).astype(np.int32)
    c = orc.OracleCollider()
    c.kind = 5
    c.mask = mask.ctypes.data
    v = orc.grid_update(P, [c], 0.0, 1e-4, _uniform_grid(n))
    assert np.all(v[mask >= 1] == 0) and np.all(v[mask < 1] == np.array([1.0, -2.0, 3.0]))


def test_float32_and_float64_oracles_agree_on_a_rollout():
    from empm_amd import scenes

    sc = scenes.small_scene(n_particles=300, n_grid=16, material=0)
    mu, lam = orc.mu_lam(sc.E, sc.nu)
    P = orc.make_params(sc.n, 16, material=0)
    cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
    res = {}
    for prec in ("f64", "f32"):
        s = orc.OracleState(sc.x, sc.params["v0"], np.zeros((sc.n, 3, 3)), np.tile(np.eye(3), (sc.n, 1, 1)), mu, lam,
                            sc.yield_stress, sc.vol * sc.rho, sc.vol, precision=prec)
        orc.rollout(P, cols, 0.0, sc.dt, 20, s)
        res[prec] = s
    assert np.abs(res["f32"].x - res["f64"].x).max() < 1e-5
    assert np.abs(res["f32"].F - res["f64"].F).max() < 1e-4


# ---------------------------------------------------------------- branches on the caller's path
def _nh_diag_tau(s, mu, lam):
    """Kirchhoff stress of kirchoff_stress_neoHookean (mpm_utils.py:18-47) on a
    diagonal F = diag(s): U = V = I, tau = diag(t_i s_i) with
    t_i = mu J^(-2/3) (s_i^2 - mean s^2) + lam/2 (J^2 - 1)."""
    J = np.prod(s)
    b = s * s
    t = mu * J ** (-2.0 / 3.0) * (b - b.mean()) + lam / 2 * (J * J - 1)
    return np.diag(t * s)


@pytest.mark.parametrize("scale", [0.98, 0.5, 1.0, 1.1])
def test_grid_damping_closed_form(scale):
    """add_damping_via_grid (mpm_utils.py:722-734) as the substep applies it
    (mpm_solver_diff.py:244-251: only when scale < 1): after normalisation
    and gravity, v_out = v - (1 - s) v = s (mv/m + dt g).  Fails for damping
    before gravity, for a factor (1 - s) in place of s, or for a scale >= 1
    that is applied anyway."""
    n, dt = 16, 1e-3
    g = np.array([0.5, 0.0, -9.8])
    P = orc.make_params(1, n, gravity=tuple(g), grid_v_damping_scale=scale)
    gf = np.array([P.gravity[0], P.gravity[1], P.gravity[2]])
    v = orc.grid_update(P, [], 0.0, dt, _uniform_grid(n))
    base = np.array([1.0, -2.0, 3.0]) + dt * gf
    expect = scale * base if scale < 1 else base
    assert np.abs(v - expect).max() < 1e-14
    # empty nodes stay at zero (the mass threshold of :432)
    empty = np.zeros((n, n, n, 4))
    assert np.all(orc.grid_update(P, [], 0.0, dt, empty) == 0)


def test_g2p_position_clamp_lands_on_the_bound():
    """g2p_differentiable's clamp (mpm_utils.py:553-565): x' = clamp(x + dt v',
    2 dx, lim - 2 dx) per axis.  A uniform velocity field without stress or
    gravity gives v' = v exactly, so particles pushed past either bound land
    on it and the others move by dt v.  Fails for bounds of dx or 3 dx, or
    for a clamp applied before the velocity update."""
    n, dt = 16, 1e-3
    dx = 1.0 / n
    v0 = np.array([5.0, -5.0, 0.1])
    x = np.array([[1 - 2 * dx - 0.002, 0.5, 0.5],    # x crosses the upper bound
                  [0.5, 2 * dx + 0.002, 0.5],         # y crosses the lower bound
                  [0.5, 0.5, 0.5],                    # stays inside
                  [1 - 2 * dx - 0.002, 2 * dx + 0.002, 0.5]])
    N = len(x)
    s = orc.OracleState(x, np.tile(v0, (N, 1)), np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)),
                        np.full(N, 3e3), np.full(N, 5e3), np.full(N, 1e9), np.full(N, 1e-4), np.full(N, 1e-7))
    P = orc.make_params(N, n, material=4, gravity=(0, 0, 0))
    o = orc.substep(P, [], 0.0, dt, s)
    assert np.abs(o.v - v0).max() < 1e-12
    lo, hi = 2 * np.float32(1.0 / n), 1.0 - 2 * np.float32(1.0 / n)
    expect = np.clip(x + dt * v0, lo, hi)
    assert np.abs(o.x - expect).max() < 1e-12
    assert o.x[0, 0] == hi and o.x[1, 1] == lo and o.x[3, 0] == hi and o.x[3, 1] == lo
    assert np.abs(o.x[2] - (x[2] + dt * v0)).max() < 1e-12


@pytest.mark.parametrize("r", [0.0, 0.3, 1.0, -0.0005, -0.01])
def test_rpic_blend_closed_form(r):
    """The RPIC blend of p2g_apic_with_stress (mpm_utils.py:396-411):
    C~ = (1 - r) C + r/2 (C - C^T); C~ = 0 when r < -0.001 (PIC).  One
    particle at rest without stress scatters mv_i = w_i m C~ (x_i - x_p); the
    quadratic B-spline's second moment sum_i w_i d_i d_i^T = dx^2/4 I then
    recovers C~ = 4 / (m dx^2) sum_i mv_i d_i^T from the P2G grid.  r = 1
    gives the skew part (C - C^T)/2; a transposed or sign-flipped blend fails."""
    n = 16
    dx = 1.0 / n
    rng = np.random.Generator(np.random.PCG64(41))
    A = rng.normal(size=(3, 3))
    xp = np.array([[0.4713, 0.5291, 0.5077]])
    m = 2.5e-4
    s = orc.OracleState(xp, np.zeros((1, 3)), A[None], np.eye(3)[None], np.full(1, 3e3), np.full(1, 5e3),
                        np.full(1, 1e9), np.full(1, m), np.full(1, 1e-7))
    P = orc.make_params(1, n, material=4, gravity=(0, 0, 0), rpic_damping=r)
    _, _, _, grid = orc.substep(P, [], 0.0, 1e-4, s, return_fe=True)
    idx = np.argwhere(grid[..., 3] > 0)
    assert len(idx) == 27
    d = idx * P.dx - xp[0]
    Ct = 4.0 / (m * dx * dx) * np.einsum("ki,kj->ij", grid[idx[:, 0], idx[:, 1], idx[:, 2], :3], d)
    expect = np.zeros((3, 3)) if r < -0.001 else (1 - r) * A + r / 2 * (A - A.T)
    assert np.abs(Ct - expect).max() < 1e-9 * max(1.0, np.abs(A).max())
    if r == 1.0:
        assert np.abs(Ct + Ct.T).max() < 1e-9


@pytest.mark.parametrize("material", [1, 5])
def test_hardening_update_closed_form(material):
    """hardening == 1 (mpm_utils.py:124-127 for material 1, :178-181 after the
    softening of material 5): Y' = Y (- softening |dgamma/n eps_hat|) +
    2 mu' xi dgamma, dgamma = n - Y / (2 mu), n = |eps_hat| + 1e-6, eps the log
    of the clamped singular values (numpy's SVD here).  The elastic F is the
    one without hardening; fails for a missing factor 2, a dgamma/n in place
    of dgamma, or a hardening that also moves F_e."""
    mu, lam, ys, xi = 3e4, 5e4, 400.0, 0.35
    F = rand_rot(RNG) @ np.diag([1.12, 0.93, 0.98]) @ rand_rot(RNG).T
    P0 = P_(material, hardening=0.0, xi=xi, softening=0.1)
    P1 = P_(material, hardening=1.0, xi=xi, softening=0.1)
    Fe0, _, _, _, ys0 = orc.constitutive(material, F, mu, lam, ys, P0)
    Fe1, _, mu1, lam1, ys1 = orc.constitutive(material, F, mu, lam, ys, P1)
    e = np.log(np.maximum(np.linalg.svd(F, compute_uv=False), 0.01))
    eh = e - e.mean()
    n = np.linalg.norm(eh) + 1e-6
    dg = n - ys / (2 * mu)
    assert 2 * mu * np.linalg.norm(eh) > ys                      # the state yields
    soft = 0.1 * np.linalg.norm(dg / n * eh) if material == 5 else 0.0
    assert abs(ys0 - (ys - soft)) < 1e-9 * ys
    assert abs(ys1 - (ys - soft + 2 * mu1 * xi * dg)) < 1e-9 * ys
    assert mu1 == mu and lam1 == lam
    assert np.abs(Fe1 - Fe0).max() < 1e-14


def test_material7_is_the_von_mises_return_with_the_nh_stress():
    """Id 7 (Neo-Hookean + von Mises, the workload's material; not a reference
    id) composes von_mises_return_mapping (mpm_utils.py:87-130) with
    kirchoff_stress_neoHookean (:18-47).  Below yield: F_e = F and tau is the
    NH closed form on a diagonal F.  Above yield on a diagonal F: F_e =
    diag(exp(eps - dgamma/n eps_hat)) and tau the NH closed form at F_e.  An
    id 7 that dispatched to the StVK or FCR stress, or skipped the return
    map, fails."""
    mu, lam = 3e4, 5e4
    s = np.array([1.06, 0.97, 0.99])
    Fe, tau, *_ = orc.constitutive(7, np.diag(s), mu, lam, 1e9, P_(7))
    assert np.abs(Fe - np.diag(s)).max() < 1e-15
    ref = _nh_diag_tau(s, mu, lam)
    assert np.abs(tau - ref).max() < 1e-9 * np.abs(ref).max()
    ys = 300.0
    Fe, tau, mu_o, lam_o, ys_o = orc.constitutive(7, np.diag(s), mu, lam, ys, P_(7))
    e = np.log(s)
    eh = e - e.mean()
    n = np.linalg.norm(eh) + 1e-6
    dg = n - ys / (2 * mu)
    assert 2 * mu * np.linalg.norm(eh) > ys
    se = np.exp(e - dg / n * eh)
    assert np.abs(Fe - np.diag(se)).max() < 1e-12
    ref = _nh_diag_tau(se, mu, lam)
    assert np.abs(tau - ref).max() < 1e-9 * np.abs(ref).max()
    assert (mu_o, lam_o, ys_o) == (mu, lam, ys)
    # and the same F_e as id 1 on a general F
    F = rand_rot(RNG) @ np.diag([1.1, 0.92, 1.0]) @ rand_rot(RNG).T
    Fe7, *_ = orc.constitutive(7, F, mu, lam, ys, P_(7))
    Fe1, *_ = orc.constitutive(1, F, mu, lam, ys, P_(1))
    assert np.abs(Fe7 - Fe1).max() == 0


def test_frozen_grid_mass_substep():
    """oracle_substep_frozen_mass at the substep's own grid mass is the plain
    substep (up to the order of the OpenMP atomics in P2G); at a scaled mass
    the grid velocities (hence v') of a stress-free, gravity-free state scale
    by its inverse."""
    rng = np.random.Generator(np.random.PCG64(8))
    N, n = 200, 16
    s = _state(N, n, rng)
    P = orc.make_params(N, n, material=4, gravity=(0, 0, 0))
    o, _, _, grid = orc.substep(P, [], 0.0, 1e-4, s.copy(), return_fe=True)
    o2 = orc.substep(P, [], 0.0, 1e-4, s.copy(), m_fixed=grid[..., 3].copy())
    for k in ("x", "v", "C", "F"):
        assert np.abs(getattr(o, k) - getattr(o2, k)).max() < 1e-13 * max(1.0, np.abs(getattr(o, k)).max())
    o3 = orc.substep(P, [], 0.0, 1e-4, s.copy(), m_fixed=2.0 * grid[..., 3])
    assert np.abs(o3.v - o.v / 2).max() < 1e-12 * np.abs(o.v).max()
    # a rollout fed its own mass history reproduces itself
    mh = np.zeros((5, n ** 3))
    a, b = s.copy(), s.copy()
    orc.rollout(P, [], 0.0, 1e-4, 5, a, m_hist_out=mh)
    orc.rollout(P, [], 0.0, 1e-4, 5, b, m_hist_in=mh)
    assert np.abs(a.x - b.x).max() < 1e-13 and np.abs(a.F - b.F).max() < 1e-13


# ---------------------------------------------------------------- restart_and_compute_F_C
def _lattice(n, k0, k1):
    """Particles on a regular lattice of spacing dx/2 at (k + 1/2) dx/2: the
    lattice is symmetric about every grid node, so a node whose B-spline
    support lies inside it averages the lattice positions to its own."""
    dx = 1.0 / n
    c = (np.arange(k0, k1) + 0.5) * dx / 2
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), c


def test_restart_F_translation_gives_identity():
    # a uniform displacement u makes every touched node's velocity u, and
    # sum_i u (x) grad w_i = u (x) sum_i grad w_i = 0: F_trial = I exactly
    # (a mistake taking w for grad w, or the displacement unweighted, fails)
    n, N = 16, 400
    x = RNG.uniform(0.25, 0.75, (N, 3))
    u = np.array([0.013, -0.021, 0.007])
    P = orc.make_params(N, n)
    F = orc.restart_F(P, x, x + u, RNG.uniform(1e-4, 1e-3, N), np.tile(np.eye(3), (N, 1, 1)))
    # (not exactly I: a node reached only by a stencil's far edge can hold
    # less than the 1e-15 mass threshold and keep v = 0, as in the
    # reference; its weight gradient is then tiny as well)
    assert np.abs(F - np.eye(3)).max() < 1e-8


def test_restart_F_linear_field_interior_closed_form():
    # displacement A x + b on the symmetric lattice: interior nodes take
    # v_i = A x_i + b, and sum_i x_i (x) grad w_i = I gives F_trial = I + A
    # exactly for particles whose stencil nodes are all interior; a missing
    # mass normalisation scales A by the node mass and fails
    n = 16
    dx = 1.0 / n
    x, c = _lattice(n, 4, 28)
    N = len(x)
    A = np.array([[0.01, 0.02, -0.005], [0.003, -0.01, 0.004], [-0.02, 0.006, 0.015]])
    b = np.array([0.001, -0.002, 0.0005])
    P = orc.make_params(N, n)
    F = orc.restart_F(P, x, x + x @ A.T + b, np.full(N, 1e-3), np.tile(np.eye(3), (N, 1, 1)))
    inner = np.all((x > c.min() + 3 * dx) & (x < c.max() - 3 * dx), axis=1)
    assert inner.sum() >= 500
    assert np.abs(F[inner] - (np.eye(3) + A)).max() < 1e-12
    # at the lattice's faces the average is one-sided: F differs from I + A there
    assert np.abs(F[~inner] - (np.eye(3) + A)).max() > 1e-3


def test_restart_F_leaves_unselected_particles():
    # only particle_selection == 0 is restarted (mpm_utils.py:847,884), and
    # the unselected ones neither deposit mass nor read the grid
    n, N = 16, 300
    x = RNG.uniform(0.3, 0.7, (N, 3))
    target = x + RNG.normal(0, 0.01, (N, 3))
    F0 = np.tile(np.eye(3), (N, 1, 1)) + RNG.normal(0, 0.1, (N, 3, 3))
    sel = np.zeros(N, np.int32)
    sel[::3] = 1
    P = orc.make_params(N, n)
    mass = RNG.uniform(1e-4, 1e-3, N)
    F = orc.restart_F(P, x, target, mass, F0, sel)
    assert np.array_equal(F[sel == 1], F0[sel == 1])
    keep = sel == 0
    Pk = orc.make_params(int(keep.sum()), n)
    Fk = orc.restart_F(Pk, x[keep], target[keep], mass[keep], F0[keep])
    assert np.abs(F[keep] - Fk).max() < 1e-14
