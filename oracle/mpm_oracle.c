/*
 * CPU oracle for the MLS-MPM substep — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference substep (rai-opensource/empm,
 * phys_sim/diff_simulator/warp_mpm/).  It is built twice from this one file:
 * REAL=double (liboracle_f64.so, the FP64 rerun used for error reports and
 * finite differences) and REAL=float (liboracle_f32.so, the FP32 restatement
 * used as the parity reference and as the CPU baseline).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to phys_sim/diff_simulator/warp_mpm/ unless stated otherwise.
 *
 * Pinning (how each part is checked independently of this code, in
 * tests/test_oracle.py): B-spline identities (partition of unity, first
 * moment, gradient sums), SVD against numpy.linalg.svd with U,V in SO(3),
 * zero-stress states and uniform-scale closed forms of every constitutive
 * model, return-map post-conditions (yield surface, volume, DP cone),
 * mass/momentum conservation of P2G, exact APIC affine reproduction, free
 * fall closed form, collider index sets.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif

#if defined(ORACLE_F32)
#define RLOG logf
#define REXP expf
#define RSQRT sqrtf
#define RFABS fabsf
#define RPOW powf
#else
#define RLOG log
#define REXP exp
#define RSQRT sqrt
#define RFABS fabs
#define RPOW pow
#endif

typedef REAL real;

/* Scalars of MPMModelStruct (mpm_data_structure.py:390-422).  Always double
 * at the boundary; converted to REAL where the reference computes in float. */
typedef struct {
    int n_particles;
    int n_grid;
    int material;
    int pad0;
    double grid_lim;
    double dx;
    double inv_dx;
    double rpic_damping;
    double grid_v_damping_scale;
    double gravity[3];
    double alpha;
    double hardening;
    double xi;
    double plastic_viscosity;
    double softening;
} oracle_params;

/* One registered grid post-process (collider).  kind:
 *   0 add_surface_collider           (mpm_solver_diff.py:437-531)
 *   1 set_velocity_on_cuboid         (mpm_solver_diff.py:814-869)
 *   2 add_bounding_box               (mpm_solver_diff.py:871-938)
 *   3 add_geometry_collider          (mpm_solver_diff.py:533-683)
 *   4 set_grid_velocity_on_geometry  (mpm_solver_diff.py:685-804)
 *   5 enforce_grid_velocity_by_mask  (mpm_solver_diff.py:1264-1289)
 * Geometry is float32 in the reference (Dirichlet_collider,
 * mpm_data_structure.py:502-537), so masks are evaluated in float here in
 * both builds: the node sets are discrete decisions that must not depend on
 * the oracle's precision. */
typedef struct {
    int kind;
    int surface_type;
    int geo_type;
    int reset;
    int active;
    int pad0;
    double start_time;
    double end_time;
    double point[3];
    double normal[3];
    double size[3];
    double velocity[3];
    double rot_axis[3];
    double angular_velocity;
    double radius;
    double half_height;
    double friction;
    const int *mask;
} oracle_collider;

/* ------------------------------------------------------------------ */
/* 3x3 helpers, row-major m[r*3+c]                                     */
/* ------------------------------------------------------------------ */
static void mat_mul(const real *a, const real *b, real *o)
{
    real t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            t[r * 3 + c] = a[r * 3 + 0] * b[0 * 3 + c] + a[r * 3 + 1] * b[1 * 3 + c] +
                           a[r * 3 + 2] * b[2 * 3 + c];
    memcpy(o, t, sizeof(t));
}
static void mat_mul_bt(const real *a, const real *b, real *o) /* a * b^T */
{
    real t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            t[r * 3 + c] = a[r * 3 + 0] * b[c * 3 + 0] + a[r * 3 + 1] * b[c * 3 + 1] +
                           a[r * 3 + 2] * b[c * 3 + 2];
    memcpy(o, t, sizeof(t));
}
static real mat_det(const real *m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}
/* U diag(d) V^T */
static void udvt(const real *U, const real *d, const real *V, real *o)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            o[r * 3 + c] = U[r * 3 + 0] * d[0] * V[c * 3 + 0] + U[r * 3 + 1] * d[1] * V[c * 3 + 1] +
                           U[r * 3 + 2] * d[2] * V[c * 3 + 2];
}

/* ------------------------------------------------------------------ */
/* SVD with U, V in SO(3): the convention wp.svd3 is used under at      */
/* mpm_utils.py:92,140,197,244,641 (R = U V^T must be a rotation,       */
/* mpm_utils.py:13).  Singular values are ordered by decreasing         */
/* magnitude and only the last may be negative (det F < 0).             */
/* Method: cyclic Jacobi on F^T F run to convergence, eigenvectors give */
/* V, then a Gram-Schmidt QR of B = F V gives U and the signed sigma.   */
/* ------------------------------------------------------------------ */
int oracle_svd3(const real *F, real *U, real *s, real *V)
{
    real A[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            A[r * 3 + c] = F[0 * 3 + r] * F[0 * 3 + c] + F[1 * 3 + r] * F[1 * 3 + c] +
                           F[2 * 3 + r] * F[2 * 3 + c];
    real W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 60; ++sweep) {
        real off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        real dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
#if defined(ORACLE_F32)
        if (off == 0 || off <= dia * 1e-16f)
            break;
#else
        if (off == 0 || off <= dia * 1e-34)
            break;
#endif
        for (int pq = 0; pq < 3; ++pq) {
            int p = pq == 2 ? 1 : 0;
            int q = pq == 0 ? 1 : 2;
            real apq = A[p * 3 + q];
            if (apq == 0)
                continue;
            real theta = (A[q * 3 + q] - A[p * 3 + p]) / (2 * apq);
            real t = (theta >= 0 ? 1 : -1) / (RFABS(theta) + RSQRT(theta * theta + 1));
            real c = 1 / RSQRT(t * t + 1);
            real sn = t * c;
            /* A <- J^T A J with J = rotation in (p,q): J_pp=c, J_pq=s, J_qp=-s, J_qq=c */
            for (int k = 0; k < 3; ++k) { /* A <- A J (columns p,q) */
                real akp = A[k * 3 + p], akq = A[k * 3 + q];
                A[k * 3 + p] = c * akp - sn * akq;
                A[k * 3 + q] = sn * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) { /* A <- J^T A (rows p,q) */
                real apk = A[p * 3 + k], aqk = A[q * 3 + k];
                A[p * 3 + k] = c * apk - sn * aqk;
                A[q * 3 + k] = sn * apk + c * aqk;
            }
            for (int k = 0; k < 3; ++k) {
                real wkp = W[k * 3 + p], wkq = W[k * 3 + q];
                W[k * 3 + p] = c * wkp - sn * wkq;
                W[k * 3 + q] = sn * wkp + c * wkq;
            }
        }
    }
    /* sort eigenvalues descending; each swap of two columns also negates one
     * so that det W stays +1 */
    real ev[3] = {A[0], A[4], A[8]};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2 - i; ++j)
            if (ev[j] < ev[j + 1]) {
                real te = ev[j];
                ev[j] = ev[j + 1];
                ev[j + 1] = te;
                for (int k = 0; k < 3; ++k) {
                    real tw = W[k * 3 + j];
                    W[k * 3 + j] = W[k * 3 + j + 1];
                    W[k * 3 + j + 1] = -tw;
                }
            }
    if (mat_det(W) < 0)
        for (int k = 0; k < 3; ++k)
            W[k * 3 + 2] = -W[k * 3 + 2];
    /* B = F V; QR by Gram-Schmidt with u3 = u1 x u2 */
    real B[9];
    mat_mul(F, W, B);
    real b1[3] = {B[0], B[3], B[6]}, b2[3] = {B[1], B[4], B[7]}, b3[3] = {B[2], B[5], B[8]};
    real n1 = RSQRT(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
    real u1[3], u2[3], u3[3];
    if (n1 > 0) {
        for (int k = 0; k < 3; ++k)
            u1[k] = b1[k] / n1;
    } else {
        u1[0] = 1;
        u1[1] = 0;
        u1[2] = 0;
    }
    real d12 = u1[0] * b2[0] + u1[1] * b2[1] + u1[2] * b2[2];
    real w2[3] = {b2[0] - d12 * u1[0], b2[1] - d12 * u1[1], b2[2] - d12 * u1[2]};
    real n2 = RSQRT(w2[0] * w2[0] + w2[1] * w2[1] + w2[2] * w2[2]);
#if defined(ORACLE_F32)
    if (n2 > 1e-30f) {
#else
    if (n2 > 1e-200) {
#endif
        for (int k = 0; k < 3; ++k)
            u2[k] = w2[k] / n2;
    } else { /* any unit vector orthogonal to u1 */
        real a[3] = {0, 0, 0};
        int m = RFABS(u1[0]) < RFABS(u1[1]) ? (RFABS(u1[0]) < RFABS(u1[2]) ? 0 : 2)
                                            : (RFABS(u1[1]) < RFABS(u1[2]) ? 1 : 2);
        a[m] = 1;
        real d = u1[0] * a[0] + u1[1] * a[1] + u1[2] * a[2];
        for (int k = 0; k < 3; ++k)
            a[k] -= d * u1[k];
        real na = RSQRT(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int k = 0; k < 3; ++k)
            u2[k] = a[k] / na;
    }
    u3[0] = u1[1] * u2[2] - u1[2] * u2[1];
    u3[1] = u1[2] * u2[0] - u1[0] * u2[2];
    u3[2] = u1[0] * u2[1] - u1[1] * u2[0];
    s[0] = n1;
    s[1] = u2[0] * b2[0] + u2[1] * b2[1] + u2[2] * b2[2];
    s[2] = u3[0] * b3[0] + u3[1] * b3[1] + u3[2] * b3[2];
    for (int k = 0; k < 3; ++k) {
        U[k * 3 + 0] = u1[k];
        U[k * 3 + 1] = u2[k];
        U[k * 3 + 2] = u3[k];
    }
    memcpy(V, W, sizeof(W));
    return 0;
}

/* ------------------------------------------------------------------ */
/* Constitutive models (mpm_utils.py:8-84)                             */
/* ------------------------------------------------------------------ */
/* kirchoff_stress_FCR, mpm_utils.py:8-15 */
static void stress_fcr(const real *F, const real *U, const real *V, real J, real mu, real lam, real *tau)
{
    real R[9], FmR[9];
    mat_mul_bt(U, V, R);
    for (int i = 0; i < 9; ++i)
        FmR[i] = F[i] - R[i];
    mat_mul_bt(FmR, F, tau);
    for (int i = 0; i < 9; ++i)
        tau[i] *= 2 * mu;
    for (int d = 0; d < 3; ++d)
        tau[d * 4] += lam * J * (J - 1);
}
/* U diag(t) V^T F^T, the common tail of mpm_utils.py:42-47,61-66,83-84 */
static void stress_udvtft(const real *F, const real *U, const real *V, const real *t, real *tau)
{
    real G[9];
    udvt(U, t, V, G);
    mat_mul_bt(G, F, tau);
}
/* kirchoff_stress_neoHookean, mpm_utils.py:18-47 */
static void stress_nh(const real *F, const real *U, const real *V, real J, const real *sig, real mu, real lam,
                      real *tau)
{
    real b[3] = {sig[0] * sig[0], sig[1] * sig[1], sig[2] * sig[2]};
    real m = (b[0] + b[1] + b[2]) / 3;
    real jp = RPOW(J, (real)(-2.0 / 3.0));
    real t[3];
    for (int i = 0; i < 3; ++i)
        t[i] = mu * jp * (b[i] - m) + lam / 2 * (J * J - 1);
    stress_udvtft(F, U, V, t, tau);
}
/* kirchoff_stress_StVK, mpm_utils.py:50-66 */
static void stress_stvk(const real *F, const real *U, const real *V, const real *sig, real mu, real lam, real *tau)
{
    real e[3];
    for (int i = 0; i < 3; ++i)
        e[i] = RLOG(sig[i] > (real)0.01 ? sig[i] : (real)0.01);
    real tr = e[0] + e[1] + e[2];
    real t[3];
    for (int i = 0; i < 3; ++i)
        t[i] = 2 * mu * e[i] + lam * tr;
    stress_udvtft(F, U, V, t, tau);
}
/* kirchoff_stress_drucker_prager, mpm_utils.py:69-84 (no clamp: log of a
 * negative singular value is NaN, as in the reference) */
static void stress_dp(const real *F, const real *U, const real *V, const real *sig, real mu, real lam, real *tau)
{
    real L = RLOG(sig[0]) + RLOG(sig[1]) + RLOG(sig[2]);
    real t[3];
    for (int i = 0; i < 3; ++i)
        t[i] = 2 * mu * RLOG(sig[i]) * (1 / sig[i]) + lam * L * (1 / sig[i]);
    stress_udvtft(F, U, V, t, tau);
}

/* von_mises_return_mapping, mpm_utils.py:87-130; with damage (material 5),
 * mpm_utils.py:133-186.  Writes yield stress / mu / lam in place like the
 * reference (mpm_utils.py:126-129,162-167,180-183). */
static void return_vm(const real *Ftr, real *mu, real *lam, real *ys, const oracle_params *P, int damage,
                      real *Fe)
{
    real U[9], V[9], s0[3];
    oracle_svd3(Ftr, U, s0, V);
    real e[3];
    for (int i = 0; i < 3; ++i)
        e[i] = RLOG(s0[i] > (real)0.01 ? s0[i] : (real)0.01);
    real temp = (e[0] + e[1] + e[2]) / 3;
    real tr = e[0] + e[1] + e[2];
    real tau[3], st = 0;
    for (int i = 0; i < 3; ++i) {
        tau[i] = 2 * (*mu) * e[i] + (*lam) * tr;
        st += tau[i];
    }
    real cd[3] = {tau[0] - st / 3, tau[1] - st / 3, tau[2] - st / 3};
    real cn = RSQRT(cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2]);
    if (cn > *ys) {
        if (damage && *ys <= 0) {
            memcpy(Fe, Ftr, 9 * sizeof(real));
            return;
        }
        real eh[3] = {e[0] - temp, e[1] - temp, e[2] - temp};
        real ehn = RSQRT(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]) + (real)1e-6;
        real dg = ehn - *ys / (2 * (*mu));
        for (int i = 0; i < 3; ++i)
            e[i] = e[i] - (dg / ehn) * eh[i];
        if (damage) {
            real q[3] = {(dg / ehn) * eh[0], (dg / ehn) * eh[1], (dg / ehn) * eh[2]};
            *ys = *ys - (real)P->softening * RSQRT(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (*ys <= 0) {
                *mu = 0;
                *lam = 0;
            }
        }
        real se[3] = {REXP(e[0]), REXP(e[1]), REXP(e[2])};
        udvt(U, se, V, Fe);
        if (P->hardening == 1)
            *ys = *ys + 2 * (*mu) * (real)P->xi * dg;
        return;
    }
    memcpy(Fe, Ftr, 9 * sizeof(real));
}

/* viscoplasticity_return_mapping_with_StVK, mpm_utils.py:189-234 */
static void return_visco(const real *Ftr, real mu, real ys, const oracle_params *P, real dt, real *Fe)
{
    real U[9], V[9], s0[3];
    oracle_svd3(Ftr, U, s0, V);
    real sg[3], b[3], e[3];
    for (int i = 0; i < 3; ++i) {
        sg[i] = s0[i] > (real)0.01 ? s0[i] : (real)0.01;
        b[i] = sg[i] * sg[i];
        e[i] = RLOG(sg[i]);
    }
    real tr = e[0] + e[1] + e[2];
    real eh[3] = {e[0] - tr / 3, e[1] - tr / 3, e[2] - tr / 3};
    real s[3] = {2 * mu * eh[0], 2 * mu * eh[1], 2 * mu * eh[2]};
    real sn = RSQRT(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    real y = sn - RSQRT((real)(2.0 / 3.0)) * ys;
    if (y > 0) {
        real muh = mu * (b[0] + b[1] + b[2]) / 3;
        real snew = sn - y / (1 + (real)P->plastic_viscosity / (2 * muh * dt));
        real en[3], se[3];
        for (int i = 0; i < 3; ++i) {
            en[i] = 1 / (2 * mu) * ((snew / sn) * s[i]) + tr / 3;
            se[i] = REXP(en[i]);
        }
        udvt(U, se, V, Fe);
        return;
    }
    memcpy(Fe, Ftr, 9 * sizeof(real));
}

/* sand_return_mapping, mpm_utils.py:237-274 */
static void return_sand(const real *Ftr, real mu, real lam, const oracle_params *P, real *Fe)
{
    real U[9], V[9], s0[3];
    oracle_svd3(Ftr, U, s0, V);
    real e[3];
    for (int i = 0; i < 3; ++i) {
        real a = RFABS(s0[i]);
        e[i] = RLOG(a > (real)1e-14 ? a : (real)1e-14);
    }
    real tr = e[0] + e[1] + e[2];
    real eh[3] = {e[0] - tr / 3, e[1] - tr / 3, e[2] - tr / 3};
    real ehn = RSQRT(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    real dg = ehn + (3 * lam + 2 * mu) / (2 * mu) * tr * (real)P->alpha;
    if (dg <= 0) {
        memcpy(Fe, Ftr, 9 * sizeof(real));
        return;
    }
    if (tr > 0) {
        mat_mul_bt(U, V, Fe);
        return;
    }
    real sn[3];
    for (int i = 0; i < 3; ++i)
        sn[i] = REXP(e[i] - eh[i] * (dg / ehn));
    udvt(U, sn, V, Fe);
}

/* compute_stress_from_F_trial, mpm_utils.py:603-665.  Material 7 is the
 * Neo-Hookean + von Mises composition (vM return map of :87-130 followed by
 * the NH stress of :18-47); it is not a reference material id. */
int oracle_constitutive(int material, const real *Ftr, real *mu, real *lam, real *ys, const oracle_params *P,
                        double dt, real *Fe, real *tau_sym)
{
    switch (material) {
    case 1:
    case 7:
        return_vm(Ftr, mu, lam, ys, P, 0, Fe);
        break;
    case 2:
        return_sand(Ftr, *mu, *lam, P, Fe);
        break;
    case 3:
        return_visco(Ftr, *mu, *ys, P, (real)dt, Fe);
        break;
    case 5:
        return_vm(Ftr, mu, lam, ys, P, 1, Fe);
        break;
    default:
        memcpy(Fe, Ftr, 9 * sizeof(real));
        break;
    }
    real J = mat_det(Fe);
    real U[9], V[9], sg[3], tau[9];
    oracle_svd3(Fe, U, sg, V);
    memset(tau, 0, sizeof(tau));
    switch (material) {
    case 0:
    case 5:
        stress_fcr(Fe, U, V, J, *mu, *lam, tau);
        break;
    case 1:
    case 3:
        stress_stvk(Fe, U, V, sg, *mu, *lam, tau);
        break;
    case 2:
        stress_dp(Fe, U, V, sg, *mu, *lam, tau);
        break;
    case 6:
    case 7:
        stress_nh(Fe, U, V, J, sg, *mu, *lam, tau);
        break;
    default: /* 4 (snow): no stress branch matches, tau = 0 */
        break;
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            tau_sym[r * 3 + c] = (tau[r * 3 + c] + tau[c * 3 + r]) / 2;
    return 0;
}

/* compute_mu_lam_from_E_nu, mpm_utils.py:277-283 */
void oracle_mu_lam(int n, const real *E, const real *nu, real *mu, real *lam)
{
    for (int p = 0; p < n; ++p) {
        mu[p] = E[p] / (2 * (1 + nu[p]));
        lam[p] = E[p] * nu[p] / ((1 + nu[p]) * (1 - 2 * nu[p]));
    }
}

/* ------------------------------------------------------------------ */
/* Quadratic B-spline weights (mpm_utils.py:368-383 and 294-303).       */
/* The reference builds w = wp.mat33(wa2/2, 0.75-wb2, wc2/2) from       */
/* vectors, i.e. columns, so w[d][i] = N_i(f_d).                        */
/* ------------------------------------------------------------------ */
static void bspline(const real *xp, real inv_dx, int *base, real *fx, real w[3][3], real dw[3][3])
{
    for (int d = 0; d < 3; ++d) {
        real g = xp[d] * inv_dx;
        base[d] = (int)(g - (real)0.5); /* wp.int truncates toward zero */
        real f = g - (real)base[d];
        fx[d] = f;
        real wa = (real)1.5 - f, wb = f - 1, wc = f - (real)0.5;
        w[d][0] = wa * wa * (real)0.5;
        w[d][1] = (real)0.75 - wb * wb;
        w[d][2] = wc * wc * (real)0.5;
        dw[d][0] = f - (real)1.5;
        dw[d][1] = -2 * (f - 1);
        dw[d][2] = f - (real)0.5;
    }
}

/* Exported for the B-spline identity tests. */
void oracle_bspline(const real *xp, double inv_dx, int *base, real *fx, real *w, real *dw)
{
    real W[3][3], D[3][3];
    bspline(xp, (real)inv_dx, base, fx, W, D);
    memcpy(w, W, sizeof(W));
    memcpy(dw, D, sizeof(D));
}

/* ------------------------------------------------------------------ */
/* Grid post-processes (colliders), one node at a time.                */
/* ------------------------------------------------------------------ */
static void collide_node(const oracle_collider *c, int ix, int iy, int iz, int n, float time, float dt, float dxf,
                         real *v)
{
    int in_window = time >= (float)c->start_time && time < (float)c->end_time;
    float pos[3] = {(float)ix * dxf, (float)iy * dxf, (float)iz * dxf};
    float off[3] = {pos[0] - (float)c->point[0], pos[1] - (float)c->point[1], pos[2] - (float)c->point[2]};
    switch (c->kind) {
    case 0: { /* add_surface_collider kernel, mpm_solver_diff.py:473-528 */
        if (!in_window)
            return;
        float d = off[0] * (float)c->normal[0] + off[1] * (float)c->normal[1] + off[2] * (float)c->normal[2];
        if (d < 0.0f) {
            if (c->surface_type == 11) {
                float z = (float)iz * dxf;
                if (z < 0.4f || z > 0.53f) {
                    v[0] = v[1] = v[2] = 0;
                } else {
                    v[0] = v[0] * (real)0.3f;
                    v[1] = 0;
                    v[2] = v[2] * (real)0.3f;
                }
            } else {
                /* sticky, slip and separate all end in v = 0 (the slip and
                 * friction results are overwritten, mpm_solver_diff.py:525-527) */
                v[0] = v[1] = v[2] = 0;
            }
        }
        return;
    }
    case 1: { /* set_velocity_on_cuboid kernel, mpm_solver_diff.py:835-858 */
        if (in_window) {
            if (fabsf(off[0]) < (float)c->size[0] && fabsf(off[1]) < (float)c->size[1] &&
                fabsf(off[2]) < (float)c->size[2]) {
                v[0] = (real)(float)c->velocity[0];
                v[1] = (real)(float)c->velocity[1];
                v[2] = (real)(float)c->velocity[2];
            }
        } else if (c->reset == 1) {
            if (time < (float)c->end_time + 15.0f * dt)
                v[0] = v[1] = v[2] = 0;
        }
        return;
    }
    case 2: { /* add_bounding_box kernel, mpm_solver_diff.py:878-935 */
        if (!in_window)
            return;
        const int pad = 3;
        int idx[3] = {ix, iy, iz};
        for (int d = 0; d < 3; ++d) {
            if (idx[d] < pad && v[d] < 0)
                v[d] = 0;
            if (idx[d] >= n - pad && v[d] > 0)
                v[d] = 0;
        }
        return;
    }
    case 3: { /* add_geometry_collider kernel, mpm_solver_diff.py:612-672 */
        if (!(in_window && c->active))
            return;
        int inside = 0;
        if (c->geo_type == 1) {
            inside = fabsf(off[0]) < (float)c->size[0] && fabsf(off[1]) < (float)c->size[1] &&
                     fabsf(off[2]) < (float)c->size[2];
        } else if (c->geo_type == 2) {
            inside = sqrtf(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]) < (float)c->radius;
        } else if (c->geo_type == 3) {
            float ax[3] = {(float)c->normal[0], (float)c->normal[1], (float)c->normal[2]};
            float pr = off[0] * ax[0] + off[1] * ax[1] + off[2] * ax[2];
            float rd[3] = {off[0] - pr * ax[0], off[1] - pr * ax[1], off[2] - pr * ax[2]};
            inside = fabsf(pr) < (float)c->half_height &&
                     sqrtf(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]) < (float)c->radius;
        } else if (c->geo_type == 4) {
            float d = off[0] * (float)c->normal[0] + off[1] * (float)c->normal[1] + off[2] * (float)c->normal[2];
            if (d < 0.0f) {
                if (c->surface_type == 11) {
                    float z = (float)iz * dxf;
                    if (z < 0.4f || z > 0.53f) {
                        v[0] = v[1] = v[2] = 0;
                    } else {
                        v[0] = v[0] * (real)0.3f;
                        v[1] = 0;
                        v[2] = v[2] * (real)0.3f;
                    }
                } else {
                    v[0] = v[1] = v[2] = 0;
                }
            }
            return;
        }
        if (inside)
            v[0] = v[1] = v[2] = 0;
        return;
    }
    case 4: { /* set_grid_velocity_on_geometry kernel, mpm_solver_diff.py:756-793 */
        if (in_window && c->active) {
            float r[3] = {(float)c->rot_axis[0], (float)c->rot_axis[1], (float)c->rot_axis[2]};
            float av = (float)c->angular_velocity;
            float cr[3] = {r[1] * off[2] - r[2] * off[1], r[2] * off[0] - r[0] * off[2], r[0] * off[1] - r[1] * off[0]};
            float vv[3] = {(float)c->velocity[0] + av * cr[0], (float)c->velocity[1] + av * cr[1],
                           (float)c->velocity[2] + av * cr[2]};
            int inside = 0;
            if (c->geo_type == 1) {
                inside = fabsf(off[0]) < (float)c->size[0] && fabsf(off[1]) < (float)c->size[1] &&
                         fabsf(off[2]) < (float)c->size[2];
            } else if (c->geo_type == 2) {
                inside = sqrtf(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]) < (float)c->radius;
            } else if (c->geo_type == 3) {
                float ax[3] = {(float)c->normal[0], (float)c->normal[1], (float)c->normal[2]};
                float pr = off[0] * ax[0] + off[1] * ax[1] + off[2] * ax[2];
                float rd[3] = {off[0] - pr * ax[0], off[1] - pr * ax[1], off[2] - pr * ax[2]};
                inside = fabsf(pr) < (float)c->half_height &&
                         sqrtf(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]) < (float)c->radius;
            } else if (c->geo_type == 4) {
                float d = off[0] * (float)c->normal[0] + off[1] * (float)c->normal[1] + off[2] * (float)c->normal[2];
                inside = d < 0.0f;
            }
            if (inside) {
                v[0] = vv[0];
                v[1] = vv[1];
                v[2] = vv[2];
            }
        } else if (c->reset == 1) {
            if (time < (float)c->end_time + 15.0f * dt)
                v[0] = v[1] = v[2] = 0;
        }
        return;
    }
    case 5: /* enforce_grid_velocity_by_mask kernel, mpm_solver_diff.py:1275-1286 */
        if (c->mask[((size_t)ix * n + iy) * n + iz] >= 1)
            v[0] = v[1] = v[2] = 0;
        return;
    default:
        return;
    }
}

/* Host-side point advance of moving colliders (mpm_solver_diff.py:795-801,
 * 860-866): float32 point += dt * velocity while the window is open. */
void oracle_advance_colliders(oracle_collider *cols, int ncol, double time, double dt)
{
    for (int k = 0; k < ncol; ++k) {
        oracle_collider *c = &cols[k];
        if (c->kind != 1 && c->kind != 4)
            continue;
        if (time >= (double)(float)c->start_time && time < (double)(float)c->end_time)
            for (int d = 0; d < 3; ++d)
                c->point[d] = (double)(float)((double)(float)c->point[d] + dt * (double)(float)c->velocity[d]);
    }
}

/* ------------------------------------------------------------------ */
/* One substep: MPMWARPDiff.p2g2p_differentiable, mpm_solver_diff.py:165-291 */
/* State s (x, v, C, F_trial) -> state s+1.  mu/lam/ys are updated in place */
/* by material 5 as in the reference.  grid: n^3*4 scratch (mv.xyz, m).     */
/* Fe_out / tau_out (optional) receive the return-mapped F and the stress. */
/* m_fixed (optional, n^3): the grid mass the normalisation divides by and  */
/* tests against 1e-15, in place of the mass P2G just accumulated (grid[.3] */
/* still receives that mass).  Holding it at an unperturbed substep's value */
/* gives the function whose derivative drops the grid-mass adjoint, as the  */
/* reference Tape does (grid_m carries no grad: mpm_data_structure.py:      */
/* 100-105 with warp_utils.py:116-120), so central differences of it check  */
/* that gradient convention.                                                */
/* ------------------------------------------------------------------ */
static int substep_impl(const oracle_params *P, const oracle_collider *cols, int ncol, double time, double dt,
                        const real *x, const real *v, const real *C, const real *Ftr, real *mu, real *lam, real *ys,
                        const real *mass, const real *vol, const int *sel, real *x1, real *v1, real *C1, real *F1,
                        real *grid, real *Fe_out, real *tau_out, const real *m_fixed)
{
    const int N = P->n_particles, n = P->n_grid;
    const size_t G = (size_t)n * n * n;
    const real inv_dx = (real)P->inv_dx, dx = (real)P->dx, rdt = (real)dt;
    real *Fe = (real *)malloc(sizeof(real) * 9 * (size_t)(N > 0 ? N : 1));
    real *tau = (real *)malloc(sizeof(real) * 9 * (size_t)(N > 0 ? N : 1));
    real *vout = (real *)malloc(sizeof(real) * 3 * G);
    if (!Fe || !tau || !vout) {
        free(Fe);
        free(tau);
        free(vout);
        return -1;
    }
    /* 1. zero_grid, mpm_utils.py:286-291 */
    memset(grid, 0, sizeof(real) * 4 * G);
    memset(vout, 0, sizeof(real) * 3 * G);

    /* 3. compute_stress_from_F_trial, mpm_utils.py:603-665 */
#pragma omp parallel for schedule(static)
    for (int p = 0; p < N; ++p) {
        if (sel && sel[p] != 0) {
            memcpy(&Fe[9 * p], &Ftr[9 * p], 9 * sizeof(real));
            memset(&tau[9 * p], 0, 9 * sizeof(real));
            continue;
        }
        oracle_constitutive(P->material, &Ftr[9 * p], &mu[p], &lam[p], &ys[p], P, dt, &Fe[9 * p], &tau[9 * p]);
    }

    /* 4. p2g_apic_with_stress, mpm_utils.py:358-423 */
    const real r = (real)P->rpic_damping;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < N; ++p) {
        if (sel && sel[p] != 0)
            continue;
        int base[3];
        real fx[3], w[3][3], dw[3][3];
        bspline(&x[3 * p], inv_dx, base, fx, w, dw);
        const real *Cp = &C[9 * p];
        real Ct[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Ct[i * 3 + j] = (1 - r) * Cp[i * 3 + j] + r / 2 * (Cp[i * 3 + j] - Cp[j * 3 + i]);
        if (r < (real)-0.001)
            memset(Ct, 0, sizeof(Ct));
        const real *S = &tau[9 * p];
        const real mp = mass[p], vp = vol[p];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    int ix = base[0] + a, iy = base[1] + b, iz = base[2] + c;
                    /* out-of-range nodes (possible only at t0, see the
                     * add_particles bounds check) are skipped */
                    if (ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n)
                        continue;
                    real dpos[3] = {((real)a - fx[0]) * dx, ((real)b - fx[1]) * dx, ((real)c - fx[2]) * dx};
                    real wt = w[0][a] * w[1][b] * w[2][c];
                    real dwt[3] = {dw[0][a] * w[1][b] * w[2][c] * inv_dx, w[0][a] * dw[1][b] * w[2][c] * inv_dx,
                                   w[0][a] * w[1][b] * dw[2][c] * inv_dx};
                    real add[3];
                    for (int i = 0; i < 3; ++i) {
                        real cd = Ct[i * 3 + 0] * dpos[0] + Ct[i * 3 + 1] * dpos[1] + Ct[i * 3 + 2] * dpos[2];
                        real ef = -vp * (S[i * 3 + 0] * dwt[0] + S[i * 3 + 1] * dwt[1] + S[i * 3 + 2] * dwt[2]);
                        add[i] = wt * mp * (v[3 * p + i] + cd) + rdt * ef;
                    }
                    real *g = &grid[4 * (((size_t)ix * n + iy) * n + iz)];
                    for (int i = 0; i < 3; ++i) {
#pragma omp atomic
                        g[i] += add[i];
                    }
#pragma omp atomic
                    g[3] += wt * mp;
                }
    }

    /* 5.-7. grid_normalization_and_gravity (mpm_utils.py:427-438),
     * add_damping_via_grid (mpm_utils.py:722-734), colliders in
     * registration order (mpm_solver_diff.py:264-278) */
    const float tf = (float)time, dtf = (float)dt, dxf = (float)P->dx;
    const real ds = (real)P->grid_v_damping_scale;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < n; ++iz) {
                size_t id = ((size_t)ix * n + iy) * n + iz;
                real *g = &grid[4 * id];
                real vv[3] = {0, 0, 0};
                const real gm = m_fixed ? m_fixed[id] : g[3];
                if (gm > (real)1e-15) {
                    real im = 1 / gm;
                    for (int i = 0; i < 3; ++i)
                        vv[i] = g[i] * im + rdt * (real)P->gravity[i];
                }
                if (P->grid_v_damping_scale < 1.0)
                    for (int i = 0; i < 3; ++i)
                        vv[i] = vv[i] - (1 - ds) * vv[i];
                for (int k = 0; k < ncol; ++k)
                    collide_node(&cols[k], ix, iy, iz, n, tf, dtf, dxf, vv);
                memcpy(&vout[3 * id], vv, sizeof(vv));
            }

    /* 8. g2p_differentiable, mpm_utils.py:502-579 */
    const real amin = dx * 2, amax = (real)P->grid_lim - dx * 2;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < N; ++p) {
        if (sel && sel[p] != 0) { /* partial_clone defaults, mpm_data_structure.py:356-387 */
            memset(&x1[3 * p], 0, 3 * sizeof(real));
            memset(&v1[3 * p], 0, 3 * sizeof(real));
            memset(&C1[9 * p], 0, 9 * sizeof(real));
            for (int i = 0; i < 9; ++i)
                F1[9 * p + i] = (i % 4 == 0) ? 1 : 0;
            continue;
        }
        int base[3];
        real fx[3], w[3][3], dw[3][3];
        bspline(&x[3 * p], inv_dx, base, fx, w, dw);
        real nv[3] = {0, 0, 0}, nC[9] = {0}, nF[9] = {0};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    int ix = base[0] + a, iy = base[1] + b, iz = base[2] + c;
                    real gv[3] = {0, 0, 0};
                    if (!(ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n))
                        memcpy(gv, &vout[3 * (((size_t)ix * n + iy) * n + iz)], sizeof(gv));
                    real dpos[3] = {(real)a - fx[0], (real)b - fx[1], (real)c - fx[2]};
                    real wt = w[0][a] * w[1][b] * w[2][c];
                    real dwt[3] = {dw[0][a] * w[1][b] * w[2][c] * inv_dx, w[0][a] * dw[1][b] * w[2][c] * inv_dx,
                                   w[0][a] * w[1][b] * dw[2][c] * inv_dx};
                    for (int i = 0; i < 3; ++i) {
                        nv[i] += gv[i] * wt;
                        for (int j = 0; j < 3; ++j) {
                            nC[i * 3 + j] += gv[i] * dpos[j] * (wt * inv_dx * 4);
                            nF[i * 3 + j] += gv[i] * dwt[j];
                        }
                    }
                }
        for (int i = 0; i < 3; ++i) {
            v1[3 * p + i] = nv[i];
            real xn = x[3 * p + i] + rdt * nv[i];
            x1[3 * p + i] = xn < amin ? amin : (xn > amax ? amax : xn);
        }
        memcpy(&C1[9 * p], nC, sizeof(nC));
        real IF[9];
        for (int i = 0; i < 9; ++i)
            IF[i] = (i % 4 == 0 ? 1 : 0) + nF[i] * rdt;
        mat_mul(IF, &Fe[9 * p], &F1[9 * p]);
    }
    if (Fe_out)
        memcpy(Fe_out, Fe, sizeof(real) * 9 * (size_t)N);
    if (tau_out)
        memcpy(tau_out, tau, sizeof(real) * 9 * (size_t)N);
    free(Fe);
    free(tau);
    free(vout);
    return 0;
}

int oracle_substep(const oracle_params *P, const oracle_collider *cols, int ncol, double time, double dt,
                   const real *x, const real *v, const real *C, const real *Ftr, real *mu, real *lam, real *ys,
                   const real *mass, const real *vol, const int *sel, real *x1, real *v1, real *C1, real *F1,
                   real *grid, real *Fe_out, real *tau_out)
{
    return substep_impl(P, cols, ncol, time, dt, x, v, C, Ftr, mu, lam, ys, mass, vol, sel, x1, v1, C1, F1, grid,
                        Fe_out, tau_out, 0);
}

/* restart_and_compute_F_C (mpm_solver_diff.py:1214-1262) without its last
 * step (x = target, set_vec3_to_vec3): zero_grid; set_F_C_p2g
 * (mpm_utils.py:843-876), the mass and mass-weighted displacement
 * target - x of every selected particle into the grid;
 * grid_normalization_and_gravity with dt = 0 (mpm_utils.py:427-438);
 * set_F_C_g2p (mpm_utils.py:879-919), F_trial = I + sum_i v_i (x) dweight_i
 * for every selected particle (F_out of other particles is not written). */
int oracle_restart_F(const oracle_params *P, const real *x, const real *target, const real *mass, const int *sel,
                     real *F_out)
{
    const int N = P->n_particles, n = P->n_grid;
    const size_t G = (size_t)n * n * n;
    const real inv_dx = (real)P->inv_dx;
    real *grid = (real *)calloc(4 * G, sizeof(real));
    if (!grid)
        return -1;
    for (int p = 0; p < N; ++p) {
        if (sel && sel[p] != 0)
            continue;
        int base[3];
        real fx[3], w[3][3], dw[3][3];
        bspline(&x[3 * p], inv_dx, base, fx, w, dw);
        const real disp[3] = {target[3 * p] - x[3 * p], target[3 * p + 1] - x[3 * p + 1],
                              target[3 * p + 2] - x[3 * p + 2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    const int ix = base[0] + a, iy = base[1] + b, iz = base[2] + c;
                    if (ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n)
                        continue;
                    const real wt = w[0][a] * w[1][b] * w[2][c];
                    real *g = &grid[4 * (((size_t)ix * n + iy) * n + iz)];
                    for (int i = 0; i < 3; ++i)
                        g[i] += wt * mass[p] * disp[i];
                    g[3] += wt * mass[p];
                }
    }
    for (size_t id = 0; id < G; ++id) {
        real *g = &grid[4 * id];
        if (g[3] > (real)1e-15) {
            const real im = 1 / g[3];
            for (int i = 0; i < 3; ++i)
                g[i] = g[i] * im; /* + dt * gravity with dt = 0 */
        } else {
            g[0] = g[1] = g[2] = 0;
        }
    }
    for (int p = 0; p < N; ++p) {
        if (sel && sel[p] != 0)
            continue;
        int base[3];
        real fx[3], w[3][3], dw[3][3];
        bspline(&x[3 * p], inv_dx, base, fx, w, dw);
        real nF[9] = {0};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    const int ix = base[0] + a, iy = base[1] + b, iz = base[2] + c;
                    if (ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n)
                        continue;
                    const real *gv = &grid[4 * (((size_t)ix * n + iy) * n + iz)];
                    const real dwt[3] = {dw[0][a] * w[1][b] * w[2][c] * inv_dx, w[0][a] * dw[1][b] * w[2][c] * inv_dx,
                                         w[0][a] * w[1][b] * dw[2][c] * inv_dx};
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            nF[i * 3 + j] += gv[i] * dwt[j];
                }
        for (int i = 0; i < 9; ++i)
            F_out[9 * p + i] = (i % 4 == 0 ? 1 : 0) + nF[i];
    }
    free(grid);
    return 0;
}

/* oracle_substep with the grid mass of the normalisation held at m_fixed. */
int oracle_substep_frozen_mass(const oracle_params *P, const oracle_collider *cols, int ncol, double time, double dt,
                               const real *x, const real *v, const real *C, const real *Ftr, real *mu, real *lam,
                               real *ys, const real *mass, const real *vol, const int *sel, real *x1, real *v1,
                               real *C1, real *F1, real *grid, const real *m_fixed)
{
    return substep_impl(P, cols, ncol, time, dt, x, v, C, Ftr, mu, lam, ys, mass, vol, sel, x1, v1, C1, F1, grid, 0,
                        0, m_fixed);
}

/* The grid stage alone (steps 5-7 of the substep) on a given (mv, m) grid:
 * exported so the collider node sets can be checked directly. */
void oracle_grid_update(const oracle_params *P, const oracle_collider *cols, int ncol, double time, double dt,
                        const real *grid, real *vout)
{
    const int n = P->n_grid;
    const float tf = (float)time, dtf = (float)dt, dxf = (float)P->dx;
    const real ds = (real)P->grid_v_damping_scale, rdt = (real)dt;
    for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < n; ++iz) {
                size_t id = ((size_t)ix * n + iy) * n + iz;
                const real *g = &grid[4 * id];
                real vv[3] = {0, 0, 0};
                if (g[3] > (real)1e-15) {
                    real im = 1 / g[3];
                    for (int i = 0; i < 3; ++i)
                        vv[i] = g[i] * im + rdt * (real)P->gravity[i];
                }
                if (P->grid_v_damping_scale < 1.0)
                    for (int i = 0; i < 3; ++i)
                        vv[i] = vv[i] - (1 - ds) * vv[i];
                for (int k = 0; k < ncol; ++k)
                    collide_node(&cols[k], ix, iy, iz, n, tf, dtf, dxf, vv);
                memcpy(&vout[3 * id], vv, sizeof(vv));
            }
}

/* A rollout of n_steps substeps from state (x, v, C, F), in place, with the
 * per-substep host collider advance and time += dt (mpm_solver_diff.py:291),
 * as the reference's host loops run it (interface.py:137-145,
 * warp_mpm_wrapper.py:278-306).  Optional: m_hist_out (n_steps * n^3)
 * receives each substep's P2G grid mass; m_hist_in (same shape) holds the
 * grid mass of every substep's normalisation at the values given (see
 * oracle_substep_frozen_mass).  Returns the final time (-1 on failure). */
static double rollout_impl(const oracle_params *P, oracle_collider *cols, int ncol, double t0, double dt, int n_steps,
                           real *x, real *v, real *C, real *F, real *mu, real *lam, real *ys, const real *mass,
                           const real *vol, const int *sel, real *m_hist_out, const real *m_hist_in)
{
    const int N = P->n_particles, n = P->n_grid;
    const size_t G = (size_t)n * n * n;
    real *x1 = (real *)malloc(sizeof(real) * 3 * (size_t)N), *v1 = (real *)malloc(sizeof(real) * 3 * (size_t)N);
    real *C1 = (real *)malloc(sizeof(real) * 9 * (size_t)N), *F1 = (real *)malloc(sizeof(real) * 9 * (size_t)N);
    real *grid = (real *)malloc(sizeof(real) * 4 * G);
    double t = t0;
    if (!x1 || !v1 || !C1 || !F1 || !grid) {
        t = -1;
        goto done;
    }
    for (int s = 0; s < n_steps; ++s) {
        if (substep_impl(P, cols, ncol, t, dt, x, v, C, F, mu, lam, ys, mass, vol, sel, x1, v1, C1, F1, grid, 0, 0,
                         m_hist_in ? m_hist_in + (size_t)s * G : 0) != 0) {
            t = -1;
            goto done;
        }
        if (m_hist_out)
            for (size_t i = 0; i < G; ++i)
                m_hist_out[(size_t)s * G + i] = grid[4 * i + 3];
        oracle_advance_colliders(cols, ncol, t, dt);
        memcpy(x, x1, sizeof(real) * 3 * (size_t)N);
        memcpy(v, v1, sizeof(real) * 3 * (size_t)N);
        memcpy(C, C1, sizeof(real) * 9 * (size_t)N);
        memcpy(F, F1, sizeof(real) * 9 * (size_t)N);
        t = t + dt;
    }
done:
    free(x1);
    free(v1);
    free(C1);
    free(F1);
    free(grid);
    return t;
}

double oracle_rollout(const oracle_params *P, oracle_collider *cols, int ncol, double t0, double dt, int n_steps,
                      real *x, real *v, real *C, real *F, real *mu, real *lam, real *ys, const real *mass,
                      const real *vol, const int *sel)
{
    return rollout_impl(P, cols, ncol, t0, dt, n_steps, x, v, C, F, mu, lam, ys, mass, vol, sel, 0, 0);
}

double oracle_rollout_mass_hist(const oracle_params *P, oracle_collider *cols, int ncol, double t0, double dt,
                                int n_steps, real *x, real *v, real *C, real *F, real *mu, real *lam, real *ys,
                                const real *mass, const real *vol, const int *sel, real *m_hist_out,
                                const real *m_hist_in)
{
    return rollout_impl(P, cols, ncol, t0, dt, n_steps, x, v, C, F, mu, lam, ys, mass, vol, sel, m_hist_out,
                        m_hist_in);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int k)
{
#ifdef _OPENMP
    if (k > 0)
        omp_set_num_threads(k);
#else
    (void)k;
#endif
}
