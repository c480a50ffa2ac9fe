/* Test infrastructure only: a driver that runs the CPU oracle under
 * AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle
 * sanitize`; tests/test_oracle_sanitize.py builds and runs it).  It includes
 * the oracle's single source, so every static function is in the checked
 * build, and drives each entry point the tests use: the SVD, every material's
 * return map and stress, substeps of small random scenes with every kind of
 * grid post-process (including the mask and the cuboid reset window),
 * partially selected particles, particles at the grid's edges, the restart
 * of F and a short rollout with the grid-mass history.  Exit status 0 and no
 * sanitizer report mean the oracle read and wrote only what it owns; the
 * SANITIZE_SELFTEST build plants a read past an array to show that the
 * sanitizer is live. */
#include "mpm_oracle.c"

#include <stdio.h>

static unsigned long long rng_state = 0x9E3779B97F4A7C15ull;

static double urand(void) /* uniform in [0, 1) (splitmix64) */
{
    unsigned long long z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void params(oracle_params *P, int N, int n, int material)
{
    memset(P, 0, sizeof(*P));
    P->n_particles = N;
    P->n_grid = n;
    P->material = material;
    P->grid_lim = 1.0;
    P->dx = 1.0 / n;
    P->inv_dx = n;
    P->rpic_damping = material == 3 ? 0.5 : 0.0;
    P->grid_v_damping_scale = 0.98;
    P->gravity[2] = -9.8;
    P->alpha = 0.3;
    P->hardening = material == 1 ? 1.0 : 0.0;
    P->xi = 0.2;
    P->plastic_viscosity = 0.5;
    P->softening = 0.1;
}

static oracle_collider collider(int kind, int geo_type)
{
    oracle_collider c;
    memset(&c, 0, sizeof(c));
    c.kind = kind;
    c.geo_type = geo_type;
    c.active = 1;
    c.start_time = 0.0;
    c.end_time = 999.0;
    for (int i = 0; i < 3; ++i) {
        c.point[i] = 0.5;
        c.size[i] = 0.05;
        c.rot_axis[i] = i == 2 ? 1.0 : 0.0;
    }
    c.normal[2] = 1.0;
    c.velocity[0] = 0.4;
    c.angular_velocity = 3.0;
    c.radius = 0.07;
    c.half_height = 0.05;
    c.friction = 0.3;
    return c;
}

int main(void)
{
    int failures = 0;
    /* SVD and every material's constitutive update on random F */
    for (int t = 0; t < 200; ++t) {
        real F[9], U[9], s[3], V[9];
        for (int k = 0; k < 9; ++k)
            F[k] = (real)((k % 4 == 0 ? 1.0 : 0.0) + 0.4 * (urand() - 0.5));
        failures += oracle_svd3(F, U, s, V) != 0;
        for (int mat = 0; mat <= 7; ++mat) {
            oracle_params P;
            params(&P, 1, 16, mat);
            real mu = 4000, lam = 6000, ys = 50, Fe[9], tau[9];
            failures += oracle_constitutive(mat, F, &mu, &lam, &ys, &P, 1e-4, Fe, tau) != 0;
        }
    }
    /* substeps of random scenes, every collider kind, per material */
    const int n = 16, N = 700;
    real *x = calloc(3 * N, sizeof(real)), *v = calloc(3 * N, sizeof(real)), *C = calloc(9 * N, sizeof(real));
    real *F = calloc(9 * N, sizeof(real)), *mu = calloc(N, sizeof(real)), *lam = calloc(N, sizeof(real));
    real *ys = calloc(N, sizeof(real)), *mass = calloc(N, sizeof(real)), *vol = calloc(N, sizeof(real));
    real *x1 = calloc(3 * N, sizeof(real)), *v1 = calloc(3 * N, sizeof(real)), *C1 = calloc(9 * N, sizeof(real));
    real *F1 = calloc(9 * N, sizeof(real)), *Fe = calloc(9 * N, sizeof(real)), *tau = calloc(9 * N, sizeof(real));
    real *grid = calloc((size_t)4 * n * n * n, sizeof(real)), *tgt = calloc(3 * N, sizeof(real));
    real *mh = calloc((size_t)8 * n * n * n, sizeof(real));
    int *sel = calloc(N, sizeof(int)), *mask = calloc((size_t)n * n * n, sizeof(int));
    for (int i = 0; i < n * n * n; ++i) mask[i] = (i % 7) == 0;
#ifdef SANITIZE_SELFTEST
    /* a planted read one past the end of x: the sanitizer must stop here */
    failures += x[3 * N + (int)(urand() * 0.5)] != 0;
#endif
    for (int mat = 0; mat <= 7; ++mat) {
        for (int p = 0; p < N; ++p) {
            for (int d = 0; d < 3; ++d) {
                /* the first particles sit at the grid's edges (clamped stencils) */
                x[3 * p + d] = (real)(p < 8 ? ((p >> d) & 1 ? 0.995 : 0.005) : 0.3 + 0.4 * urand());
                v[3 * p + d] = (real)(urand() - 0.5);
            }
            for (int k = 0; k < 9; ++k) {
                C[9 * p + k] = (real)(0.2 * (urand() - 0.5));
                F[9 * p + k] = (real)((k % 4 == 0 ? 1.0 : 0.0) + 0.1 * (urand() - 0.5));
            }
            mu[p] = 4000, lam[p] = 6000, ys[p] = 50;
            mass[p] = 1e-3, vol[p] = 1e-6;
            sel[p] = p % 5 == 0;
        }
        oracle_params P;
        params(&P, N, n, mat);
        oracle_collider cols[10];
        int nc = 0;
        cols[nc++] = collider(0, 0);
        cols[0].surface_type = mat % 3; /* sticky, slip, separate */
        cols[nc++] = collider(1, 0);
        cols[1].reset = 1;
        cols[nc++] = collider(2, 0);
        for (int g = 1; g <= 4; ++g) cols[nc++] = collider(3, g);
        cols[nc++] = collider(4, 2);
        cols[nc++] = collider(5, 0);
        cols[nc - 1].mask = mask;
        cols[nc++] = collider(0, 0);
        cols[nc - 1].surface_type = 11; /* cut */
        failures += oracle_substep(&P, cols, nc, 0.0, 1e-4, x, v, C, F, mu, lam, ys, mass, vol, sel, x1, v1, C1, F1,
                                   grid, Fe, tau) != 0;
        for (int i = 0; i < n * n * n; ++i) mh[i] = grid[4 * i + 3]; /* the substep's node masses */
        failures += oracle_substep_frozen_mass(&P, cols, nc, 0.0, 1e-4, x, v, C, F, mu, lam, ys, mass, vol, sel, x1,
                                               v1, C1, F1, grid, mh) != 0;
        for (int p = 0; p < 3 * N; ++p) tgt[p] = x[p] + (real)(0.002 * (urand() - 0.5));
        failures += oracle_restart_F(&P, x, tgt, mass, sel, F1) != 0;
        oracle_advance_colliders(cols, nc, 0.0, 1e-4);
        oracle_rollout_mass_hist(&P, cols, nc, 0.0, 1e-4, 4, x, v, C, F, mu, lam, ys, mass, vol, sel, mh, NULL);
        oracle_rollout(&P, cols, nc, 4e-4, 1e-4, 3, x, v, C, F, mu, lam, ys, mass, vol, NULL);
    }
    free(x), free(v), free(C), free(F), free(mu), free(lam), free(ys), free(mass), free(vol);
    free(x1), free(v1), free(C1), free(F1), free(Fe), free(tau), free(grid), free(tgt), free(mh);
    free(sel), free(mask);
    printf("oracle sanitizer check: %d failing calls\n", failures);
    return failures != 0;
}
