/*
 * empm.h — C-ABI of libempm.so, the MI355X (gfx950) implementation of the
 * differentiable MLS-MPM substep of rai-opensource/empm and its adjoint.
 *
 * The reference runs this path as Warp kernels launched from Python; every
 * entry point below replaces one of those launches (or a host loop of them)
 * and says which.  Reference paths are relative to
 * phys_sim/diff_simulator/warp_mpm/.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (hipMalloc'd or torch CUDA
 *    storage) unless marked "host".  Particle arrays are AoS float32 exactly
 *    as the reference's torch/Warp views: x, v are (N,3); C, F are (N,3,3)
 *    row-major.  Grid arrays are owned by the context.
 *  - The caller owns every buffer it passes; the library never frees caller
 *    memory.  All device allocation happens in empm_create.
 *  - Every call is stream-ordered on `stream` (a hipStream_t, NULL = the
 *    default stream).  A context is not thread-safe; distinct contexts are.
 *  - Return 0 on success or a negative EMPM_E_* code; the message is in
 *    empm_last_error() (thread-local).  No exception crosses the ABI.
 */
#ifndef EMPM_H
#define EMPM_H

#ifdef __cplusplus
extern "C" {
#endif

#define EMPM_ABI_VERSION 8
#define EMPM_MAX_COLLIDERS 16

enum {
    EMPM_OK = 0,
    EMPM_E_INVALID_ARG = -1,
    EMPM_E_HIP = -2,
    EMPM_E_OOM = -3,
    EMPM_E_UNSUPPORTED_MATERIAL = -4,
    EMPM_E_NONFINITE = -5,
    EMPM_E_UNSUPPORTED = -6
};

/* Failure bits of one simulation (empm_read_status) */
#define EMPM_STATUS_NONFINITE 1 /* a substep produced a non-finite x, v or F_trial */
#define EMPM_STATUS_CFL 2       /* a particle moved more than a cell in one substep (|v| dt > dx) */

/* Adjoint flags */
#define EMPM_ADJ_EXACT 0
/* Drop the grid-mass adjoint, as the reference Tape does (grid_m has no
 * grad buffer: mpm_data_structure.py:100-105 with warp_utils.py:116-120). */
#define EMPM_ADJ_DROP_GRID_MASS 1

typedef struct empm_ctx empm_ctx;

typedef struct {
    int device;        /* HIP device ordinal */
    int n_grid;        /* grid resolution n (n^3 nodes) */
    int max_particles; /* capacity for per-particle scratch */
    int pad0;
} empm_create_desc;

/* MPMModelStruct (mpm_data_structure.py:390-422) plus the per-particle mass,
 * volume and selection of MPMStateStruct (mpm_data_structure.py:13-39).
 * material: 0 jelly (FCR), 1 metal (vM + StVK), 2 sand (DP), 3 foam
 * (viscoplastic + StVK), 4 snow (no stress), 5 plasticine (vM with damage +
 * FCR), 6 neo-hookean, 7 neo-hookean + von Mises (composition of 1's return
 * map with 6's stress; not a reference id). */
typedef struct {
    int n_particles;
    int n_grid;
    int material;
    int pad0;
    float grid_lim;
    float dx;
    float inv_dx;
    float rpic_damping;
    float grid_v_damping_scale;
    float gravity[3];
    float alpha;
    float hardening;
    float xi;
    float plastic_viscosity;
    float softening;
    float *mu;           /* N; written in place by material 5 */
    float *lam;          /* N; written in place by material 5 */
    float *yield_stress; /* N; written in place by material 5 / hardening */
    const float *mass;   /* N */
    const float *vol;    /* N */
    const int *selection; /* N or NULL; only selection[p]==0 is simulated */
} empm_model;

/* One particle state: the fields p2g2p_differentiable reads and writes
 * (particle_x, particle_v, particle_C, particle_F_trial).  Fe is optional:
 * when non-NULL, a forward substep reading this state writes the
 * return-mapped F there (the reference's particle_F, written by
 * compute_stress_from_F_trial at mpm_utils.py:614-631); when NULL a context
 * scratch buffer is used.  empm_rollout_adj reads it back for substeps whose
 * grid was recorded (empm_grid_rec); other adjoint uses ignore it. */
typedef struct {
    float *x;  /* N*3 */
    float *v;  /* N*3 */
    float *C;  /* N*9 */
    float *F;  /* N*9, F_trial */
    float *Fe; /* N*9 or NULL */
    /* 3*N or NULL: the yield stress, mu and lam (N floats each, in that
     * order) that the substep reading this state starts from.  Used only for
     * simulations whose return map rewrites them (material 5's softening,
     * mpm_utils.py:162-167; hardening == 1, :124-127,178-181), for which they
     * are state: a forward substep records its starting values here before
     * updating the model arrays, a replay starts from states[0].mstate, and
     * an adjoint substep reads them (the model arrays hold only the latest
     * values).  The reference keeps them in its model arrays only, which its
     * Tape cannot rewind. */
    float *mstate;
} empm_particles;

/* The grid of one forward substep, recorded for reverse mode so that the
 * adjoint does not recompute P2G and the boundary pass: the reference keeps
 * grid_m / grid_v_in / grid_v_out of every substep alive for its Tape
 * (interface.py:137-145 with mpm_data_structure.py:356-387).  `bins`
 * (optional) also keeps the substep's particle binning, so the reverse
 * sweep does not re-bin. */
typedef struct {
    float *mv_m;  /* n^3*4: (m v, m) after P2G, node-major float4 */
    float *v_out; /* n^3*4: grid velocity after the boundary pass (w unused) */
    int *bins;    /* empm_bin_record_ints(n_grid, N) ints, or NULL */
} empm_grid_rec;

/* One grid post-process, in registration order.  kind:
 *   0 surface collider       add_surface_collider      mpm_solver_diff.py:437-531
 *   1 velocity on cuboid     set_velocity_on_cuboid    mpm_solver_diff.py:814-869
 *   2 bounding box           add_bounding_box          mpm_solver_diff.py:871-938
 *   3 geometry collider      add_geometry_collider     mpm_solver_diff.py:533-683
 *   4 grid geometry mover    set_grid_velocity_on_geometry mpm_solver_diff.py:685-804
 *   5 mask                   enforce_grid_velocity_by_mask mpm_solver_diff.py:1264-1289
 * geo_type: 1 cube, 2 sphere, 3 cylinder, 4 halfspace (kinds 3, 4).
 * Kinds 1 and 4 move: the rollout entry points advance `point` by
 * dt*velocity after each substep while start_time <= t < end_time
 * (mpm_solver_diff.py:795-801,860-866). */
typedef struct {
    int kind;
    int surface_type;
    int geo_type;
    int reset;
    int active;
    int pad0;
    float start_time;
    float end_time;
    float point[3];
    float normal[3];
    float size[3];
    float velocity[3];
    float rot_axis[3];
    float angular_velocity;
    float radius;
    float half_height;
    float friction;
    const int *mask; /* device, n^3 int32, kind 5 only */
} empm_collider;

/* A batch: n_cand independent simulations advanced by the same kernel
 * launches over concatenated arenas (the population of a CMA-ES step or a
 * gradient batch; no counterpart in the reference, which runs candidates one
 * after another, phys_sim/engine/trainer_warp_mpm.py:598-663).  Simulation c owns the next
 * models[c].n_particles particles of the arena (in order) and a grid of its
 * own models[c].n_grid^3 nodes; grids may differ in size and material.  The
 * per-particle arrays below span the whole arena; the per-particle pointers
 * inside models[] are ignored.  All simulations share one time, dt and
 * collider table (each applies it in its own grid). */
typedef struct {
    int n_cand;
    int n_particles; /* arena total: sum of models[c].n_particles */
    const empm_model *models; /* host array of n_cand */
    float *mu;
    float *lam;
    float *yield_stress;
    const float *mass;
    const float *vol;
    const int *selection; /* or NULL */
} empm_batch;

/* Capacity of a batch context: at most max_cand simulations, max_particles
 * particles in all, each grid at most max_grid^3 nodes. */
typedef struct {
    int device;
    int max_cand;
    int max_particles;
    int max_grid;
} empm_batch_create_desc;

/* ABI version (EMPM_ABI_VERSION). */
int empm_abi_version(void);

/* Message of the last failed call on this thread. */
const char *empm_last_error(void);

/* Allocate a context (grid buffers and per-particle scratch) on a device.
 * Replaces the per-rollout/per-substep allocations of MPMStateStruct.init /
 * init_grid / partial_clone (mpm_data_structure.py:41-117,356-387). */
int empm_create(empm_ctx **out, const empm_create_desc *desc);

/* Allocate a context for batches (empm_batch_rollout_*).  No counterpart
 * in the reference. */
int empm_create_batch(empm_ctx **out, const empm_batch_create_desc *desc);

/* Free a context (synchronises its device first). */
void empm_destroy(empm_ctx *ctx);

/* Replay each call of empm_rollout_fwd / _fwd_replay / _adj (and their
 * batch variants) as a HIP graph (enable != 0) instead of launching its
 * kernels one by one.  A call's launch sequence is captured once, cached in
 * the context (the 8 most recently used) under everything it bakes in (the
 * state, record and gradient pointers, substep counts, flags and launch
 * shapes) and replayed with one launch on a stream of the context's own,
 * ordered after the caller's stream's earlier work and before its later
 * work.  Times and collider tables are read from a device schedule uploaded
 * per call, and model parameters from the device simulation table, so
 * replays with new values need no new capture.  Results are those of the
 * eager launches.  Disabled while per-kernel profiling is on.  No
 * counterpart in the reference, whose host loop launches every Warp kernel
 * of every substep (mpm_solver_diff.py:165-291). */
int empm_set_graphs(empm_ctx *ctx, int enable);

/* Number of graph captures and replays of the context so far. */
int empm_graph_stats(empm_ctx *ctx, long long *captures, long long *replays);

/* Launch parameters of a context's binned kernels, picked per workload by
 * the autotuner (empm_amd.autotune) or set by hand.  Results do not depend
 * on them beyond the order of float atomic sums.  No counterpart in the
 * reference (Warp picks its own launch shapes, mpm_solver_diff.py:165-291). */
typedef struct {
    int scatter_wg;      /* threads per workgroup of the block-owned scatters (the waves sharing one block's
                          * LDS tile): 128, 256 or 512; default 256 */
    int scatter_per_cu;  /* scatter workgroups launched per CU, capped by occupancy (0: as many as fit) */
    int heavy_chunks;    /* blocks of more than this many 64-particle chunks are scattered first (0: every
                          * block counts as heavy, so the list keeps binning order); default 8 */
    int dynamic_rounds;  /* with more active blocks than this per scatter workgroup, blocks are handed out
                          * by a work counter instead of in fixed strides; default 4 */
    int node_grid_cap;   /* workgroups of a grid node pass (they stride over the near nodes); default 4096 */
} empm_tuning;

/* Set (validated; EMPM_E_INVALID_ARG names the bad field) or read the
 * context's tuning.  Setting it drops the context's cached graphs, whose
 * launch shapes it changes. */
int empm_set_tuning(empm_ctx *ctx, const empm_tuning *tuning);
int empm_get_tuning(empm_ctx *ctx, empm_tuning *tuning);

/* Deterministic mode (enable != 0): a call's results no longer depend on the
 * order in which float atomic sums arrive, so the same inputs give the same
 * bits run after run.  The binning then orders each cell's particles by
 * index (not by the order of its counting atomics), and the block-owned
 * scatters run 256-thread workgroups (whatever the tuning says) whose waves
 * each sum into an LDS tile of their own, added in wave order.  Slower (one
 * more binning pass, four tiles per scatter workgroup); for tests and
 * debugging.  Drops the context's cached graphs when the mode changes.  No
 * counterpart in the reference, whose P2G float atomic_add
 * (mpm_utils.py:420-423) sums in arrival order. */
int empm_set_deterministic(empm_ctx *ctx, int enable);

/* Failure flags of the context's simulations, set on the device by every
 * forward substep since the last read (a per-rollout check that costs one
 * read, not one synchronisation per substep): status[c] (host, n entries)
 * receives simulation c's EMPM_STATUS_* bits; clear != 0 resets them.
 * Synchronises `stream`.  Returns EMPM_E_NONFINITE when any simulation has
 * EMPM_STATUS_NONFINITE.  No counterpart in the reference, which has no
 * failure detection: a Neo-Hookean particle with J <= 0 turns NaN
 * (mpm_utils.py:38) and the rollout continues silently. */
int empm_read_status(empm_ctx *ctx, int *status, int n, int clear, void *stream);

/* mu, lam from E, nu: compute_mu_lam_from_E_nu (mpm_utils.py:277-283). */
int empm_mu_lam(int n, const float *E, const float *nu, float *mu, float *lam, void *stream);

/* Adjoint of empm_mu_lam: dE += dmu dmu/dE + dlam dlam/dE (same for nu). */
int empm_mu_lam_adj(int n, const float *E, const float *nu, const float *dmu, const float *dlam, float *dE, float *dnu,
                    void *stream);

/* mass = density * vol: get_float_array_product as launched at
 * interface.py:125-134 (mpm_data_structure.py:719-726). */
int empm_mass_from_density(int n, const float *density, const float *vol, float *mass, void *stream);

/* Nearest neighbour in B of every point of A: dist[i] = min_j |a_i - b_j|^2
 * and idx[i] = the smallest j attaining it.  This is the knn_points(K = 1)
 * step of pytorch3d.loss.chamfer_distance, which the training loss calls per
 * frame on the visible points (phys_sim/engine/trainer_warp_mpm.py:746-752).
 * a: n x 3, b: m x 3, float32 device arrays; dist: n floats, idx: n ints;
 * m >= 1 when n >= 1.  workspace (device, 4-byte aligned, may be NULL) of
 * empm_nn_sqdist_workspace(n, m) bytes lets the search split B over more
 * workgroups when A alone is too few threads to fill the GPU; without it
 * the search runs unsplit (same result, slower for small n). */
int empm_nn_sqdist(int n, const float *a, int m, const float *b, float *dist, int *idx, void *workspace,
                   size_t workspace_bytes, void *stream);

/* Workspace bytes empm_nn_sqdist uses for n points against m (0: none). */
size_t empm_nn_sqdist_workspace(int n, int m);

/* empm_nn_sqdist that allocates nothing and takes no workspace: the split
 * search's partial results live in a library-owned device array of 8 MiB
 * per GPU (2^20 (split, query) pairs; held until the process ends), which
 * holds every split plan the search makes, so the result equals
 * empm_nn_sqdist's with a workspace.  No counterpart in the reference: it is
 * for callers that cannot allocate (HIP graph capture of the training loss,
 * serving loops, harnesses that forbid hipMalloc).  Runs on the current
 * device; `stream` must belong to it.  Calls that overlap on one GPU share
 * that array and race: order them on one stream.  Returns EMPM_E_INVALID_ARG
 * for bad arguments or a split plan that would exceed the array. */
int empm_nn_sqdist_noalloc(int n, const float *a, int m, const float *b, float *dist, int *idx, void *stream);

/* empm_nn_sqdist over n_seg segments in one launch (the frames of the
 * training loss, trainer_warp_mpm.py:743-752): segment g pairs the rows
 * [a_off[g], a_off[g+1]) of a with the rows [b_off[g], b_off[g+1]) of b.
 * a_off, b_off: DEVICE arrays of n_seg + 1 ints; n_a = a_off[n_seg];
 * max_a, max_b: the longest segment of a and of b (host values, they size
 * the grid; a longer segment is searched only up to max_a rows).  idx is
 * global into b; a row whose b segment is empty gets +inf and -1.
 * workspace as for empm_nn_sqdist, of
 * empm_nn_sqdist_segments_workspace(n_seg, n_a, max_a, max_b) bytes. */
int empm_nn_sqdist_segments(int n_seg, const int *a_off, const int *b_off, int n_a, int max_a, int max_b,
                            const float *a, const float *b, float *dist, int *idx, void *workspace,
                            size_t workspace_bytes, void *stream);
size_t empm_nn_sqdist_segments_workspace(int n_seg, int n_a, int max_a, int max_b);

/* One substep state `in` -> state `out` (`out` must not alias `in`):
 * MPMWARPDiff.p2g2p_differentiable (mpm_solver_diff.py:165-291), i.e.
 * zero_grid, compute_stress_from_F_trial, p2g_apic_with_stress,
 * grid_normalization_and_gravity, add_damping_via_grid, the colliders `cols`
 * (host array, evaluated at `time`) and g2p_differentiable.  Collider points
 * are NOT advanced (the caller owns them, as the reference's modify_bc). */
int empm_substep_fwd(empm_ctx *ctx, const empm_model *model, const empm_collider *cols, int ncol, float time,
                     float dt, const empm_particles *in, const empm_particles *out, void *stream);

/* Adjoint of empm_substep_fwd at state `in`: given d_out = dL/d(out), writes
 * d_in = dL/d(in) (overwritten) and accumulates (+=) the parameter adjoints
 * d_mu, d_lam, d_ys, d_mass, d_vol (any may be NULL).  The substep's grid is
 * recomputed internally.  Replaces the Warp Tape replay of one substep's
 * launches (interface.py:197, warp_utils.py:93-127).  flags: EMPM_ADJ_*.
 * For material 5 and hardening == 1, whose return map rewrites (yield stress,
 * mu, lam), those are state: d_ys, d_mu, d_lam hold on entry the cotangents
 * of the updated values and on return those of the values the substep
 * started from (for other materials this is the same as accumulating); the
 * starting values are in->mstate, or the model arrays when it is NULL. */
int empm_substep_adj(empm_ctx *ctx, const empm_model *model, const empm_collider *cols, int ncol, float time,
                     float dt, const empm_particles *in, const empm_particles *d_out, const empm_particles *d_in,
                     float *d_mu, float *d_lam, float *d_ys, float *d_mass, float *d_vol, int flags, void *stream);

/* restart_and_compute_F_C (mpm_solver_diff.py:1214-1262) on the positions
 * x (N*3) and target positions target (N*3): zero_grid, set_F_C_p2g
 * (mpm_utils.py:843-876: mass and mass-weighted displacement target - x
 * into the grid), grid_normalization_and_gravity with dt = 0
 * (mpm_utils.py:427-438: no gravity, and no damping or colliders) and
 * set_F_C_g2p (mpm_utils.py:879-919): F_trial[p] = I + sum_i v_i (x) grad w_ip
 * (N*9, written for particles with selection == 0 only; the new C is not
 * stored, as in the reference).  Positions are not written: the reference
 * then sets x = target (set_vec3_to_vec3, mpm_solver_diff.py:1255-1261). */
int empm_restart_deformation(empm_ctx *ctx, const empm_model *model, const float *x, const float *target, float *F_trial,
                   void *stream);

/* n_steps substeps in one call: states[k] -> states[k+1] for k < n_steps
 * (states: host array of n_steps+1 descriptors; consecutive entries must not
 * alias, entries further apart may, e.g. a two-buffer ping-pong).  *time is
 * advanced by dt per substep and moving colliders (kinds 1, 4) in `cols`
 * (host, ncol entries) are advanced in place, as the reference host loop
 * does (warp_mpm_wrapper.py:278-306, interface.py:137-145).  If cols_hist /
 * times_hist are non-NULL (host, n_steps*ncol / n_steps entries) the
 * collider table and time each substep saw are recorded for the adjoint.
 * If grids is non-NULL (host, n_steps entries) substep k writes its grid
 * into grids[k]; with states[k].Fe non-NULL as well, the substep is fully
 * recorded and empm_rollout_adj differentiates it without recompute.
 * states[k].mstate, when non-NULL, receives the material state substep k
 * starts from (material 5 / hardening; see empm_particles). */
int empm_rollout_fwd(empm_ctx *ctx, const empm_model *model, empm_collider *cols, int ncol, double *time, double dt,
                     int n_steps, const empm_particles *states, const empm_grid_rec *grids, empm_collider *cols_hist,
                     double *times_hist, void *stream);

/* Forward replay of n_steps substeps with the recorded per-substep collider
 * tables and times of empm_rollout_fwd (no collider advance): the recompute
 * step of checkpointed reverse mode (the reference keeps every state alive
 * instead, interface.py:137-145).  grids: as for empm_rollout_fwd.  For
 * material 5 / hardening the replay starts from states[0].mstate (required)
 * and rewrites the model arrays as the forward does. */
int empm_rollout_fwd_replay(empm_ctx *ctx, const empm_model *model, const empm_collider *cols_hist, int ncol,
                            const double *times_hist, double dt, int n_steps, const empm_particles *states,
                            const empm_grid_rec *grids, void *stream);

/* Reverse sweep over n_steps substeps whose inputs are states[0..n_steps-1]
 * (host descriptors), with the per-substep collider tables and times from
 * empm_rollout_fwd.  On entry adj[0] holds dL/d(state n_steps); on return
 * adj[n_steps % 2] holds dL/d(state 0).  Parameter adjoints accumulate.
 * If grids is non-NULL (host, n_steps entries, as written by the forward)
 * and states[k].Fe is non-NULL, substep k uses the recorded grid and Fe;
 * otherwise its P2G and boundary pass are recomputed.  For material 5 /
 * hardening every states[k].mstate is required and d_ys, d_mu, d_lam carry
 * the cotangents of the material state (see empm_substep_adj): on return
 * they hold dL/d(values before substep 0). */
int empm_rollout_adj(empm_ctx *ctx, const empm_model *model, const empm_collider *cols_hist, int ncol,
                     const double *times_hist, double dt, int n_steps, const empm_particles *states,
                     const empm_grid_rec *grids, const empm_particles *adj, float *d_mu, float *d_lam, float *d_ys,
                     float *d_mass, float *d_vol, int flags, void *stream);

/* Order particles by the grid cell their B-spline stencil starts at:
 * perm[k] = index of the k-th particle in cell order (device, n_particles
 * ints).  Positions are those of any state.  The scatter kernels reduce
 * contributions of same-cell neighbours within a wavefront before their
 * global atomics, so running a rollout on particles in this order cuts the
 * atomic count by the particles-per-cell factor; results do not depend on
 * the order.  No counterpart in the reference (it scatters in caller order,
 * mpm_utils.py:358-423). */
int empm_sort_by_cell(empm_ctx *ctx, int n_particles, const float *x, float inv_dx, int *perm, void *stream);

/* empm_rollout_fwd / _replay / _adj over a batch: every substep runs each
 * stage once for all simulations (one launch per stage, per material for the
 * constitutive stages).  States, grid records and parameter adjoints span
 * the arenas (grid records: sum of n_grid^3 nodes; bins:
 * empm_batch_sizes).  Replaces the per-candidate rollout loop of the
 * reference's CMA-ES step (phys_sim/engine/trainer_warp_mpm.py:598-663). */
int empm_batch_rollout_fwd(empm_ctx *ctx, const empm_batch *batch, empm_collider *cols, int ncol, double *time,
                           double dt, int n_steps, const empm_particles *states, const empm_grid_rec *grids,
                           empm_collider *cols_hist, double *times_hist, void *stream);
int empm_batch_rollout_fwd_replay(empm_ctx *ctx, const empm_batch *batch, const empm_collider *cols_hist, int ncol,
                                  const double *times_hist, double dt, int n_steps, const empm_particles *states,
                                  const empm_grid_rec *grids, void *stream);
int empm_batch_rollout_adj(empm_ctx *ctx, const empm_batch *batch, const empm_collider *cols_hist, int ncol,
                           const double *times_hist, double dt, int n_steps, const empm_particles *states,
                           const empm_grid_rec *grids, const empm_particles *adj, float *d_mu, float *d_lam,
                           float *d_ys, float *d_mass, float *d_vol, int flags, void *stream);

/* empm_sort_by_cell for a batch: perm (device, n_particles ints) orders the
 * arena by (simulation, block, cell), so the particles of simulation c keep
 * the range its models[c].n_particles gives them. */
int empm_batch_sort_by_cell(empm_ctx *ctx, const empm_batch *batch, const float *x, int *perm, void *stream);

/* Sizes of a batch's arenas: total grid nodes and the ints of one binning
 * record. */
int empm_batch_sizes(const empm_batch *batch, long long *n_nodes, long long *bin_record_ints);

/* Size in ints of one substep's binning record (empm_grid_rec.bins) for a
 * grid of resolution n_grid and n_particles particles; -1 on bad arguments.
 * Every substep bins its input particles by stencil cell before P2G, so the
 * scatter never meets a particle outside its block's tile.  No counterpart
 * in the reference (it scatters unsorted particles with global atomics,
 * mpm_utils.py:358-423). */
long long empm_bin_record_ints(int n_grid, int n_particles);

/* Kernel ids for per-kernel timing */
enum {
    EMPM_KERNEL_P2G = 0,      /* P2G scatter (momentum, mass, stress impulse) */
    EMPM_KERNEL_GRID = 1,     /* grid normalise / gravity / colliders */
    EMPM_KERNEL_G2P = 2,      /* G2P gather + state update + next substep's stress */
    EMPM_KERNEL_G2P_ADJ = 3,  /* adjoint of G2P, per particle */
    EMPM_KERNEL_GRID_ADJ = 4, /* adjoint of the grid stage */
    EMPM_KERNEL_P2G_ADJ = 5,  /* adjoint of P2G + stress (gather) */
    EMPM_KERNEL_STRESS = 6,   /* stress of the first substep of a call */
    EMPM_KERNEL_G2P_ADJ_SCATTER = 7, /* adjoint of G2P: scatter of dL/dv_out */
    EMPM_KERNEL_BIN = 8,      /* per-substep binning (4 launches, timed as one) */
    EMPM_NUM_KERNELS = 9
};

/* Instrumentation (no counterpart in the reference, whose per-stage
 * ScopedTimer syncs the device five times per substep,
 * mpm_solver_diff.py:214-289): time the kernels whose bit is set in
 * kernel_mask with HIP events recorded on their launch stream.  0 disables. */
int empm_profile_enable(empm_ctx *ctx, unsigned kernel_mask);

/* Resolve pending events of one kernel (waits for them) and return the
 * cumulative launch count and milliseconds since the last reset. */
int empm_profile_read(empm_ctx *ctx, int kernel_id, long *launches, double *total_ms);

/* Zero the cumulative counters. */
int empm_profile_reset(empm_ctx *ctx);

/* Name of a kernel id ("" if out of range). */
const char *empm_kernel_name(int kernel_id);

/* ---- Linear-blend skinning of Gaussians on the simulated particles -------
 * (third_party/gaussian_splatting/dynamic_utils.py).  The particles are the
 * "bones".  All arrays are float32 device arrays, points as n x 3 rows;
 * index arrays are n x k rows of idx_bytes = 4 (int32) or 8 (int64, what
 * torch.topk returns) byte integers. */

/* The k nearest rows of r (m x 3) for every row of q (n x 3), nearest first:
 * idx (n x k ints) and the Euclidean distances dist (n x k); equal distances
 * keep the smaller index first.  skip_first = 1 finds k + 1 and drops the
 * first (the query itself when q is r).  1 <= k + skip_first <= 32 and
 * m >= k + skip_first.  Replaces the torch.topk over torch.cdist / a dense
 * distance matrix of get_topk_indices (dynamic_utils.py:248-265, with
 * skip_first = 1) and knn_weights / knn_weights_sparse (:268-302). */
int empm_knn(int n, const float *q, int m, const float *r, int k, int skip_first, int *idx, float *dist,
             void *stream);

/* weights[i, j] = 1 / (|xyz_i - bones_idx[i,j]| + 1e-6), normalised over j:
 * calc_weights_vals_from_indices (dynamic_utils.py:282-292).  k <= 32;
 * bones: n_bones x 3.  A row holding an index outside [0, n_bones) is NaN. */
int empm_lbs_weights(int n, const float *xyz, int k, const void *idx, int idx_bytes, int n_bones,
                     const float *bones, float *weights, void *stream);

/* Per-bone transforms of interpolate_motions_speedup (dynamic_utils.py:
 * 321-416): for bone b with adjacent bones relations[b, :] (n_bones x n_adj),
 * F = sum_j ((bones + motions)[a_j] - (bones + motions)[b]) (bones[a_j] -
 * bones[b])^T, R = the rotation closest to F, q = its unit quaternion
 * (w, x, y, z).  bone_rec: 16 floats per bone, 64-byte aligned:
 * [q.w q.x q.y q.z  b.x b.y b.z 0  m.x m.y m.z 0  0 0 0 0];
 * a bone with a relation index outside [0, n_bones) gets a NaN rotation.
 * Ranks (torch.linalg.matrix_rank's tolerance) decide which rotations the
 * reference keeps (dynamic_utils.py:331-405): one per bone of rank 2 or 3,
 * or, when any bone has rank 1, one per rank-1 bone (e_x turned to the
 * direction of F's range about their cross product; that direction's sign,
 * torch.svd's in the reference, is taken with its largest-magnitude
 * component positive); the kept set is applied to all bones when it holds
 * every bone or exactly one (a broadcast), else every rotation is the
 * identity.  rank_state (device, 8 ints) receives what empm_lbs_points needs
 * to apply this: [0] the number of bones of rank < 2, [1] of rank 1, [2] of
 * rank >= 2, [3] and [4] the largest rank-1 and rank >= 2 bone index (-1 if
 * none), [5] the rank-1 bones whose second singular value exceeds 1e-8
 * (the reference's allclose check fails and every rotation is the
 * identity), [6..7] 0. */
int empm_lbs_bones(int n_bones, const float *bones, const float *motions, int n_adj, const void *relations,
                   int idx_bytes, float *bone_rec, int *rank_state, void *stream);

/* Per-Gaussian blend of interpolate_motions_speedup (dynamic_utils.py:
 * 434-485): xyz_out = sum_j w_j (R_j (xyz - b_j) + m_j + b_j) and, when quat
 * (n x 4, w x y z, 16-byte aligned) is given, quat_out = normalize(sum_j w_j
 * q_j) * quat (Hamilton product).  weights (n x k) may be null: the weights
 * of empm_lbs_weights are then computed from the bones in bone_rec and xyz.
 * bone_rec holds n_bones records and rank_state (8 device ints) the ranks
 * empm_lbs_bones found: R_j is bone j's own rotation, one bone's rotation
 * broadcast to all, or the identity, as the reference assigns them; a row
 * holding an index outside [0, n_bones) is NaN.  Outputs must not alias
 * inputs. */
int empm_lbs_points(int n, const float *xyz, const float *quat, int k, const void *idx, int idx_bytes,
                    const float *weights, int n_bones, const float *bone_rec, const int *rank_state,
                    float *xyz_out, float *quat_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EMPM_H */
