/*
 * pmpm.h -- C ABI of the MI355X-native AnisoMPM substep (libpmpm.so).
 *
 * The reference has no FFI: its step is the Python method
 * AnisoMPM.substep (anisompm_py/src/anisompm.py:593-803).  Each entry point
 * below replaces one piece of that method (or of its helpers) and is what a
 * Python ctypes binding calls; the binding itself is peel_mpm_amd/_lib.py.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers borrowed from the caller
 *     (PyTorch tensors); the library never allocates device memory.  The
 *     caller also provides the workspace (pmpm_workspace_bytes) and calls
 *     pmpm_workspace_init once before the first step on it.
 *   - Particle arrays keep the reference's row-major per-particle layout:
 *     vectors (N,3), matrices (N,3,3), scalars (N,), flags (N,) uint8.
 *   - dtype is PMPM_F32 or PMPM_F64 and applies to every floating array.
 *   - stream is a hipStream_t passed as void* (NULL = default stream); all
 *     work is enqueued on it and the call returns without synchronising.
 *   - Every entry point returns 0 on success and a negative PMPM_E* code on
 *     failure; pmpm_last_error() then describes it.  Nothing throws.
 *   - Not thread-safe per workspace; two steps on one workspace must be
 *     ordered on one stream.
 */
#ifndef PMPM_H
#define PMPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PMPM_F32 = 0, PMPM_F64 = 1 };
enum { PMPM_OK = 0, PMPM_EARG = -1, PMPM_EHIP = -2, PMPM_ESIZE = -3 };

enum { PMPM_MAX_COLLIDERS = 8, PMPM_MAX_VEL_BC = 8 };
enum { PMPM_COLLIDER_SPHERE = 1, PMPM_COLLIDER_HALFSPACE = 2 };
enum { PMPM_STICKY = 0, PMPM_SLIP = 1 };

/* One grid collider, evaluated at the step's time (anisompm.py:828-867).
 * p0 = centre/origin at time t, p1 = at t + 1e-4 (the collider velocity is
 * the forward difference (p1 - p0) / 1e-4, formed on the device in dtype). */
typedef struct {
    int32_t kind;        /* PMPM_COLLIDER_SPHERE or PMPM_COLLIDER_HALFSPACE */
    int32_t mode;        /* PMPM_STICKY or PMPM_SLIP */
    double friction;
    double radius;       /* sphere only */
    double p0[3];
    double p1[3];
    double normal[3];    /* half-space only: unit outward normal */
} pmpm_collider;

/* Step configuration (AnisoMPM.__init__ kwargs, anisompm.py:304-327, plus
 * the host-side part of the damage gate, anisompm.py:647). */
typedef struct {
    int32_t dtype;
    int32_t n_grid;
    double grid_lim, dx, inv_dx, dt;
    double gravity[3];
    double grid_damp;
    int32_t use_fclamp;          /* f_clamp is not None */
    double fclamp_lo, fclamp_hi;
    int32_t do_damage;           /* allow.any() and step % damage_every == 0 */
    double dt_damage;            /* dt * damage_every */
    int32_t n_colliders;
    pmpm_collider colliders[PMPM_MAX_COLLIDERS];
    int32_t n_vel_bc;            /* velocity-BC table used with bc_id (1-based) */
    double vel_bc[PMPM_MAX_VEL_BC][3];
    int32_t stable_sort;         /* 1: within-bucket order = ascending particle index */
    int32_t export_active;       /* 1: write the dense active-node mask (slow colliders) */
    int32_t exact_sigma_eig;     /* 1: Cauchy-stress eigenpairs by a second Jacobi, as the
                                    reference (anisompm.py:653); 0: closed form from the
                                    polar stretch, which shares its eigenvectors */
    /* Batched form: n_scenes independent scenes, each on its own n_grid^3 grid,
     * their particles concatenated; scene s owns particles
     * [scene_offsets[s], scene_offsets[s+1]) (device int64, n_scenes+1 entries).
     * scene_damage (device uint8 per scene, may be NULL) gates the damage update
     * per scene when do_damage is set.  n_scenes <= 1: one scene, offsets unused. */
    int32_t n_scenes;
    const int64_t *scene_offsets;
    const uint8_t *scene_damage;
    /* Damage update.  PMPM_DAMAGE_EXPLICIT: the reference's explicit update
     * (anisompm.py:679-685).  PMPM_DAMAGE_PCG: an extension with no
     * counterpart in the reference -- one backward-Euler step of the same
     * phase-field equation on the cubic grid, particles as quadrature points,
     * lumped mass:  (diag(m) + K) d_nodes = b  with
     *   m_i = sum_p N_ip (1 + c_p (1 + D_p)),  c_p = dt_damage / eta_p,
     *   K_ij = sum_p c_p l0_p^2 grad N_ip . grad N_jp,  b_i = sum_p N_ip (d_p + c_p D_p),
     * solved matrix-free by Jacobi-preconditioned CG (D = diag(A)) until the
     * preconditioned relative residual sqrt(r^T D^-1 r / b^T D^-1 b) is at most
     * pcg_tol (at most pcg_max_iter iterations), then interpolated to the
     * particles, clamped to [0,1] and made irreversible as the explicit update
     * is.  One scene only.  Outside stream capture the call reads the
     * solver's convergence flag between chunks of iterations (so it
     * synchronises the stream a few times); under capture it enqueues all
     * pcg_max_iter iterations, those after convergence exiting at once. */
    int32_t damage_solver;
    double pcg_tol;
    int32_t pcg_max_iter;
    /* Adjoint without recompute.  pmpm_substep with keep_for_adjoint = 1 also
     * keeps the grid pass's per-node mass sums in ws; pmpm_substep_adjoint
     * with keep_for_adjoint = 1 then takes binning, P2G records and grid from
     * ws as that forward call left them (same inputs, same ws, nothing run on
     * ws in between) instead of recomputing them. */
    int32_t keep_for_adjoint;
    /* Optional DEVICE pointer to a pmpm_boundary_table.  When set, the
     * kernels read the colliders and velocity-BC values from it when they
     * run, instead of from colliders[] and vel_bc[] above (n_colliders and
     * n_vel_bc still count them).  A step captured in a HIP graph then
     * follows moving colliders and time-dependent velocities: the caller
     * writes the table before each replay.  No counterpart in the reference
     * (its closures run on the host every step, anisompm.py:601, :744). */
    const struct pmpm_boundary_table *boundary_table;
    /* Binning carried to the next step.  No counterpart in the reference
     * (its step does not sort particles).  bin_next = 1: G2P also takes the
     * next step's cell keys and per-cell counts from the positions it writes
     * (out->x), into ws (not with keep_for_adjoint: PMPM_EARG).
     * prebinned = 1: the caller vouches that in->x holds, unchanged, the
     * positions the previous step on ws wrote with bin_next (no edit in
     * between); the step then skips its count pass.  The library takes that
     * up only when the previous call on ws set bin_next with out->x == in->x
     * and the same particle count, in the same stream capture (or both
     * outside one), and counts afresh otherwise; an edit of x in place it
     * cannot see, so a false claim after one gives a wrong step.  The
     * binning is the same either way (with stable_sort, bit for bit). */
    int32_t bin_next;
    int32_t prebinned;
} pmpm_step_params;

/* The device-resident form of the step's boundary values (see
 * pmpm_step_params.boundary_table). */
typedef struct pmpm_boundary_table {
    pmpm_collider colliders[PMPM_MAX_COLLIDERS];
    double vel_bc[PMPM_MAX_VEL_BC][3];
} pmpm_boundary_table;

enum { PMPM_DAMAGE_EXPLICIT = 0, PMPM_DAMAGE_PCG = 1 };

/* Particle state and materials at the start of the step (AnisoMPM attributes,
 * anisompm.py:374-463). n_interface is read only where interface_mask != 0.
 * bc_id may be NULL; else a nonzero bc_id[p] = k replaces v[p] by vel_bc[k-1]
 * for the transfer, as the reference's particle_bc hooks do (anisompm.py:601). */
typedef struct {
    int64_t n;
    const void *x, *v, *F, *C, *d, *lap, *D_hist;
    const void *vol0, *mass, *mu, *lam, *sigma_c, *eta, *zeta, *l0, *residual, *A0;
    const uint8_t *allow, *interface_mask;
    const void *n_interface;
    const uint8_t *bc_id;
    /* Uniform material fields.  Bit k of uniform_mask set: every particle holds
     * the same value of field k, given in uniform_value[k] (A0: the 3x3
     * matrix in uniform_A0, row-major), and the step reads that value instead
     * of the array, whose pointer may then be NULL.  Bits, in order: vol0,
     * mass, mu, lam, sigma_c, eta, zeta, l0, residual, A0 (PMPM_U_*).  The
     * values must be the fields' own (converted from the step's dtype), so
     * the step computes the same bits either way.  No counterpart in the
     * reference (its fields are per-particle tensors, though add_object fills
     * each from one scalar); pmpm_substep_adjoint takes no uniform fields. */
    uint32_t uniform_mask;
    double uniform_value[9];
    double uniform_A0[9];
} pmpm_particles_in;

enum {
    PMPM_U_VOL0 = 0, PMPM_U_MASS = 1, PMPM_U_MU = 2, PMPM_U_LAM = 3, PMPM_U_SIGMA_C = 4,
    PMPM_U_ETA = 5, PMPM_U_ZETA = 6, PMPM_U_L0 = 7, PMPM_U_RESIDUAL = 8, PMPM_U_A0 = 9
};

/* Outputs.  x, v, F, C, lap are always written, and so are tau_last and
 * J_last unless both are NULL.  d, D_hist, phi, phi_shear are written only
 * when do_damage (the reference then rebinds them, anisompm.py:665-685);
 * otherwise they may be NULL, and phi and phi_shear may be NULL together
 * also when do_damage.  The four diagnostics (tau_last, J_last, phi,
 * phi_shear; anisompm.py:665, :673, :705-706) feed nothing else in the step:
 * a caller that reads them only after the last substep of a loop passes
 * NULL on the others and saves their 48 bytes per particle of stores.
 * Outputs must not alias inputs. */
typedef struct {
    void *x, *v, *F, *C, *lap, *d, *D_hist, *phi, *phi_shear, *tau_last, *J_last;
} pmpm_particles_out;

/* ---- workspace ---------------------------------------------------------- */
/* Bytes of workspace for up to max_n particles over n_scenes grids of n_grid^3. */
size_t pmpm_workspace_bytes(int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype);
/* Zero the workspace and record, by its address, the shape it was set up
 * for.  Every call that takes ws afterwards (steps, adjoint, pmpm_bin, the
 * stats readers) needs the same n_grid, n_scenes and dtype and at most max_n
 * particles; any n <= max_n is served (the sections that depend on the
 * particle count are laid out last), anything else returns PMPM_EARG before
 * work is enqueued, as does a ws this call never set up.  Call it again for
 * a buffer reused with another shape. */
int pmpm_workspace_init(void *ws, size_t ws_bytes, int64_t max_n, int32_t n_grid, int32_t n_scenes,
                        int32_t dtype, void *stream);

/* ---- the step ------------------------------------------------------------ */
/* One substep: AnisoMPM.substep (anisompm.py:593-803) minus the Python
 * particle_bc hooks (replaced by bc_id/vel_bc) and with device colliders. */
int pmpm_substep(const pmpm_step_params *p, const pmpm_particles_in *in,
                 const pmpm_particles_out *out, void *ws, size_t ws_bytes, void *stream);
/* Steps on ws that started from a carried count (prebinned taken up)
 * since ws was set up.  No counterpart in the reference. */
int pmpm_carried_steps(const void *ws, int64_t *steps);
/* The same step split at the collider call (anisompm.py:736-746) so a host
 * may apply arbitrary collider callables to the grid velocity in between:
 * begin = bin + stress/damage + P2G + grid update (+ device colliders);
 * end   = G2P + cubic damage Laplacian + advection/F update/clamp. */
int pmpm_substep_begin(const pmpm_step_params *p, const pmpm_particles_in *in,
                       const pmpm_particles_out *out, void *ws, size_t ws_bytes, void *stream);
int pmpm_substep_end(const pmpm_step_params *p, const pmpm_particles_in *in,
                     const pmpm_particles_out *out, void *ws, size_t ws_bytes, void *stream);
/* No counterpart in the reference.  pmpm_substep without a workspace
 * argument and without allocating, for callers that cannot call hipMalloc
 * or an allocating library routine: code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  The workspace is
 * PMPM_NOALLOC_SCRATCH_BYTES (64 MiB) of device memory the library itself
 * holds, a static device array that every GPU the library runs on keeps
 * until the process ends -- enough for a 64^3-grid scene of up to about
 * 217,000 fp32 particles (pmpm_workspace_bytes(n, n_grid, n_scenes, dtype)
 * is the need; a larger step returns PMPM_ESIZE before anything is
 * enqueued).  Every call on a GPU shares that GPU's scratch and leaves its
 * contents behind; the call clears what must start at zero itself.  Calls
 * that overlap on one GPU (two streams or two threads) race.  The PCG
 * damage mode enqueues all pcg_max_iter iterations (no host polling);
 * keep_for_adjoint is refused (PMPM_EARG): the adjoint takes a caller-owned
 * workspace. */
#define PMPM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int pmpm_substep_noalloc(const pmpm_step_params *p, const pmpm_particles_in *in,
                         const pmpm_particles_out *out, void *stream);
/* Byte offsets into ws of the grid velocity (n_grid^3 x 4: vx,vy,vz,gd) and
 * of the dense active-node mask (n_grid^3 uint8, valid after begin with
 * export_active=1), for a host that resolves colliders itself. */
int pmpm_grid_views(int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype,
                    size_t *gv_offset, size_t *active_offset);
/* Active quadratic nodes (grid_m > 1e-12) and cubic nodes (sum w > 1e-12) of
 * the last step on ws; synchronises the stream. */
int pmpm_last_step_stats(const void *ws, int64_t max_n, int32_t n_grid, int32_t n_scenes,
                         int32_t dtype, int64_t *n_active_q, int64_t *n_active_c, void *stream);

/* Iterations and final relative residual sqrt(r^T D^-1 r / b^T D^-1 b) of the most recent PCG
 * damage solve on ws (zeros before the first one); synchronises the stream.
 * No counterpart in the reference (its damage update is explicit). */
int pmpm_damage_solve_info(const void *ws, int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype,
                           int32_t *iterations, double *rel_residual, void *stream);

/* ---- launch shape (autotuning) ------------------------------------------ */
/* Launch shape of the step's kernels, process-wide; zero fields take the
 * defaults.  No counterpart in the reference (its step is a sequence of
 * torch ops): these are the knobs the autotuner (peel_mpm_amd/autotune.py)
 * searches.  A step captured in a HIP graph keeps the shape it was captured
 * with; set the shape between steps, not while a step is being enqueued on
 * another thread. */
typedef struct {
    int32_t bucket_wgs;     /* work-groups of the persistent per-bucket kernels (P2G, G2P,
                               PCG drive/apply): a multiple of 8 in [8, 16384]; default 2048 */
    int32_t block_wgs;      /* work-groups of the per-node-block kernels (grid pass, PCG
                               fold/update) = the number of PCG partial dot products:
                               [1, 2048]; default 2048 */
    int32_t p2g_pair;       /* particles a lane scatters per arena update in P2G: 1 or 2;
                               default 2 */
    int32_t const_threads;  /* threads per work-group of the per-particle constitutive
                               kernel: 64, 128 or 256; default 256 */
    int32_t p2g_waves;      /* wavefronts (each with its own LDS arena) per bucket in P2G:
                               1, 2 or 4 (4: fp32 only, fp64 takes 2); default 2 */
    int32_t g2p_threads;    /* threads per work-group (one bucket) of G2P: 128, 256 or 512;
                               default 256 */
    int32_t cell_block;     /* cells along z of a P2G block: 4 (one 4x4x4-cell bucket per
                               work-group, a 7x7x7-node arena each) or 8 (the two buckets
                               adjacent along z in one wavefront, their cells packed onto the
                               lanes by particle count, into one 7x7x11-node arena, which the
                               grid pass folds); default 4.  The PCG operator keeps 4. */
    int32_t p2g_split;      /* P2G work split with cell_block 4: 1 (each cell on one lane,
                               p2g_waves wavefronts per bucket sharing its particles), 2
                               (two wavefronts per bucket sharing its channels, quadratic and
                               cubic, each with two arenas: a cell of two or more passes split
                               over them and the halves packed onto the lanes by load;
                               p2g_waves is then not used) or 3 (as 2, with the channels
                               balanced between the wavefronts: mv_x, mv_y on one and m, mv_z
                               and the cubic pair on the other); default 1 */
    int32_t const_stream;   /* where the explicit-damage step runs its constitutive kernel:
                               1 (on the launch stream after the binning; its P2G records at
                               the particles' sorted positions) or 2 (on a second stream of
                               the library's, one per device, beside the binning's scan and
                               sort; the records in particle order, P2G reading them through
                               the sort permutation).  The PCG damage mode, steps that keep
                               their records for the adjoint, pmpm_substep_noalloc and a
                               batched step that counts its cells afresh with a per-scene
                               damage gate take 1.  Steps enqueued from several host threads
                               on one device share that device's second stream.  Default 1 */
} pmpm_launch_config;
/* NULL restores the defaults.  Out-of-range values: PMPM_EARG, nothing changed. */
int pmpm_set_launch_config(const pmpm_launch_config *cfg);
int pmpm_get_launch_config(pmpm_launch_config *cfg);

/* ---- adjoint ------------------------------------------------------------- */
/* Adjoints of a step's outputs (read) ... */
typedef struct {
    const void *x, *v, *F, *C, *d, *lap, *D_hist;
} pmpm_adjoint_in;
/* ... and of its inputs (written), plus parameter adjoints (accumulated +=). */
typedef struct {
    void *x, *v, *F, *C, *d, *lap, *D_hist;
    void *mu, *lam, *A0, *sigma_c;
} pmpm_adjoint_out;

/* Bytes of the adjoint's workspace: 6 values per grid node, 10 per particle. */
size_t pmpm_adjoint_workspace_bytes(int64_t max_n, int32_t n_grid, int32_t dtype);
/* Backward of one substep (vector-Jacobian product of anisompm.py:593-803).
 * No counterpart in the reference, which has no gradient code.  in = the
 * step's inputs as passed to pmpm_substep; fwd_out = its outputs: v and C
 * are read, d, D_hist, phi, phi_shear, tau_last, J_last are rewritten by the
 * recompute (with the stable sort, to the same values) unless
 * p->keep_for_adjoint reuses the forward call's workspace.  p must have
 * do_damage as in the forward call.  Supported: one scene, the explicit
 * damage update, f_clamp (through the strain limit's spectral derivative),
 * device colliders (sticky, or slip with friction; torch's subgradient
 * conventions at the clamp kinks); not interface particles (the directional
 * split is not differentiated).  Velocity-BC particles (bc_id) get
 * v_bar = 0.  With keep_for_adjoint the forward call keeps per node m, sum w
 * and mv/m in ws for it. */
int pmpm_substep_adjoint(const pmpm_step_params *p, const pmpm_particles_in *in,
                         const pmpm_particles_out *fwd_out, const pmpm_adjoint_in *gin,
                         const pmpm_adjoint_out *gout, void *ws, size_t ws_bytes, void *aws,
                         size_t aws_bytes, void *stream);

/* ---- pieces exposed for function-level parity --------------------------- */
/* Binning: counting sort of particles by cell, cells numbered bucket-major:
 * key = bucket * 64 + (cx%4)*16 + (cy%4)*4 + (cz%4), where (cx,cy,cz) =
 * floor(x * inv_dx), the reference's cell (anisompm.py:574), clamped to the
 * grid, and bucket = the 4x4x4-cell block (cx/4, cy/4, cz/4) numbered
 * (bx*nb + by)*nb + bz with nb = ceil(n_grid/4).  perm (N) lists particles
 * cell by cell; with stable != 0 it equals a stable argsort of key.  key (N)
 * and cell_start (64*nb^3 + 1 offsets into perm) may be NULL. */
int pmpm_bin(const void *x, int32_t dtype, int64_t n, int32_t n_grid, double inv_dx, int32_t stable,
             int32_t *key, int32_t *perm, int32_t *cell_start, void *ws, size_t ws_bytes,
             void *stream);
/* Stencil of each particle as the step computes it: quadratic base cell and
 * 27 node indices, cubic base cell and 64 node indices (clamp-to-edge), and
 * the weights (anisompm.py:544-590).  Any output may be NULL.
 * base_q, base_c: (N,3) int32; idx27 (N,27), idx64 (N,64) int32;
 * w (N,3,3), wc (N,4,3), wc2 (N,4,3) in dtype. */
int pmpm_stencil(const void *x, int32_t dtype, int64_t n, int32_t n_grid, double inv_dx,
                 int32_t *base_q, int32_t *idx27, int32_t *base_c, int32_t *idx64, void *w,
                 void *wc, void *wc2, void *stream);
/* Register-resident batched cyclic Jacobi on symmetric 3x3 (anisompm.py:48-95):
 * A (N,3,3) -> evals (N,3) ascending, V (N,3,3) eigenvectors as columns. */
int pmpm_sym_eig(const void *A, int32_t dtype, int64_t n, void *evals, void *V, void *stream);

/* Live kernel timing with HIP events on the launch stream (for benchmarks):
 * pmpm_timing(1) clears and starts recording around each kernel group
 * (0 binning, 1 stress+P2G, 2 grid update, 3 G2P, 4 the PCG damage solve,
 * 5 one CG iteration within it, 6 the constitutive kernel when it runs on
 * the second stream, launch shape const_stream 2: group 1 then holds P2G
 * alone); pmpm_timing_read synchronises and returns each group's summed
 * milliseconds and launch count (n <= 7 groups).  Adds two event records
 * per group per step. */
int pmpm_timing(int32_t enable);
int pmpm_timing_read(double *total_ms, int64_t *launches, int32_t n);

const char *pmpm_last_error(void);
const char *pmpm_version(void);
/* sizeof of pmpm_collider, pmpm_step_params, pmpm_particles_in,
 * pmpm_particles_out, pmpm_frame_in, pmpm_frame_obs, pmpm_pcg64,
 * pmpm_boundary_table (in that order) for binding-layout checks; returns the number of sizes written (at
 * most n). */
int pmpm_abi_sizes(size_t *sizes, int32_t n);

/* ---- after the step: the renderer's inputs -------------------------------- */
/* Gaussian carry, replacing the torch ops of deform_and_render
 * (anisompm_py/src/run_orange.py:116-128): splat g rides on particle
 * p = bind[g] (int64),
 *   pos[g]   = (x[p] / grid_lim - 0.5) / scale_to_sim + center + F[p] off0[g]
 *   cov[g]   = F[p] cov0[g] F[p]^T
 *   color_out[g] = color[g] (1 - crack_tint d[p])   (color copied when crack_tint <= 0)
 * x (N,3), F (N,3,3), d (N,) are the step's particle state; off0 (G,3),
 * cov0 (G,3,3), color (G,3) the splats' reference offsets, covariances and
 * colours; pos (G,3), cov (G,3,3) are written, color_out (G,3) when not NULL
 * (d may be NULL when crack_tint <= 0).  bind must index [0, n_particles).
 * One kernel, enqueued on stream. */
int pmpm_gaussian_carry(int32_t dtype, int64_t n_gauss, int64_t n_particles, const int64_t *bind,
                        const void *x, const void *F, const void *d, const void *off0,
                        const void *cov0, const void *color, double grid_lim, double scale_to_sim,
                        const double *center, double crack_tint, void *pos, void *cov,
                        void *color_out, void *stream);

/* ---- after the step: the peel test's per-frame observables --------------- */
/* The reductions of the reference's peel frame loop
 * (anisompm_py/src/peel_test.py:271-310) on the device, with no host
 * synchronisation per frame.  Per call (one frame), from the step's state:
 *   phi_pp = max(phi_pp, phi), psh_pp = max(psh_pp, phi_shear)    (per particle)
 *   peaks[0..2] = max(peaks[0..2], max phi over interface, flesh, peel)
 *   peaks[3] = max(peaks[3], max phi_shear over the interface)
 *   peaks[4] = max(peaks[4], phi_int / (max(phi_flesh, phi_peel) + 1e-9))   (routing)
 *   onset: the first frame with mean(d[interface] > 0.5) > 0.05; on it the
 *          snapshot buffers receive d, phi_pp, psh_pp, x, tau_last, J_last
 *   log[10 * frame ..]: mean and max of d over the interface, the broken
 *          fraction, max d over flesh and peel, the three region phi maxima,
 *          mean y of the grip (cap) particles, any NaN in x (1.0 or 0.0)
 * Maxima propagate NaN as torch's do.  phi and phi_shear may be NULL (a step
 * without the damage update): then the running maxima and peaks stay and the
 * log row holds NaN for the phi maxima.  Every region must hold particles,
 * as the reference requires.  No counterpart as a function in the reference
 * (its frame loop does these reductions with torch operations and a host
 * read per frame). */
typedef struct {
    const void *phi, *phi_shear;   /* (N,) or NULL */
    const void *d, *x;             /* (N,), (N,3) */
    const void *tau_last, *J_last; /* (N,3,3), (N,); NULL: not snapshotted */
    const int8_t *region;          /* (N,): 0 flesh, 1 interface, 2 peel */
    const uint8_t *cap;            /* (N,): grip particles */
} pmpm_frame_in;
typedef struct {
    void *phi_pp, *psh_pp;         /* (N,): running per-particle maxima (updated) */
    void *d_e, *phipp_e, *pshpp_e, *x_e, *tau_e, *J_e;  /* onset snapshot (tau_e, J_e may be NULL) */
    double *peaks;                 /* device, 5 values (start at 0) */
    int32_t *onset;                /* device, 2 values: onset frame (start at -1), snapshot-this-frame */
    double *log;                   /* device, 10 values per frame */
} pmpm_frame_obs;
/* Bytes of the per-call scratch (partial reductions) pmpm_observe_frame needs. */
size_t pmpm_observe_workspace_bytes(void);
int pmpm_observe_frame(int32_t dtype, int64_t n, int32_t frame, const pmpm_frame_in *in,
                       const pmpm_frame_obs *obs, void *ws, size_t ws_bytes, void *stream);

/* ---- Option A directional damage (the variational pair) --------------------- */
/* energy() and directional_pk1() (anisompm_py/src/anisompm.py:129-187): per
 * particle the energy density Psi (psi, (N,)) and its first Piola-Kirchhoff
 * stress P = dPsi/dF (P, (N,3,3)) for F (N,3,3), reference interface normals
 * n0 (N,3) and damage d (N,); mu, lam, gamma_I, gamma_II per particle from
 * the *_arr arrays (N,) or, when an array is NULL, the scalar.  Either output
 * may be NULL.  One kernel, enqueued on stream. */
int pmpm_directional_pk1(int32_t dtype, int64_t n, const void *F, const void *n0, const void *d, const void *mu_arr,
                         double mu, const void *lam_arr, double lam, const void *gamma_I_arr, double gamma_I,
                         const void *gamma_II_arr, double gamma_II, void *P, void *psi, void *stream);

/* directional_kirchhoff (anisompm_py/src/anisompm.py:190-227) as the step
 * uses it for interface particles (:693-703), through the step's own code:
 * tau = Q tau_c Q + gI P tau_c P + gII (P tau_c Q + Q tau_c P), P = m m^T,
 * Q = I - P, m = F n0 / (|F n0| + 1e-9), gI = gII = (1 - d)^2 (1 - k_r) + k_r
 * (k_r per particle from res_arr (N,), or res when it is NULL); with
 * tension_gate, gI = 1 where m^T tau_c m <= 0.  tau_c, F, tau (N,3,3); d
 * (N,); n0 (N,3).  One kernel, enqueued on stream. */
int pmpm_directional_kirchhoff(int32_t dtype, int64_t n, const void *tau_c, const void *F, const void *d, const void *n0,
                               const void *res_arr, double res, int32_t tension_gate, void *tau, void *stream);

/* ---- scene builders (setup) ------------------------------------------------ */
/* The reference's particle-set constructors, bit-exact with their NumPy /
 * SciPy results.  Jitter draws come from NumPy's default_rng bit generator
 * (PCG64); the caller seeds it (NumPy's SeedSequence) and passes the seeded
 * state: the 128-bit LCG state and increment. */
typedef struct {
    uint64_t state_hi, state_lo, inc_hi, inc_lo;
} pmpm_pcg64;

/* Bytes of scratch the builders' stable compaction needs for n candidates. */
size_t pmpm_compact_workspace_bytes(int64_t n_candidates);

/* ball_particles (anisompm_py/src/peel_test.py:128-137): the points of the
 * lattice g x g x g (g: (n,) fp64, device; candidates in meshgrid "ij" C
 * order) with sqrt(x^2 + y^2 + z^2) < R, each moved by (u - 0.5) js (u: the
 * generator's draws, three per kept point in order) and by center (host,
 * 3), rounded to fp32 into out ((cap, 3), device).  *count (host) gets the
 * number of points (also when it exceeds cap, which fails with
 * PMPM_ESIZE).  n^3 must stay below 2^31 (n <= 1290).  Synchronises stream
 * once (the count); the points are written on stream. */
int pmpm_ball_particles(const double *g, int64_t n, double R, double js, const double *center, const pmpm_pcg64 *rng,
                        float *out, int64_t cap, int64_t *count, void *ws, size_t ws_bytes, void *stream);

/* grid_block (anisompm_py/src/test_block.py:41-47): out ((nx ny nz, 3) fp64,
 * device) = the "ij" meshgrid of xs (nx), ys (ny), zs (nz) (device, fp64 --
 * the per-axis linspaces), raveled in C order. */
int pmpm_grid_block(const double *xs, const double *ys, const double *zs, int64_t nx, int64_t ny, int64_t nz,
                    double *out, void *stream);

/* fill_interior (anisompm_py/src/orange_setup.py:65-92), in two calls.
 * Points w ((n,3), device, w_dtype PMPM_F32 or PMPM_F64); arithmetic in
 * ct_dtype, the type NumPy's promotion gives the caller's points, centre and
 * scale (fp32 only when all three are fp32).
 * pmpm_fill_bounds: sim = (w - center) * scale + 0.5; writes lo (host, 3) =
 * the minimum of sim per axis and dims (host, 3) = max floor((sim - lo) / dx)
 * + 3 per axis.  Workspace: pmpm_fill_workspace_bytes(NULL).
 * pmpm_fill_interior: marks the points' voxels in a dims grid, closes it
 * (dilation, then erosion, by the 6-neighbour cross, 0 outside the array),
 * fills its holes (background not connected through background to the
 * array's faces), and writes one point per solid voxel in C order, lo +
 * (ijk + 0.5) dx, moved by (u - 0.5) js when rng is not NULL, rounded to fp32
 * into out ((cap, 3), device).  *count as for pmpm_ball_particles.  The
 * voxel grid must hold fewer than 2^31 voxels.
 * Workspace: pmpm_fill_workspace_bytes(dims).  Both synchronise stream. */
size_t pmpm_fill_workspace_bytes(const int64_t *dims);
int pmpm_fill_bounds(int32_t w_dtype, int32_t ct_dtype, const void *w, int64_t n, const double *center, double scale,
                     double dx, double *lo, int64_t *dims, void *ws, size_t ws_bytes, void *stream);
int pmpm_fill_interior(int32_t w_dtype, int32_t ct_dtype, const void *w, int64_t n, const double *center, double scale,
                       double dx, const double *lo, const int64_t *dims, double js, const pmpm_pcg64 *rng, float *out,
                       int64_t cap, int64_t *count, void *ws, size_t ws_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PMPM_H */
