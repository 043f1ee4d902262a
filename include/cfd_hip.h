/* C-ABI of libcfdhip.so: the MI355X (gfx950) implementation of the explicit finite-volume
 * residual-and-update path of rameshkolluru43/CFD_Solver_withCUDA.
 *
 * Two families of entry points:
 *   1. Native, mesh-handle based (cfd_*): the reference's CPU driver path
 *      (Form_Cells -> Apply_Boundary_Conditions -> Evaluate_Cell_Net_Flux_1O/2O ->
 *      Runge_Kutta_Method / Explicit_Method -> Estimate_Error -> Update) on a tiled, renumbered
 *      mesh resident in HBM.
 *   2. Reference-layout kernels (launch_*): the reference's CUDA wrapper names and argument
 *      layouts (per-cell 4-face arrays, caller-owned device pointers), declared in
 *      cfd_hip_ref.h.
 *
 * Conventions: every function returns 0 on success, a positive hipError_t value for a HIP
 * failure, or a negative CFD_E_* code for invalid arguments (cfd_last_error() gives a message).
 * Stream arguments are hipStream_t passed as void* (NULL = the null stream). State arrays at
 * this interface are in the caller's (input-file) cell order; the library renumbers internally.
 */
#ifndef CFD_HIP_H
#define CFD_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_E_ARG (-1)
#define CFD_E_MESH (-2)
#define CFD_E_STATE (-3)
#define CFD_E_NOTCONV (-4)

typedef struct cfd_mesh_s cfd_mesh_t;
typedef struct cfd_solver_s cfd_solver_t;

/* Scheme selection (the reference's Dissipation_Type / Is_Second_Order / Time_Accurate /
 * Local_Time_Stepping / CFL flags, src/Configuration_Read.cpp:23-87, made explicit). */
typedef struct {
    int flux;      /* 0 Roe (Harten entropy fix), 1 HLLC, 2 LLF, 3 AUSM */
    int order;     /* 1 first order, 2 MUSCL, 3 WENO5 (WENO_Reconstruction, src/WENO2D.cpp:9-79,
                      on the conserved variables along lines of quadrilaterals through opposite
                      faces; MUSCL on faces without a complete five-cell line; native iteration) */
    int limiter;   /* 0 none, 1 minmod, 2 Van Leer, 3 superbee, 4 van Albada (differentiable form,
                      slopes of differences below 1e-2 not limited; MUSCL faces) */
    int gradient;  /* 0 Green-Gauss, 1 weighted least squares (MUSCL faces) */
    int roe_bdiss; /* Roe dissipation on boundary faces (0 = the reference's behaviour) */
    double cfl;
    int bc[6];     /* ghost type for face kinds 1..5 (LEFT, RIGHT, TOP, BOTTOM, WALL):
                      1 inlet (rho,P given, interior velocity), 2 exit (copy), 3 slip wall,
                      4 symmetry, 5 freestream, 6 characteristic far field,
                      7 the reference's subsonic inlet, 8 no-slip adiabatic wall (velocity
                      negated, rho and P copied), 9 back-pressure exit (characteristic, static
                      pressure p_exit; supersonic outflow copies), 10 total-condition subsonic
                      inlet (total pressure, total density and direction of fs; static pressure
                      from the interior) */
    double fs[4];  /* prescribed primitive state (rho, u, v, p) */
    int rk;        /* 1 forward Euler (reference mode), 3 TVD-RK3, 4 classical RK4 */
    int lts;       /* local time stepping */
    int ref_mode;  /* 1: the reference's iteration (per-stage global dt, delta form, LTS in
                      stage 1 only, exit ghosts from U^n); 0: native fused stages */
    int block;     /* workgroup size (64..1024, power of two; 256 default) */
    int xcd_remap; /* order of tiles over the 8 XCDs (workgroups are dealt to them in turn):
                      0 block order, 1 one contiguous run of tiles per XCD, k = 2..4096 runs of
                      k consecutive tiles dealt to the XCDs in turn; results do not depend on it */
    int grad_lds;  /* stage the tile's gradients in LDS (1) or read them through L2 (0) */
    /* Limiter freezing for steady marches (order 2, native mode): once Error[0] <= freeze_tol,
     * or once Error[0] has not dropped below 0.9 x its best value for freeze_stall steps, the
     * next step records each face's limiter factor psi = slope / (grad W . d) and later steps
     * reuse it (slope = psi * grad W . d). 0 and 0 disable it. */
    double freeze_tol;
    int freeze_stall;
    /* Reference iteration (ref_mode = 1) at order 2 with the Roe flux: the reference's
     * Evaluate_Cell_Net_Flux_2O (Roe dissipation on MinMod-limited differences,
     * src/Roe_Scheme.cpp:257-447, src/Limiters.cpp:97-200) with Limiter_Zeta / Limiter_Zeta1. */
    double zeta, zeta1;
    /* Laminar viscous terms (order 2 or 3, native mode): 0 off (Euler), 1 constant viscosity mu,
     * 2 Sutherland's law in T = p / (rho r_gas). Stress tensor and heat flux as the reference's
     * Viscous_Flux_on_Face (src/Grid_Refine_Functions.cpp:418-479) with dimensional mu,
     * conductivity k = mu gamma r_gas / ((gamma - 1) prandtl); the local time step adds the
     * viscous spectral radius of Viscous_Time_Step_1 (src/Time_Step.cpp:103-109). */
    int viscous;
    double mu, prandtl, r_gas;
    /* Positivity fallback of the native MUSCL stages (order 2, inviscid; single and batched
     * solvers): after each stage, a cell whose new state has a non-positive density or pressure,
     * or a density or
     * pressure outside 1e-3 max(|hi|, hi - lo) of the range [lo, hi] of its own and its
     * face neighbours' stage inputs, has the stage recomputed with first-order states on all its
     * faces (an a-posteriori, MOOD-type check; no counterpart in the reference, whose second-order
     * schemes have no safeguard). 2 applies the positivity test only, without the range test. The
     * range's bounds are the stage inputs rounded to FP32. With it on, a single solver's device
     * buffer of the state alternates between two arrays from step to step (query cfd_state_device
     * after each step); a batch copies the last stage's output back over its marching tiles. */
    int positivity;
    /* Static pressure of back-pressure exits (ghost type 9); 0 = the prescribed state's p. */
    double p_exit;
} cfd_config_t;

const char *cfd_last_error(void);
int cfd_version(void);

/* Device-memory helpers for callers without their own HIP runtime bindings. */
int cfd_device_alloc(size_t bytes, void **ptr);
int cfd_device_free(void *ptr);
int cfd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int cfd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int cfd_synchronize(void);

/* ---- mesh -------------------------------------------------------------------------------
 * Builds the reference connectivity (Read_VTK_Grid + Form_Cells, src/Read_Gmsh_File.cpp:11-621,
 * src/Grid_Computations.cpp:319-827) from points [3*npts] and cells [4*ncells] (triangles padded
 * with -1), renumbers cells (renumber = 1: Hilbert order) into tiles of at most `tile` cells and
 * `face_budget` faces (0 = no face limit) and uploads the face lists to the device. */
int cfd_mesh_create(int npts, const double *pts3, int ncells, const int *cells4, int tile,
                    int face_budget, int renumber, cfd_mesh_t **out);
/* Same, host only (no device upload; such a mesh cannot back a solver). */
int cfd_mesh_create_host(int npts, const double *pts3, int ncells, const int *cells4, int tile,
                         int face_budget, int renumber, cfd_mesh_t **out);
void cfd_mesh_destroy(cfd_mesh_t *m);
/* out[9] = nphys, nghost, n_inlet, n_exit, n_wall, nfaces, ntiles, nface_entries, max_tile_faces */
int cfd_mesh_counts(const cfd_mesh_t *m, int *out9);
/* The reference's per-cell arrays (4 slots per cell) and boundary triplet lists. */
int cfd_mesh_export(const cfd_mesh_t *m, int *nf, int *nodes4, int *nbr4, double *nrm8,
                    double *len4, double *ctr3_all, double *area_all, double *inv_area_all,
                    double *dist4, int *kind4, int *inlet3, int *exit3, int *wall3);
int cfd_mesh_permutation(const cfd_mesh_t *m, int *new_to_old);
/* Tiled face lists: tile_cstart/tile_fstart[ntiles+1], fL/fR[nface_entries] (renumbered cells;
 * fR = -kind on a boundary face), slots4[4*nphys] ((local face << 1) | cell-is-right, 0xFFFF
 * unused). */
int cfd_mesh_layout_export(const cfd_mesh_t *m, int *tile_cstart, int *tile_fstart, int *fL,
                           int *fR, unsigned short *slots4);

/* ---- solver -------------------------------------------------------------------------------- */
int cfd_solver_create(cfd_mesh_t *m, const cfd_config_t *cfg, cfd_solver_t **out);
void cfd_solver_destroy(cfd_solver_t *s);
int cfd_solver_configure(cfd_solver_t *s, const cfd_config_t *cfg);
/* conserved state U[4*nphys] (rho, rho u, rho v, rho E), host memory, caller's cell order */
int cfd_set_state(cfd_solver_t *s, const double *U);
int cfd_get_state(cfd_solver_t *s, double *U);
/* device pointer of the resident state (renumbered order, double4 per cell) */
int cfd_state_device(cfd_solver_t *s, void **d_U);

/* One residual evaluation of the current state: R[4*nphys] (net flux, the reference's
 * Cells_Net_Flux), dt[nphys] (local time step, Cells[i].del_t), grad[8*nphys] (primitive
 * gradients, order 2; may be NULL). Host outputs in the caller's order. */
int cfd_residual(cfd_solver_t *s, double *R, double *dt, double *grad, void *stream);

/* Allocation-free residual. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It runs cfd_residual's kernels on a caller's device state d_U (renumbered cell order,
 * 4 doubles per cell; cfd_mesh_permutation gives the order) and writes device outputs: d_R
 * [4*nphys] net flux and d_dt [nphys] local time step (may be NULL), both renumbered, and d_dtmin
 * [1] their minimum (may be NULL). The gradients and the time-step partials live in a device
 * array of the library (one per GPU, held until the process ends), so meshes are limited to
 * CFD_NOALLOC_MAX_CELLS physical cells; larger ones are rejected with CFD_E_ARG. Order 1 or 2,
 * native iteration (ref_mode 0); the mesh must live on the current device. Calls that overlap on
 * one GPU (several streams or threads at once) share that array and race; calls in sequence do
 * not, since every call overwrites what it reads. */
#define CFD_NOALLOC_MAX_CELLS (1 << 20)
int cfd_residual_noalloc(const cfd_mesh_t *m, const cfd_config_t *cfg, const void *d_U, void *d_R,
                         void *d_dt, void *d_dtmin, void *stream);

/* Sets the limiter-freezing state: 0 off (limiter evaluated), 1 record psi on the next step,
 * 2 use the recorded psi. With freeze_tol = freeze_stall = 0 the state only changes here. */
int cfd_set_freeze(cfd_solver_t *s, int mode);

/* Ghost states of the current state (Apply_Boundary_Conditions, src/Boundary_Conditions.cpp:12-61):
 * U_ghost[4*nghost] (host, may be NULL to keep them on the device) in the reference's ghost
 * order, ghost k = cell id nphys + k. The residual and step entries form these states on the fly;
 * this entry exposes them. */
int cfd_apply_bc(cfd_solver_t *s, double *U_ghost, void *stream);

/* nsteps solver iterations. err_hist[4*nsteps] (the reference's Error[0..3] per step) and
 * dt_hist[nsteps] (the step's global dt, or min local dt with LTS) may be NULL; without them the
 * steps run without host synchronisation. */
int cfd_step(cfd_solver_t *s, int nsteps, double *err_hist, double *dt_hist, void *stream);

/* Steady march until Error[0] <= tol (checked every check_every steps) or max_steps.
 * Returns 0 when converged, CFD_E_NOTCONV otherwise; *steps and err4 report the end point. */
int cfd_run_steady(cfd_solver_t *s, double tol, int max_steps, int check_every, int *steps,
                   double *err4, void *stream);

/* Times `reps` back-to-back launches on `stream` with HIP events: which = 0 the gradient pass,
 * 1 the fused flux/stage pass (first stage, output to scratch; the state is unchanged; with the
 * positivity fallback its detecting pass),
 * 2 whole steps (advances the state), 3 the flux pass in residual mode (R and the local time
 * step out, as cfd_residual). *ms = total elapsed milliseconds. */
int cfd_time_pass(cfd_solver_t *s, int which, int reps, float *ms, void *stream);

/* ---- batched steady solves ------------------------------------------------------------------
 * ninst independent meshes concatenated into one device mesh and marched together with local
 * time stepping until each instance's Error[0] <= tol. Instance i has points
 * pts3[3*pt_off[i] .. 3*pt_off[i+1]) and cells cells4[4*cell_off[i] .. 4*cell_off[i+1]) (node ids
 * local to the instance) and its own prescribed state fs4[4*i..4*i+3] (rho, u, v, p) for the
 * inlet / freestream / far-field ghosts; the rest of cfg is shared (cfg->fs is ignored;
 * cfg->lts must be 1, rk 3 or 4). cfg->block = 0 autotunes (tile, face budget, block) on the
 * largest instances; 64/128/256 fixes block = tile and face budget = 2*block. Meshes are built
 * on `nthreads` host threads. The state starts at each instance's freestream. */
typedef struct cfd_batch_s cfd_batch_t;
int cfd_batch_create(int ninst, const int *pt_off, const double *pts3, const int *cell_off,
                     const int *cells4, const double *fs4, const cfd_config_t *cfg, int nthreads,
                     cfd_batch_t **out);
void cfd_batch_destroy(cfd_batch_t *b);
/* out8 = ninst, cells, tiles, face entries, tile, face budget, block, max faces per tile;
 * secs2 = mesh build seconds, autotune seconds (either may be NULL) */
int cfd_batch_info(const cfd_batch_t *b, int *out8, double *secs2);
/* Where the launch shape came from: *source = 0 given (cfg->block), 1 timed by the autotuner
 * (each candidate marched 200 steps from freestream, then 5 timed steps), 2 read from the autotune
 * cache (a JSON file, CFD_AUTOTUNE_CACHE or $HOME/.cache/cfdhip_autotune.json, keyed by the GPU's
 * architecture and compute units and the batch's instance count, total and largest cell counts
 * in powers of two, and scheme). */
int cfd_batch_tuning(const cfd_batch_t *b, int *source);
/* U[4*total cells] conserved, instances concatenated in order, each in its caller's cell order */
int cfd_batch_set_state(cfd_batch_t *b, const double *U);
int cfd_batch_get_state(cfd_batch_t *b, double *U);
/* March all instances (one hipGraph launch per step; the host reads the instance status every
 * check_every steps). Per instance: steps[ninst] taken, err4[4*ninst] last Error[0..3],
 * status[ninst] (0 converged, 1 still marching at max_steps, -1 non-finite), conv_s[ninst]
 * wall seconds at the check where it was first seen converged; *wall_s total. Any output may be
 * NULL. Returns 0 when every instance converged, CFD_E_NOTCONV otherwise. */
int cfd_batch_solve(cfd_batch_t *b, double tol, int max_steps, int check_every, int *steps,
                    double *err4, int *status, double *conv_s, double *wall_s, void *stream);
/* Per-instance limiter-freezing mode that every later solve starts with (needs freeze_tol or
 * freeze_stall in cfg): 0 = evaluated limiter until the freezing trigger (the default), 1 = record
 * the limiter factors on the first step and march frozen from the second (for starts already
 * close to the solution). mode = NULL restores 0 for all. No counterpart in the reference. */
int cfd_batch_set_freeze(cfd_batch_t *b, const int *mode);
/* Progress report of long solves: fn(steps done, instances still marching, user) after every
 * `every_checks` status checks of cfd_batch_solve (fn = NULL or every_checks = 0: none). It runs
 * on the calling thread, between graph launches. */
typedef void (*cfd_progress_fn)(int steps_done, int marching, void *user);
int cfd_batch_set_progress(cfd_batch_t *b, cfd_progress_fn fn, void *user, int every_checks);
/* Courant number of later solves (the local time steps scale with it). */
int cfd_batch_set_cfl(cfd_batch_t *b, double cfl);
/* How later solves launch the gradient and flux/stage passes (the captured step is rebuilt):
 * 0 (default) one workgroup per tile of the batch, those of converged instances exiting at once;
 * 1 persistent workgroups, as many as the GPU holds at once, each walking the tiles still
 * marching (order 1 or 2, inviscid). Results are identical. */
int cfd_batch_set_launch(cfd_batch_t *b, int mode);
/* Error[0] of every instance at each check of the last solve: *nchecks checks, e0[check*ninst + i]
 * (e0 may be NULL to query the count). The batched counterpart of the reference's per-iteration
 * Error rows (src/output_files.cpp:558-575). */
int cfd_batch_history(const cfd_batch_t *b, int *nchecks, double *e0);
/* Times nsteps graph-launched steps of all instances (advances the state). */
int cfd_batch_time_steps(cfd_batch_t *b, int nsteps, float *ms);

/* ---- coordinate-MLP surrogate (warm start) ---------------------------------------------------
 * n_in features -> depth SiLU hidden layers of `width` (depth 1..15) -> n_out outputs; BF16 weights and
 * activations on MFMA with FP32 accumulation; Adam on FP32 master weights. Parameters are
 * concatenated per layer: W_l row-major [out x in], then b_l. Features and targets are float
 * [n x n_in] / [n x n_out] host arrays, normalised inside as (v - mean) * scale (targets and
 * predictions in physical units at this interface). max_rows bounds a minibatch (rounded up
 * to a multiple of 512). */
typedef struct cfd_mlp_s cfd_mlp_t;
int cfd_mlp_create(int n_in, int width, int depth, int n_out, int max_rows, unsigned long long seed,
                   cfd_mlp_t **out);
void cfd_mlp_destroy(cfd_mlp_t *m);
int cfd_mlp_num_params(const cfd_mlp_t *m, long long *n_weights, long long *n_biases);
int cfd_mlp_set_params(cfd_mlp_t *m, const float *W, const float *b);  /* resets Adam moments */
int cfd_mlp_get_params(cfd_mlp_t *m, float *W, float *b);
int cfd_mlp_set_normalization(cfd_mlp_t *m, const float *in_mean, const float *in_scale,
                              const float *out_mean, const float *out_scale);
/* One forward + backward over all n rows (n a multiple of 512): gradients of the mean squared
 * error of the normalised outputs, and that loss. */
int cfd_mlp_gradients(cfd_mlp_t *m, const float *feat, const float *targ, int n, float *gW,
                      float *gb, double *loss);
/* `steps` Adam steps on minibatches of `batch` rows drawn (hash of seed, step, row) from the n
 * samples; learning rate cosine-decays from lr to lr_final. loss_hist[ceil(steps/log_every)]
 * (may be NULL) records the minibatch loss of every log_every-th step. */
int cfd_mlp_train(cfd_mlp_t *m, const float *feat, const float *targ, int n, int steps, int batch,
                  float lr, float lr_final, unsigned long long seed, float *loss_hist, int log_every);
/* How cfd_mlp_train launches a step: 0 (default) kernel by kernel, 1 one training step captured
 * in a hipGraph and replayed (step index, learning rate and bias corrections read from device
 * memory; for callers that replay it themselves). Results are identical. */
int cfd_mlp_set_launch(cfd_mlp_t *m, int mode);
/* Predictions out[n x n_out] (physical units) for feat[n x n_in]. */
int cfd_mlp_predict(cfd_mlp_t *m, const float *feat, int n, float *out);
/* Times `reps` forward + backward passes of `batch` rows on the training data (HIP events). */
int cfd_mlp_time_step(cfd_mlp_t *m, int batch, int reps, float *ms);

#ifdef __cplusplus
}
#endif
#endif
