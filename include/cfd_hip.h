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
    int order;     /* 1 first order, 2 MUSCL */
    int limiter;   /* 0 none, 1 minmod, 2 Van Leer, 3 superbee (order 2) */
    int gradient;  /* 0 Green-Gauss, 1 weighted least squares (order 2) */
    int roe_bdiss; /* Roe dissipation on boundary faces (0 = the reference's behaviour) */
    double cfl;
    int bc[6];     /* ghost type for face kinds 1..5 (LEFT, RIGHT, TOP, BOTTOM, WALL):
                      1 inlet (rho,P given, interior velocity), 2 exit (copy), 3 slip wall,
                      4 symmetry, 5 freestream, 6 characteristic far field,
                      7 the reference's subsonic inlet */
    double fs[4];  /* prescribed primitive state (rho, u, v, p) */
    int rk;        /* 1 forward Euler (reference mode), 3 TVD-RK3, 4 classical RK4 */
    int lts;       /* local time stepping */
    int ref_mode;  /* 1: the reference's iteration (per-stage global dt, delta form, LTS in
                      stage 1 only, exit ghosts from U^n); 0: native fused stages */
    int block;     /* workgroup size (64..1024, power of two; 256 default) */
    int xcd_remap; /* map consecutive tiles onto one XCD */
    int grad_lds;  /* stage the tile's gradients in LDS (1) or read them through L2 (0) */
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
 * 1 the fused flux/stage pass (first stage, output to scratch; the state is unchanged),
 * 2 whole steps (advances the state). *ms = total elapsed milliseconds. */
int cfd_time_pass(cfd_solver_t *s, int which, int reps, float *ms, void *stream);

#ifdef __cplusplus
}
#endif
#endif
