/* CPU FP64 oracle for the explicit finite-volume residual-and-update path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * path (cfd_solver_withcuda_amd) never calls it.
 *
 * It restates, from the reference rameshkolluru43/CFD_Solver_withCUDA:
 *   - connectivity/geometry of Read_VTK_Grid + Form_Cells (src/Read_Gmsh_File.cpp:11-621,
 *     src/Grid_Computations.cpp:319-442,523-827,957-1142);
 *   - primitive recovery (src/Basic_Functions.cpp:85-142);
 *   - ghost-cell BCs (src/Wall_Boundary_Conditions.cpp:8-124, src/Inlet_Boundary_Conditions.cpp:7-74,
 *     src/Exit_Boundary_Conditions.cpp:68-94);
 *   - average flux + Roe / LLF dissipation (src/Average_Interface_Flux.cpp:20-82,
 *     src/Roe_Scheme.cpp:6-256, src/LLF.cpp:8-83), Roe-2O limited differences
 *     (src/Roe_Scheme.cpp:257-447, src/Limiters.cpp:58-200), AUSM (src/Ausm_Flux.cpp:5-204);
 *   - local time step (src/Time_Step.cpp:293-329), TVD-RK3 delta form and forward Euler
 *     (src/Numerical_Method.cpp:26-103,322-457), error norm (src/Error_Estimate_Update.cpp:10-44);
 *   - Green-Gauss gradient (src/Grid_Refine_Functions.cpp:221-246);
 *   - HLLC (CUDA_KERNELS/HLLC_LLF_Flux_Cuda_Kernels.cu:25-208) and the limiter functions
 *     (CUDA_KERNELS/Limiter_Cuda_Kernels.cu:89-157), applied to a proper slope ratio.
 * and defines the pieces the reference lacks (face-based MUSCL, WLSQ gradients, classical RK4,
 * LTS-consistent RK3, freestream/far-field ghosts).
 *
 * Parity pins: see the header of cfd_oracle.c for which test pins each function.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OMesh OMesh;

/* ---- mesh ---------------------------------------------------------------------------- */
OMesh *om_build(int npts, const double *pts3, int ncells, const int *cells4);
void om_free(OMesh *m);
int om_nfaces(const OMesh *m);
void om_counts(const OMesh *m, int *nphys, int *nghost, int *n_in, int *n_ex, int *n_wall);
void om_export(const OMesh *m, int *nf, int *nodes4, int *nbr4, double *nrm8, double *len4,
               double *ctr3_all, double *area_all, double *inv_area_all, double *dist4, int *kind4,
               int *inl, int *exl, int *wal);

/* ---- boundary-condition types for face kinds ------------------------------------------- */
enum {
    OBC_INLET_REF = 1,  /* rho, P prescribed, interior velocity (reference supersonic inlet) */
    OBC_EXIT = 2,       /* copy interior state (reference supersonic exit) */
    OBC_WALL = 3,       /* slip wall, mirrored velocity (reference inviscid wall) */
    OBC_SYMMETRY = 4,   /* same arithmetic as the reference symmetry condition */
    OBC_FREESTREAM = 5, /* all four primitives prescribed */
    OBC_FARFIELD = 6,   /* characteristic (Riemann-invariant) far field */
    OBC_SUBSONIC_INLET_REF = 7, /* reference subsonic inlet, incl. its energy expression */
    OBC_WALL_NOSLIP = 8, /* no-slip adiabatic wall: velocity negated, rho and p copied */
    OBC_SUBSONIC_EXIT = 9, /* back-pressure exit (Subsonic_Exit_Boundary_Condition's characteristic
                              face state, ghost 2 U_face - U_cell) */
    OBC_SUBSONIC_INLET = 10 /* total-condition inlet: total p and rho and direction of fs, static p
                               from the interior */
};
enum { OFLUX_ROE = 0, OFLUX_HLLC = 1, OFLUX_LLF = 2, OFLUX_AUSM = 3 };
enum { OLIM_NONE = 0, OLIM_MINMOD = 1, OLIM_VANLEER = 2, OLIM_SUPERBEE = 3, OLIM_VANALBADA = 4 };
enum { OGRAD_GG = 0, OGRAD_WLSQ = 1 };

typedef struct {
    int flux;        /* OFLUX_* */
    int order;       /* 1, 2 (MUSCL) or 3 (WENO5 on lines of quads, MUSCL elsewhere) */
    int limiter;     /* OLIM_* */
    int gradient;    /* OGRAD_* */
    int roe_bdiss;   /* 1: Roe dissipation on boundary faces; 0: none (reference behaviour) */
    double cfl;
    int bc[6];       /* bc[kind] for face kinds 1..5 (LEFT, RIGHT, TOP, BOTTOM, WALL) */
    double fs[4];    /* prescribed primitive state (rho, u, v, p) for inlet/freestream/far field */
    int freeze;      /* limiter freezing (order 2): 0 off, 1 record phi = slope / (grad W . d),
                        rounded to FP32, into phi; 2 use the recorded phi (slope = phi * grad W . d) */
    float *phi;      /* [8 * faces]: per face, side L (4 variables) then side R */
    int viscous;     /* laminar viscous terms (order 2): 0 off, 1 constant mu, 2 Sutherland */
    double mu;       /* dynamic viscosity (viscous == 1) */
    double prandtl;  /* Prandtl number: conductivity k = mu gamma r_gas / ((gamma - 1) Pr) */
    double r_gas;    /* gas constant: T = p / (rho r_gas) */
    int positivity;  /* a-posteriori first-order fallback of the native MUSCL stages (order 2): a
                        cell whose stage state has a non-positive density or pressure, or a density
                        or pressure outside the relaxed range of its own and its face neighbours'
                        stage-input values, is recomputed with first-order states on all its faces;
                        2: the positivity test alone (no range test) */
    double p_exit;   /* static pressure of back-pressure exits (OBC_SUBSONIC_EXIT); 0: fs[3] */
} ONatCfg;

/* ---- primitive / ghost helpers --------------------------------------------------------- */
void or_prim(const double *U4, double *prim11);
/* WENO5-JS of one variable (src/WENO2D.cpp:9-79); shift 0 right edge of c, 1 left edge. */
double or_weno5(double a, double b, double c, double d, double e, int shift);
void on_ghost_state(int bc, const double *Win5, double nx, double ny, const double *fs5,
                    double *Wg5);
/* conserved ghost state of a cell with conserved state U4 across a face with unit normal n */
void on_ghost_cons(int bc, const double *U4, double nx, double ny, const double *fs5, double *Ug4);
double or_limiter_slope(int limiter, double dminus, double dplus);

/* ---- face-loop residual (the spec of the GPU kernels) ------------------------------------
 * U[4*nphys] conserved state. R[4*nphys] out, dt[nphys] out (local time step from the
 * cell-centre state), grad[8*nphys] out (primitive gradients; only written when order 2).
 * U_exit_src: if non-NULL, exit ghosts copy this state instead of U (the reference's RK
 * stages copy U_Cells, which still holds U^n). */
void on_residual(const OMesh *m, const ONatCfg *cfg, const double *U, const double *U_exit_src,
                 double *R, double *dt, double *grad);
/* cells recomputed at first order by the positivity fallback during the last on_step */
int on_last_fallback_cells(void);
/* threads of the residual's parallel loops (OMP_NUM_THREADS; all host threads by default) */
int or_threads(void);
void on_gradients(const OMesh *m, const ONatCfg *cfg, const double *W5, double *grad);
void on_face_flux(const ONatCfg *cfg, int boundary, const double *WL5, const double *WR5,
                  double nx, double ny, double dl, double *F4);

/* ---- time integration -------------------------------------------------------------------
 * on_step:   native step. rk = 3 (Shu-Osher TVD-RK3) or 4 (classical RK4); lts = 1 uses the
 *            per-cell dt from U^n in every stage, lts = 0 uses min_i dt_i (from U^n).
 *            err4 = reference Error norm of dU, dt_out = global dt (or min local dt).
 * on_refstep: the reference's iteration (BC, RK3 delta form with per-stage global dt and
 *            LTS only in stage 1, or forward Euler when rk == 1), residual by the face loop. */
void on_step(const OMesh *m, const ONatCfg *cfg, int rk, int lts, double *U, double *err4,
             double *dt_out);
void on_refstep(const OMesh *m, const ONatCfg *cfg, int rk, int lts, double *U, double *err4,
                double *dt_out);
void or_error(int n, const double *dU, const double *U, double *err4);
void or_rk_stage(int rk, int stage, int n, const double *Un, const double *Us, double *acc,
                 const double *R, const double *h, double *out);

/* ---- reference per-cell functions (cell-centric, same arithmetic order) --------------- */
void or_gg_gradient(const OMesh *m, const double *prim_all11, int var, double *grad2);
void or_ausm_cell(const OMesh *m, const double *U_all, double *R);
void or_roe2o_residual(const OMesh *m, const double *U_all, const double *prim_all11,
                       double zeta, double zeta1, double cfl, double *R, double *dt);
void or_apply_bc_ref(const OMesh *m, int inlet_bc, double rho_in, double p_in, double *U_all,
                     double *prim_all11, const double *U_exit_src);

#ifdef __cplusplus
}
#endif
#endif
