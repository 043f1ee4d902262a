/* Reference-layout entry points of libcfdhip.so.
 *
 * Same names, argument order and array layouts as the CUDA wrappers of
 * rameshkolluru43/CFD_Solver_withCUDA (CUDA_KERNELS/*.h), as a C ABI:
 *   - return int: 0 (hipSuccess), a positive hipError_t, or CFD_E_ARG (-1) for bad arguments;
 *     num_cells (or list length) 0 returns 0 without launching, as the reference's do;
 *   - the reference's trailing `int block_size = 256` is an explicit argument (0 means 256),
 *     followed by a hipStream_t passed as void* (NULL = null stream);
 *   - C++ references to the BC structs become const pointers.
 * Arrays are caller-owned device memory: U[N*4] conserved, P[N*6] = (rho,u,v,P,T,a),
 * neighbours[N*4] (a value outside [0,N) marks a boundary face), normals[N*8], face lengths[N*4],
 * per-face outputs [(N*4 + face)*4 + var].
 */
#ifndef CFD_HIP_REF_H
#define CFD_HIP_REF_H

#ifdef __cplusplus
extern "C" {
#endif

/* Limiter_Cuda_Kernels.h:21-29 */
typedef enum {
    LIMITER_MINMOD = 0,
    LIMITER_VANLEER = 1,
    LIMITER_SUPERBEE = 2,
    LIMITER_VANALBADA = 3,
    LIMITER_VENKATAKRISHNAN = 4,
    LIMITER_BARTH_JESPERSEN = 5
} LimiterType;

/* Boundary_Conditions_Cuda_Kernels.h:21-48 */
typedef struct { double Rho, u, v, P, T; } InletCondition_GPU;
typedef struct { double P; } ExitCondition_GPU;
typedef struct { double Rho_inf, u_inf, v_inf, P_inf, Mach; } FarfieldCondition_GPU;

/* ---- fluxes: Advanced_Flux_Schemes_Cuda.h:71-115 -------------------------------------- */
int launch_roe_flux(const double *d_U_cells, const double *d_P_cells, const int *d_face_neighbors,
                    const double *d_face_normals, const double *d_face_areas, double *d_flux_output,
                    int num_cells, double gamma, int block_size, void *stream);
int launch_roe_flux_2nd_order(const double *d_U_cells, const double *d_P_cells,
                              const int *d_face_neighbors, const double *d_face_normals,
                              const double *d_face_areas, const double *d_limited_gradients,
                              const double *d_cell_centers, const double *d_face_centers,
                              double *d_flux_output, int num_cells, double gamma, int block_size,
                              void *stream);
int launch_hllc_flux(const double *d_U_cells, const double *d_P_cells, const int *d_face_neighbors,
                     const double *d_face_normals, const double *d_face_areas, double *d_flux_output,
                     int num_cells, double gamma, int block_size, void *stream);
int launch_llf_flux(const double *d_U_cells, const double *d_P_cells, const int *d_face_neighbors,
                    const double *d_face_normals, const double *d_face_areas, double *d_flux_output,
                    int num_cells, double gamma, int block_size, void *stream);

/* ---- limiters: Limiter_Cuda_Kernels.h:122-209 ------------------------------------------ */
int launch_minmod_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                          const double *d_cell_distances, double *d_limited_gradients, int num_cells,
                          double limiter_zeta, int use_three_arg, int block_size, void *stream);
int launch_vanleer_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                           const double *d_cell_distances, double *d_limited_gradients,
                           int num_cells, double limiter_zeta, int block_size, void *stream);
int launch_superbee_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                            const double *d_cell_distances, double *d_limited_gradients,
                            int num_cells, double limiter_zeta, int block_size, void *stream);
int launch_vanalbada_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                             const double *d_cell_distances, double *d_limited_gradients,
                             int num_cells, double limiter_zeta, int block_size, void *stream);
int launch_venkatakrishnan_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                                   const double *d_cell_distances, const double *d_cell_volumes,
                                   double *d_limited_gradients, int num_cells, double K_constant,
                                   int block_size, void *stream);
int launch_barth_jespersen_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                                   const double *d_cell_distances,
                                   const double *d_unlimited_gradients,
                                   double *d_limited_gradients, int num_cells, int block_size,
                                   void *stream);
int launch_generic_limiter(const double *d_U_cells, const int *d_cell_neighbors,
                           const double *d_cell_distances, double *d_limited_gradients,
                           int num_cells, LimiterType limiter_type, double limiter_zeta,
                           int block_size, void *stream);
const char *get_limiter_name(LimiterType limiter_type);

/* ---- reconstruction: Reconstruction_Schemes_Cuda.h:54-63 ------------------------------ */
int launch_muscl_reconstruction(const double *d_U_cells, const int *d_cell_neighbors,
                                const double *d_cell_distances, const double *d_limited_gradients,
                                double *d_Q_left, double *d_Q_right, int num_cells, double kappa,
                                int block_size, void *stream);

/* ---- boundary conditions (in place on the listed cells):
 *      Boundary_Conditions_Cuda_Kernels.h:138-218 --------------------------------------- */
int launch_subsonic_inlet_bc(double *d_U_cells, const int *d_inlet_cell_list,
                             const double *d_face_normals, const InletCondition_GPU *bc_data,
                             int num_inlet_cells, double gamma, int block_size, void *stream);
int launch_supersonic_inlet_bc(double *d_U_cells, const int *d_inlet_cell_list,
                               const InletCondition_GPU *bc_data, int num_inlet_cells,
                               double gamma, int block_size, void *stream);
int launch_subsonic_exit_bc(double *d_U_cells, const int *d_exit_cell_list,
                            const double *d_face_normals, const ExitCondition_GPU *bc_data,
                            int num_exit_cells, double gamma, int block_size, void *stream);
int launch_supersonic_exit_bc(double *d_U_cells, const int *d_exit_cell_list, int num_exit_cells,
                              int block_size, void *stream);
int launch_viscous_wall_bc(double *d_U_cells, const int *d_wall_cell_list, int num_wall_cells,
                           int block_size, void *stream);
int launch_inviscid_wall_bc(double *d_U_cells, const int *d_wall_cell_list,
                            const double *d_face_normals, int num_wall_cells, int block_size,
                            void *stream);
int launch_symmetry_bc(double *d_U_cells, const int *d_symmetry_cell_list,
                       const double *d_face_normals, int num_symmetry_cells, int block_size,
                       void *stream);
int launch_farfield_bc(double *d_U_cells, const int *d_farfield_cell_list,
                       const double *d_face_normals, const FarfieldCondition_GPU *bc_data,
                       int num_farfield_cells, double gamma, int block_size, void *stream);

/* ---- time integration. The reference has kernels only (runge_kutta_stage_kernel,
 *      CUDA_KERNELS/Cuda_Kernels.h:4-33; tvd_rk3_kernel, Time_Integration_Cuda_Kernels.cu:94);
 *      these are the wrappers in the same style. ----------------------------------------- */
/* delta-form TVD-RK3 stage (1..3) on U[N*4]: stage 1 Urk = U - dt/A (R - Rv); stage 2
 * Urk = 1/4 (Urk + 3U - dt/A (R - Rv)); stage 3 Urk = 2/3 (Urk - U - dt/A (R - Rv)) = dU.
 * d_viscous_flux may be NULL. */
int launch_tvd_rk3_stage(const double *d_U_cells, const double *d_net_flux,
                         const double *d_viscous_flux, double *d_U_cells_rk,
                         const double *d_inv_area, double min_dt, int num_cells, int stage,
                         int block_size, void *stream);
/* classical RK4 stage (1..4): stage 1 K = R, Us = U - dt/(2A) R; 2: K += 2R, Us = U - dt/(2A) R;
 * 3: K += 2R, Us = U - dt/A R; 4: Us = U - dt/(6A) (K + R) = U^{n+1}. */
int launch_rk4_stage(const double *d_U_cells, const double *d_net_flux, double *d_U_stage,
                     double *d_K, const double *d_inv_area, double dt, int num_cells, int stage,
                     int block_size, void *stream);

#ifdef __cplusplus
}
#endif
#endif
