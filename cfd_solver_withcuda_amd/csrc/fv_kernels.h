// Launchers of the native finite-volume kernels (fv_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cfd_types.h"

namespace cfd {

enum { MODE_RESID = 0, MODE_STAGE = 1 };

// Arguments of the flux pass. MODE_RESID uses R_out/dt_out/dtmin_out; MODE_STAGE uses the rest.
struct StageArgs {
    int rk;      // 3 or 4
    int stage;   // 0-based stage index
    int lts;     // per-cell time step (stored in hbuf at stage 0) instead of *dt_global
    int last;    // last stage: emit error partials / next-step dt
    const double4 *Un;   // state at the start of the step
    double4 *Uout;       // stage output
    const double4 *acc_in;  // RK4 accumulator read by stages 1..3
    double4 *acc_out;       // RK4 accumulator written by stages 0..2 (may equal acc_in)
    double *hbuf;        // per-cell dt/area (LTS)
    const double *dt_global;
    double *R_out;
    double *dt_out;
    double *dt_partial;  // [ntiles] per-tile minimum of the local time steps
    double *err_partial; // [4*ntiles] per-tile sums of the reference error terms
    const double4 *Uexit;  // exit ghosts copy this state when set (reference RK stages)
    // Limiter freezing (order 2): freeze[instance] = 0 off, 1 record psi, 2 use recorded psi;
    // psi [nfe][8] = slope / (grad W . d) per face entry, side (L, R) and variable (FP32).
    const int *freeze;
    float *psi;
    // Positivity fallback of native MUSCL stages: the detecting launch writes pflag[cell] (1 = the
    // stage state failed the admissibility check) and marks in dirty[tile] the tiles owning such a
    // cell or one of its face neighbours; a compaction launch lists the marked tiles in
    // dirty[ntiles .. 2 ntiles) with their count in dirty[2 ntiles + cur] (and zeroes
    // dirty[2 ntiles + next], the next stage's count, so that no count is reset while a launch
    // still reads it); the recomputing launch walks that list with persistent workgroups, redoes
    // those tiles (with first-order states on the faces of flagged cells) and clears their marks.
    // One allocation (2 ntiles + 4 ints), so that the stage kernels take no further arguments.
    uint8_t *pflag;
    int *dirty;
    int pos_range;  // bit 0: the check includes the neighbourhood range test; bits 1-2 cur, 3-4 next
};

size_t grad_smem_bytes(const DevMesh &m, int grad_type);
size_t flux_smem_bytes(const DevMesh &m, int order, int grad_lds);
// LDS of one flux launch (geom: the launch forms time steps and stages the face geometry).
size_t flux_launch_smem(const DevMesh &m, int order, int grad_lds, int block, bool geom);
// Positivity fallback (order 2, MODE_STAGE, inviscid; fv_pos.hip): pass 1 = the stage with
// detection, pass 2 = the recomputation of marked tiles. Inputs and outputs of a stage must not
// alias (pass 2 reads the stage's inputs again).
hipError_t launch_flux_pos(const DevMesh &m, const DevCfg &c, int flux, int lim, int pass,
                           const double4 *Us, const double *grad, const StageArgs &st, int block,
                           hipStream_t s);
hipError_t launch_grad(const DevMesh &m, const DevCfg &c, int grad_type, const double4 *Us,
                       double *grad, int block, hipStream_t s);
hipError_t launch_flux(const DevMesh &m, const DevCfg &c, int flux, int order, int lim, int mode,
                       const double4 *Us, const double *grad, const StageArgs &st, int block,
                       hipStream_t s);
// The viscous variants of the flux pass (fv_visc.hip), launched by launch_flux when
// c.viscous is set at order 2; sm is the dynamic LDS of the launch.
hipError_t launch_flux_visc(const DevMesh &m, const DevCfg &c, int flux, int lim, int mode,
                            const double4 *Us, const double *grad, const StageArgs &st, int block,
                            size_t sm, hipStream_t s);
// WENO5 variants (fv_weno.hip), order 3: every flux, MUSCL with limiter `lim` on faces without
// a complete line stencil.
hipError_t launch_flux_weno(const DevMesh &m, const DevCfg &c, int flux, int lim, int mode,
                            const double4 *Us, const double *grad, const StageArgs &st, int block,
                            size_t sm, hipStream_t s);
// WENO5 with viscous terms (fv_weno_visc.hip).
hipError_t launch_flux_weno_visc(const DevMesh &m, const DevCfg &c, int flux, int lim, int mode,
                                 const double4 *Us, const double *grad, const StageArgs &st, int block,
                                 size_t sm, hipStream_t s);
hipError_t launch_fill(double *p, double v, int n, hipStream_t s);
hipError_t launch_err_finalize(const double *partial, int n, double *err4, hipStream_t s);
hipError_t launch_ghosts(int n, const int *gcell, const double2 *gnrm, const int *gkind,
                         const DevCfg &c, const double4 *Us, const double4 *Uexit, double4 *Ug,
                         hipStream_t s);
hipError_t launch_batch_status(int ninst, const int *inst_tile, const double *err_partial, double tol,
                               int *status, int *steps, double *err4, int *freeze, double *best,
                               int *since, double freeze_tol, int freeze_stall, hipStream_t s);
// dst = src over the cells of the tiles still marching (m.tile_list, m.tile_count); grid_tiles
// bounds the marching count.
hipError_t launch_batch_commit(const DevMesh &m, int grid_tiles, const double4 *src, double4 *dst, hipStream_t s);
hipError_t launch_batch_list(int ninst, int ntiles, const int *inst_tile, const int *tile_inst,
                             const int *status, int *inst_off, int *tile_list, int *tile_count,
                             hipStream_t s);
// Freezing state machine of one solve after a step whose error norm is err4[0] (see
// fv_kernels.hip, freeze_update_kernel).
hipError_t launch_freeze_update(const double *err4, int *freeze, double *best, int *since,
                                double freeze_tol, int freeze_stall, hipStream_t s);
// The reference's per-cell arrays (renumbered cells; neighbours >= nc are ghosts nc + k).
struct RefCells {
    const int *nf;        // [nc] faces per cell
    const int *nbr;       // [4 nc]
    const double *dist;   // [4 nc] centre-to-neighbour-centre distances
    const double *nrm;    // [8 nc] unit normals
    const double *len;    // [4 nc] face lengths
    const double *area;   // [nc]
};
// Evaluate_Cell_Net_Flux_2O with ROE_2O: R, dt (and per-block dt minima into dt_partial,
// ceil(nc / 256) of them, when not null) from U_all = physical cells then ghosts.
hipError_t launch_roe2o(int nc, const RefCells &rc, const double4 *U_all, double zeta, double zeta1,
                        double cfl, double4 *R, double *dt, double *dt_partial, hipStream_t s);
hipError_t launch_min_reduce(const double *partial, int n, double *out, hipStream_t s);
hipError_t launch_ref_update(int nc, int kind, int lts, const double *inv_area, const double4 *Un,
                             const double4 *Us, const double4 *R, const double *dt_cell,
                             const double *dt_global, double4 *Uout, double *err_partial,
                             hipStream_t s);

}  // namespace cfd
