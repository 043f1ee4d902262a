// Plain data shared by the host library and the HIP kernels of the finite-volume path.
#pragma once
#include <stdint.h>

namespace cfd {

// Boundary-condition types (ghost states), assigned per reference face kind.
enum BcType {
    BC_NONE = 0,
    BC_INLET_REF = 1,           // rho, P prescribed; interior velocity
    BC_EXIT = 2,                // copy of the interior state
    BC_WALL = 3,                // slip wall, mirrored velocity
    BC_SYMMETRY = 4,            // same arithmetic as the wall
    BC_FREESTREAM = 5,          // all primitives prescribed
    BC_FARFIELD = 6,            // Riemann-invariant far field
    BC_SUBSONIC_INLET_REF = 7,  // the reference's subsonic inlet
    BC_WALL_NOSLIP = 8,         // no-slip adiabatic wall: velocity negated, rho and p copied
    BC_SUBSONIC_EXIT = 9,       // back-pressure exit (characteristic, prescribed static pressure)
    BC_SUBSONIC_INLET = 10      // total-condition inlet (total p, total rho, direction prescribed)
};
// Reference face kinds (src/Read_Gmsh_File.cpp:318-323).
enum FaceKind { KIND_INTERNAL = 0, KIND_LEFT = 1, KIND_RIGHT = 2, KIND_TOP = 3, KIND_BOTTOM = 4, KIND_WALL = 5 };
enum FluxType { FLUX_ROE = 0, FLUX_HLLC = 1, FLUX_LLF = 2, FLUX_AUSM = 3 };
enum LimiterType { LIM_NONE = 0, LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_SUPERBEE = 3, LIM_VANALBADA = 4 };
enum GradType { GRAD_GG = 0, GRAD_WLSQ = 1 };

// Device view of a tiled mesh. Cells are renumbered so that tiles are contiguous ranges of
// `tile` cells; each tile lists every face touching one of its cells (faces between two tiles
// appear in both lists with identical data, so both tiles compute identical fluxes).
struct DevMesh {
    int nc;        // physical cells
    int ntiles;
    int tile;      // maximum cells per tile
    int max_tile_faces;
    int nfe;       // face entries (all tiles)
    const int *tile_cstart;   // [ntiles+1] first cell of each tile
    const int *tile_fstart;   // [ntiles+1] first face entry of each tile
    // Face entry e is one 48-byte record (three 16-byte loads):
    //   fgeo[3e]   = {L, R} (two int32) and nx;  fgeo[3e+1] = {ny, dl};  fgeo[3e+2] = {dx, dy}
    // L = left cell (renumbered), R = right cell or -(face kind) on a boundary face, (nx, ny) the
    // unit normal from L, dl the length, (dx, dy) the centre offset x_R - x_L (ghost: reflected
    // centre).
    const double2 *fgeo;      // [3*nfe] (flux pass)
    // The gradient pass reads only what it needs, from compact arrays:
    const int2 *fLR;          // [nfe] {L, R}
    const double *fnx, *fny, *fdl;  // [nfe] unit normal and length (Green-Gauss, WLSQ ghosts)
    const double2 *fw;        // [2*nfe] least-squares coefficient vectors of L (2e) and R (2e+1)
    // WENO (order 3) only: per face entry the line stencil {L-1, L-2, R+1, R+2} (renumbered
    // cells through opposite faces of quadrilaterals), all -1 where the line is incomplete
    const int4 *fst;          // [nfe]
    const uint16_t *slots;    // [4*nc] (local face << 1) | (cell is R); 0xFFFF = none
    const double *area, *inv_area;  // [nc]
    // Batched meshes (several independent instances concatenated; null for one mesh):
    const int *tile_inst;     // [ntiles] instance of each tile
    const double4 *inst_fs;   // [ninst] prescribed state (rho, u, v, p) of each instance
    const int *tile_list;     // [ntiles] tiles of the instances still marching
    const int *tile_count;    // [1] entries of tile_list in use
};

// Per-launch scheme settings.
struct DevCfg {
    int bc[6];       // bc[kind] for kinds 1..5
    double fs[4];    // prescribed state (rho, u, v, p)
    double cfl;
    int roe_bdiss;
    int xcd_remap;   // tile order over the XCDs: 0 block order, 1 one contiguous run per XCD,
                     // k >= 2 runs of k consecutive tiles dealt to the XCDs in turn
    int grad_lds;    // stage the tile's gradients in LDS (1) or read them through L2 (0)
    // laminar viscous terms (second order only): 0 off, 1 constant viscosity mu, 2 Sutherland
    int viscous;
    double mu;       // dynamic viscosity (viscous == 1)
    double r_gas;    // gas constant: T = p / (rho r_gas)
    double kfac;     // conductivity / viscosity = cp / Pr = gamma r_gas / ((gamma - 1) Pr)
    double inv_pr;   // 1 / Pr
    double p_exit;   // static pressure of back-pressure exits (0: the prescribed state's p)
    int persist;     // tile passes: 0 one workgroup per tile, 1 persistent workgroups (launch only)
    int grid_tiles;  // > 0: launch grids cover this many tiles (the marching ones; launch only)
};

}  // namespace cfd
