// Flux pass of the native finite-volume path, shared by the translation units that
// instantiate it (fv_kernels.hip: inviscid first and second order; fv_visc.hip: second order
// with viscous terms; fv_weno.hip and fv_weno_visc.hip: WENO5, order 3, without and with viscous
// terms), plus the tile helpers the other kernels use.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fv_device.h"
#include "fv_kernels.h"

// Compile-time tuning of the flux pass (defaults chosen by measurement on MI355X):
//   CFD_FLUX_WAVES_PER_EU   waves per SIMD the register allocation must allow (0: compiler's choice)
//   CFD_FLUX_PREFETCH_CELL  1: load the gather phase's cell inputs at kernel entry, 0: in the gather
#ifndef CFD_FLUX_WAVES_PER_EU
#define CFD_FLUX_WAVES_PER_EU 0
#endif
#ifndef CFD_FLUX_PREFETCH_CELL
#define CFD_FLUX_PREFETCH_CELL 0
#endif
#if CFD_FLUX_WAVES_PER_EU > 0
#define CFD_FLUX_ATTR __attribute__((amdgpu_waves_per_eu(CFD_FLUX_WAVES_PER_EU)))
#else
#define CFD_FLUX_ATTR
#endif
//   CFD_POS_WAVES_PER_EU  minimum waves per SIMD of the positivity fallback's detecting pass (0: the
//                         compiler's choice)
#ifndef CFD_POS_WAVES_PER_EU
#define CFD_POS_WAVES_PER_EU 0
#endif
#if CFD_POS_WAVES_PER_EU > 0
#define CFD_POS_ATTR __attribute__((amdgpu_waves_per_eu(POS == 1 ? CFD_POS_WAVES_PER_EU : 1)))
#else
#define CFD_POS_ATTR
#endif

namespace cfd {

namespace {

__device__ __forceinline__ int tile_of_block(int b, int ntiles, int remap) {
    if (!remap) return b;
    // Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
    if (remap == 1) {
        // each XCD takes one contiguous run of tiles
        const int per = (ntiles + 7) / 8;
        return (b % 8) * per + b / 8;
    }
    // runs of `remap` consecutive tiles dealt to the XCDs in turn
    const int j = b / 8;
    return ((j / remap) * 8 + (b % 8)) * remap + (j % remap);
}

// Tiles to process: all, or for a batched mesh the list of tiles still marching.
__device__ __forceinline__ int tiles_to_run(const DevMesh &m) { return m.tile_list ? *m.tile_count : m.ntiles; }

// Virtual workgroups covering n tiles (a multiple of 8 with the XCD-aware order).
__host__ __device__ __forceinline__ int tile_span(int n, int remap) {
    if (remap <= 0) return n;
    if (remap == 1) return 8 * ((n + 7) / 8);
    return 8 * remap * (((n + remap - 1) / remap + 7) / 8);
}

// Tile of virtual workgroup v: its XCD-aware position in the tile sequence, or in the list of
// tiles still marching. -1 when v has no tile. The XCD-aware order is a bijection of [0, span)
// only: a batched launch keeps the grid of all tiles while fewer still march, and workgroups
// beyond the span must not map onto a tile again.
__device__ __forceinline__ int virtual_tile(const DevMesh &m, int v, int n, int remap) {
    if (v >= tile_span(n, remap)) return -1;
    const int b = tile_of_block(v, n, remap);
    if (b >= n) return -1;
    return m.tile_list ? m.tile_list[b] : b;
}

// Prescribed (freestream / inlet) state of a tile: the configuration's, or on a batched mesh the
// tile's instance's, then the exits' static pressure (the configuration's, else that state's p).
// Kept in a local array (the kernel parameters stay read-only).
__device__ __forceinline__ void tile_fs(const DevMesh &m, const DevCfg &cfg, int tile, double *fsv) {
    if (m.inst_fs) {
        const double4 f = m.inst_fs[m.tile_inst[tile]];
        fsv[0] = f.x;
        fsv[1] = f.y;
        fsv[2] = f.z;
        fsv[3] = f.w;
    } else {
        for (int q = 0; q < 4; q++) fsv[q] = cfg.fs[q];
    }
    fsv[4] = (cfg.p_exit > 0.0) ? cfg.p_exit : fsv[3];
}

__device__ __forceinline__ void load_state(const double4 *__restrict__ U, int c, FState &s) {
    double4 u = U[c];
    state_from_U(u.x, u.y, u.z, u.w, s);
}

// Block-wide sum of 4 values per thread into out[0..3] (fixed tree order). `red` holds
// 4 * blockDim.x doubles of LDS.
__device__ __forceinline__ void block_sum4(double *red, const double *v, double *out) {
    const int t = threadIdx.x, n = blockDim.x;
    for (int k = 0; k < 4; k++) red[k * n + t] = v[k];
    __syncthreads();
    for (int s = n / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 4; k++) red[k * n + t] += red[k * n + t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 4; k++) out[k] = red[k * n];
}

__device__ __forceinline__ double block_min(double *red, double v) {
    const int t = threadIdx.x, n = blockDim.x;
    red[t] = v;
    __syncthreads();
    for (int s = n / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = fmin(red[t], red[t + s]);
        __syncthreads();
    }
    return red[0];
}

// Tile of renumbered cell c (tiles are contiguous cell ranges; binary search of tile_cstart).
__device__ __forceinline__ int tile_of_cell(const DevMesh &m, int c) {
    int lo = 0, hi = m.ntiles;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (m.tile_cstart[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

// reference error term (src/Error_Estimate_Update.cpp:10-28)
__device__ __forceinline__ double err_term(double d, double u) {
    double t = (d < 1e-9) ? fabs(d) : (fabs(d) / u);
    return t * t;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Face records. Every load that does not depend on LDS is issued at kernel entry: the face
// records of a thread's first two faces and its cell's phase-C inputs are held in registers
// across phase A, so a workgroup pays one memory round trip for them instead of one per phase.
// ------------------------------------------------------------------------------------------
struct FaceRec {
    int L, R;
    double nx, ny, dl, dx, dy;
};

// Face entry e's record: three 16-byte loads ({L, R, nx}, {ny, dl}, {dx, dy}); the first is
// always loaded (it holds the cells), the others when asked for.
template <bool GEOM, bool CENTRE>
__device__ __forceinline__ void load_face(const DevMesh &m, int e, FaceRec &f) {
    const int4 a = reinterpret_cast<const int4 *>(m.fgeo)[3 * (size_t)e];
    f.L = a.x;
    f.R = a.y;
    f.nx = __hiloint2double(a.w, a.z);
    if constexpr (GEOM) {
        const double2 b = m.fgeo[3 * (size_t)e + 1];
        f.ny = b.x;
        f.dl = b.y;
    }
    if constexpr (CENTRE) {
        const double2 c = m.fgeo[3 * (size_t)e + 2];
        f.dx = c.x;
        f.dy = c.y;
    }
}

// The gradient pass's face entry: cells, and for Green-Gauss the normal and length.
template <bool GG>
__device__ __forceinline__ void load_face_grad(const DevMesh &m, int e, FaceRec &f) {
    const int2 lr = m.fLR[e];
    f.L = lr.x;
    f.R = lr.y;
    if constexpr (GG) {
        f.nx = m.fnx[e];
        f.ny = m.fny[e];
        f.dl = m.fdl[e];
    }
}

// ------------------------------------------------------------------------------------------
// Flux pass. MODE_RESID writes R and the local time step; MODE_STAGE applies one stage of
// TVD-RK3 / RK4 in place of writing R. One cell per thread (tile <= blockDim).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void load_grad(const double *__restrict__ grad, int c, double *g) {
    const double2 *gp = reinterpret_cast<const double2 *>(grad + 8 * (size_t)c);
    for (int q = 0; q < 4; q++) {
        double2 v = gp[q];
        g[2 * q] = v.x;
        g[2 * q + 1] = v.y;
    }
}

template <int FLUX, int ORDER, int LIM, bool VISC, int POS = 0>
__device__ __forceinline__ void flux_face(const DevMesh &m, const DevCfg &cfg,
                                          const double4 *__restrict__ Us,
                                          const double *__restrict__ grad, const double4 *Uexit,
                                          const double *sW, const double *sU, const double *sG,
                                          double *sF, bool glds, int T, int FT, int c0, int nct,
                                          int j, const FaceRec &f, int fz, float *psi, int e,
                                          const double *fsv, const uint8_t *pflag = nullptr,
                                          double *sNb = nullptr) {
    const double nx = f.nx, ny = f.ny, dl = f.dl;
    if constexpr (POS == 2) {
        // positivity fallback, recomputation: a face of a flagged cell takes the two cells'
        // first-order states (conserved values as stored, ghost of the left state)
        const bool bnd = f.R < 0;
        if (pflag[f.L] || (!bnd && pflag[f.R])) {
            FState SL, SR;
            load_state(Us, f.L, SL);
            if (bnd) ghost_state(cfg.bc[-f.R], SL, nx, ny, fsv, SR);
            else load_state(Us, f.R, SR);
            double F[4];
            face_flux<FLUX>(bnd, cfg.roe_bdiss, SL, SR, nx, ny, dl, F);
            for (int q = 0; q < 4; q++) sF[q * FT + j] = F[q];
            return;
        }
    }
    const int lL = f.L - c0;
    const bool locL = (unsigned)lL < (unsigned)nct;
    FState SL, SR;
    // Tile cells come from LDS: primitives always, conserved values at first order and for WENO
    // (the MUSCL face states form their conserved values from the reconstructed primitives).
    if (locL) {
        for (int q = 0; q < 5; q++) SL.W[q] = sW[q * T + lL];
        if constexpr (ORDER != 2)
            for (int q = 0; q < 4; q++) SL.U[q] = sU[q * T + lL];
    } else {
        load_state(Us, f.L, SL);
    }
    const bool boundary = f.R < 0;
    int bc = 0;
    bool locR = false;
    int lR = 0;
    if (boundary) {
        bc = cfg.bc[-f.R];
        // ghost of the cell-centre state (at second order only its primitives are used: the
        // L side's reconstruction differences toward it)
        if (bc == BC_EXIT && Uexit) {
            FState src;
            load_state(Uexit, f.L, src);
            ghost_state(bc, src, nx, ny, fsv, SR);
        } else {
            ghost_state(bc, SL, nx, ny, fsv, SR);
        }
    } else {
        lR = f.R - c0;
        locR = (unsigned)lR < (unsigned)nct;
        if (locR) {
            for (int q = 0; q < 5; q++) SR.W[q] = sW[q * T + lR];
            if constexpr (ORDER != 2)
                for (int q = 0; q < 4; q++) SR.U[q] = sU[q * T + lR];
        } else {
            load_state(Us, f.R, SR);
        }
    }
    if constexpr (POS == 1) {
        // positivity fallback, detection: density and pressure of the face's two cells (the
        // cell's own values on a boundary face, whose other side is a ghost) for the gather's
        // neighbourhood range, rounded to FP32 (the range's margin is 1e-3 relative)
        float *nb = reinterpret_cast<float *>(sNb);
        nb[j] = (float)SL.W[0];
        nb[FT + j] = (float)SL.W[3];
        nb[2 * FT + j] = (float)(boundary ? SL.W[0] : SR.W[0]);
        nb[3 * FT + j] = (float)(boundary ? SL.W[3] : SR.W[3]);
    }
    double F[4];
    if constexpr (ORDER == 3) {
        // WENO5 on the conserved variables along the line of quadrilaterals through the face,
        // cells L-2, L-1, L | R, R+1, R+2, where the layout found one; MUSCL below otherwise
        const int4 q4 = boundary ? make_int4(-1, -1, -1, -1) : m.fst[e];
        if (q4.x >= 0) {
            auto cons = [&](int c, double *u) {
                const int l = c - c0;
                if ((unsigned)l < (unsigned)nct) {
                    for (int q = 0; q < 4; q++) u[q] = sU[q * T + l];
                } else {
                    const double4 v = Us[c];
                    u[0] = v.x;
                    u[1] = v.y;
                    u[2] = v.z;
                    u[3] = v.w;
                }
            };
            double Lm2[4], Lm1[4], Rp1[4], Rp2[4];
            cons(q4.y, Lm2);
            cons(q4.x, Lm1);
            cons(q4.z, Rp1);
            cons(q4.w, Rp2);
            double UL[4], UR[4];
            for (int q = 0; q < 4; q++) {
                UL[q] = weno5(Lm2[q], Lm1[q], SL.U[q], SR.U[q], Rp1[q], 0);
                UR[q] = weno5(Lm1[q], SL.U[q], SR.U[q], Rp1[q], Rp2[q], 1);
            }
            FState fL, fR;
            weno_face_state(UL, SL, fL);
            weno_face_state(UR, SR, fR);
            face_flux<FLUX>(false, cfg.roe_bdiss, fL, fR, nx, ny, dl, F);
            if constexpr (VISC) {
                double gL[8], gR[8];
                if (glds && locL) {
                    for (int q = 0; q < 8; q++) gL[q] = sG[q * T + lL];
                } else {
                    load_grad(grad, f.L, gL);
                }
                if (glds && locR) {
                    for (int q = 0; q < 8; q++) gR[q] = sG[q * T + lR];
                } else {
                    load_grad(grad, f.R, gR);
                }
                double Fv[4];
                viscous_flux(cfg, SL.W, SR.W, gL, gR, nx, ny, dl, f.dx, f.dy, Fv);
                for (int q = 0; q < 4; q++) F[q] = F[q] - Fv[q];
            }
            for (int q = 0; q < 4; q++) sF[q * FT + j] = F[q];
            return;
        }
    }
    if constexpr (ORDER >= 2) {
        double gL[8];
        if (glds && locL) {
            for (int q = 0; q < 8; q++) gL[q] = sG[q * T + lL];
        } else {
            load_grad(grad, f.L, gL);
        }
        double gR[8];
        FState fL, fR;
        // frozen limiter factors of this face entry: side L in pv[0..3], side R in pv[4..7]
        // (8 contiguous floats per entry, two 16-byte loads)
        float pv[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float4 *p4 = fz ? reinterpret_cast<float4 *>(psi) + 2 * (size_t)e : nullptr;
        if (fz == 2) {
            const float4 a = p4[0], b = p4[1];
            pv[0] = a.x; pv[1] = a.y; pv[2] = a.z; pv[3] = a.w;
            pv[4] = b.x; pv[5] = b.y; pv[6] = b.z; pv[7] = b.w;
        }
        reconstruct_side<LIM>(SL, SR, gL, f.dx, f.dy, 1.0, fz, pv, fL);
        if (boundary) {
            if (bc == BC_EXIT && Uexit) {
                FState src;
                load_state(Uexit, f.L, src);
                ghost_state(bc, src, nx, ny, fsv, fR);
            } else {
                ghost_state(bc, fL, nx, ny, fsv, fR);
            }
            if constexpr (VISC)
                for (int q = 0; q < 8; q++) gR[q] = gL[q];
        } else {
            if (glds && locR) {
                for (int q = 0; q < 8; q++) gR[q] = sG[q * T + lR];
            } else {
                load_grad(grad, f.R, gR);
            }
            reconstruct_side<LIM>(SR, SL, gR, f.dx, f.dy, -1.0, fz, pv + 4, fR);
        }
        if (fz == 1) {
            p4[0] = make_float4(pv[0], pv[1], pv[2], pv[3]);
            p4[1] = make_float4(pv[4], pv[5], pv[6], pv[7]);
        }
        face_flux<FLUX>(boundary, cfg.roe_bdiss, fL, fR, nx, ny, dl, F);
        if constexpr (VISC) {
            // viscous flux from the cell-centre states (ghost on a boundary face), subtracted
            double Fv[4];
            viscous_flux(cfg, SL.W, SR.W, gL, gR, nx, ny, dl, f.dx, f.dy, Fv);
            for (int q = 0; q < 4; q++) F[q] = F[q] - Fv[q];
        }
    } else {
        face_flux<FLUX>(boundary, cfg.roe_bdiss, SL, SR, nx, ny, dl, F);
    }
    for (int q = 0; q < 4; q++) sF[q * FT + j] = F[q];
}

// FZ: the stage reads st.freeze and can record or use frozen limiter factors (launched only when a
// freezing state is set; without it the kernel holds none of that code). Exit ghosts from
// st.Uexit (the reference iteration's stages) exist in residual mode only.
//
// PERSIST: the launch has fewer workgroups than tiles (what the GPU holds at once) and each
// workgroup walks the virtual workgroup indices blockIdx.x, blockIdx.x + gridDim.x, ... up to the
// tiles still marching, so a batch whose instances have mostly converged launches no empty
// workgroups. With a grid that is a multiple of 8 a workgroup stays on one XCD and keeps that
// XCD's run of tiles. Without it the loop runs once (one tile per workgroup).
template <int FLUX, int ORDER, int LIM, int MODE, bool VISC, int POS = 0, bool FZ = false, bool PERSIST = false>
__global__ void __launch_bounds__(256) CFD_FLUX_ATTR CFD_POS_ATTR flux_tile_kernel(DevMesh m, DevCfg cfg,
                                                        const double4 *__restrict__ Us,
                                                        const double *__restrict__ grad,
                                                        StageArgs st) {
    extern __shared__ double smem[];
    // the recomputing pass of the positivity fallback walks the list of marked tiles
    constexpr bool DLIST = (POS == 2 && PERSIST);
    const int par = (st.pos_range >> 1) & 3;
    const int ntr = DLIST ? st.dirty[2 * m.ntiles + par] : tiles_to_run(m);
    const int span = DLIST ? ntr : (PERSIST ? tile_span(ntr, cfg.xcd_remap) : 0);
    int v = blockIdx.x;
    if (DLIST && v >= span) return;
    do {
    const int tile = DLIST ? st.dirty[m.ntiles + v] : virtual_tile(m, v, ntr, cfg.xcd_remap);
    if (tile < 0) continue;
    if (PERSIST && v != (int)blockIdx.x) __syncthreads();  // the previous tile's LDS reads are done
    if constexpr (POS == 2) {
        if (!st.dirty[tile]) continue;  // no face of this tile touches a flagged cell
    }
    double fsv[5];
    tile_fs(m, cfg, tile, fsv);
    const int T = m.tile;
    const int FT = m.max_tile_faces;
    const int B = blockDim.x;
    const int c0 = m.tile_cstart[tile];
    const int nct = m.tile_cstart[tile + 1] - c0;
    const int fs = m.tile_fstart[tile], nft = m.tile_fstart[tile + 1] - fs;
    const bool glds = (ORDER >= 2) && cfg.grad_lds;
    const int fz = (FZ && ORDER == 2 && st.freeze) ? st.freeze[m.tile_inst ? m.tile_inst[tile] : 0] : 0;
    const double4 *Uexit = (MODE == MODE_RESID) ? st.Uexit : nullptr;
    double *sW = smem;                          // [5][T] primitives (rho, u, v, p, a)
    double *sU = sW + 5 * T;                    // [4][T] conserved (first order and WENO)
    double *sF = sU + (ORDER != 2 ? 4 * T : 0); // [4][FT] face fluxes
    double *sG = sF + 4 * FT;                   // [8][T] (MUSCL, LDS-staged gradients)
    double *sN = sG + (glds ? 8 * T : 0);       // [3][FT] face normal and length (time steps)
    // phase C forms time steps from the face geometry: staged here from the face records of
    // phase B instead of read again from HBM (the launch sizes LDS for it only where it is used)
    const bool needN = (MODE == MODE_RESID) || (st.lts && st.stage == 0) || (st.last && st.dt_partial);
    double *sNb = needN ? sN + 3 * FT : sN;     // float [4][FT] rho, p of both cells (positivity detection)
    auto stageN = [&](int j, const FaceRec &f) {
        if (needN) {
            sN[j] = f.nx;
            sN[FT + j] = f.ny;
            sN[2 * FT + j] = f.dl;
        }
    };
    const int k = threadIdx.x;
    const bool hasCell = k < nct;
    const int c = c0 + k;

    // ---- prefetch -----------------------------------------------------------------------
    double4 u4 = make_double4(1.0, 0.0, 0.0, 1.0), un4 = make_double4(0, 0, 0, 0);
    ushort4 sl4 = make_ushort4(0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF);
    double ar = 0.0, ia = 0.0, hprev = 0.0;
    // the gather phase's cell inputs: loaded here (in flight during phases A and B) or there
    auto load_cell_inputs = [&]() {
        sl4 = reinterpret_cast<const ushort4 *>(m.slots)[c];
        ar = m.area[c];
        if constexpr (MODE == MODE_STAGE) {
            un4 = st.Un[c];
            ia = m.inv_area[c];
            if (st.lts && st.stage > 0) hprev = st.hbuf[c];
        }
    };
    if (hasCell) {
        u4 = Us[c];
        if (CFD_FLUX_PREFETCH_CELL) load_cell_inputs();
    }
    FaceRec f0{}, f1{};
    const bool h0 = k < nft, h1 = k + B < nft;
    if (h0) load_face<true, ORDER >= 2>(m, fs + k, f0);
    if (h1) load_face<true, ORDER >= 2>(m, fs + k + B, f1);

    // ---- A: stage the tile's states ----------------------------------------------------
    if (hasCell) {
        FState s;
        state_from_U(u4.x, u4.y, u4.z, u4.w, s);
        for (int q = 0; q < 5; q++) sW[q * T + k] = s.W[q];
        if constexpr (ORDER != 2)
            for (int q = 0; q < 4; q++) sU[q * T + k] = s.U[q];
        if (glds) {
            double g[8];
            load_grad(grad, c, g);
            for (int q = 0; q < 8; q++) sG[q * T + k] = g[q];
        }
    }
    __syncthreads();

    // ---- B: one flux per face ----------------------------------------------------------
    if (h0) {
        stageN(k, f0);
        flux_face<FLUX, ORDER, LIM, VISC, POS>(m, cfg, Us, grad, Uexit, sW, sU, sG, sF, glds, T, FT, c0, nct, k, f0, fz, st.psi, fs + k, fsv, st.pflag, sNb);
    }
    if (h1) {
        stageN(k + B, f1);
        flux_face<FLUX, ORDER, LIM, VISC, POS>(m, cfg, Us, grad, Uexit, sW, sU, sG, sF, glds, T, FT, c0, nct, k + B, f1, fz, st.psi, fs + k + B, fsv, st.pflag, sNb);
    }
    for (int j = k + 2 * B; j < nft; j += B) {
        FaceRec f;
        load_face<true, ORDER >= 2>(m, fs + j, f);
        stageN(j, f);
        flux_face<FLUX, ORDER, LIM, VISC, POS>(m, cfg, Us, grad, Uexit, sW, sU, sG, sF, glds, T, FT, c0, nct, j, f, fz, st.psi, fs + j, fsv, st.pflag, sNb);
    }
    __syncthreads();

    // ---- C: gather per cell, time step, stage update -----------------------------------
    double errp[4] = {0.0, 0.0, 0.0, 0.0};
    double dtmin = 1.7976931348623157e308;
    if (hasCell) {
        if (!CFD_FLUX_PREFETCH_CELL) load_cell_inputs();
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        const uint16_t sls[4] = {sl4.x, sl4.y, sl4.z, sl4.w};
        for (int f = 0; f < 4; f++) {
            const uint16_t sl = sls[f];
            if (sl == 0xFFFF) break;
            const int j = sl >> 1;
            const bool isR = sl & 1;
            for (int q = 0; q < 4; q++) {
                double v = sF[q * FT + j];
                r[q] += isR ? -v : v;
            }
        }
        // spectral radius sum_f (|u.n_f| + a) dl_f of the cell's own state, in its face order,
        // only where a time step is formed (residual mode, first stage of local time stepping)
        double denom = 0.0;
        if (MODE == MODE_RESID || (st.lts && st.stage == 0)) {
            const double w1 = sW[1 * T + k], w2 = sW[2 * T + k], w4 = sW[4 * T + k];
            double sdl2 = 0.0;
            for (int f = 0; f < 4; f++) {
                const uint16_t sl = sls[f];
                if (sl == 0xFFFF) break;
                const int j = sl >> 1;
                denom += (fabs(w1 * sN[j] + w2 * sN[FT + j]) + w4) * sN[2 * FT + j];
                if constexpr (VISC) sdl2 += sN[2 * FT + j] * sN[2 * FT + j];
            }
            if constexpr (VISC) denom += viscous_dt_term(cfg, sW[0 * T + k], sW[3 * T + k], sdl2, ar);
        }
        // Inviscid_Time_Step (src/Time_Step.cpp:293-329), plus the viscous term
        const double dtc = (denom <= 0.0 || !isfinite(denom)) ? 0.0 : cfg.cfl * ar / denom;
        if constexpr (MODE == MODE_RESID) {
            reinterpret_cast<double4 *>(st.R_out)[c] = make_double4(r[0], r[1], r[2], r[3]);
            if (st.dt_out) st.dt_out[c] = dtc;
            dtmin = fmin(dtmin, dtc);
        } else {
            double h;
            if (st.lts) {
                if (st.stage == 0) {
                    h = dtc * ia;
                    st.hbuf[c] = h;
                } else {
                    h = hprev;
                }
            } else {
                h = (*st.dt_global) * ia;
            }
            const double Un[4] = {un4.x, un4.y, un4.z, un4.w};
            double Uo[4];
            if (st.rk == 4) {
                double4 a4 = (st.stage == 0) ? make_double4(0, 0, 0, 0) : st.acc_in[c];
                double acc[4] = {a4.x, a4.y, a4.z, a4.w};
                for (int q = 0; q < 4; q++) {
                    switch (st.stage) {
                    case 0: acc[q] = r[q]; Uo[q] = Un[q] - 0.5 * h * r[q]; break;
                    case 1: acc[q] += 2.0 * r[q]; Uo[q] = Un[q] - 0.5 * h * r[q]; break;
                    case 2: acc[q] += 2.0 * r[q]; Uo[q] = Un[q] - h * r[q]; break;
                    default: Uo[q] = Un[q] - (h / 6.0) * (acc[q] + r[q]); break;
                    }
                }
                if (st.stage < 3) st.acc_out[c] = make_double4(acc[0], acc[1], acc[2], acc[3]);
            } else {
                const double us[4] = {u4.x, u4.y, u4.z, u4.w};
                for (int q = 0; q < 4; q++) {
                    switch (st.stage) {
                    case 0: Uo[q] = Un[q] - h * r[q]; break;
                    case 1: Uo[q] = 0.75 * Un[q] + 0.25 * (us[q] - h * r[q]); break;
                    default: Uo[q] = (1.0 / 3.0) * Un[q] + (2.0 / 3.0) * (us[q] - h * r[q]); break;
                    }
                }
            }
            if constexpr (POS == 1) {
                // positivity fallback, detection: the stage state must have a positive density and
                // pressure within the relaxed range of the cell's and its face neighbours' stage
                // inputs (the same test as the oracle's stage_fallback_flags)
                double rn, pn;
                rho_p_of(Uo, rn, pn);
                const float *nb = reinterpret_cast<const float *>(sNb);
                double rlo = (double)(float)sW[k], rhi = rlo, plo = (double)(float)sW[3 * T + k], phi = plo;
                for (int f = 0; f < 4; f++) {
                    const uint16_t sl = sls[f];
                    if (sl == 0xFFFF) break;
                    const int j = sl >> 1;
                    const bool isR = sl & 1;
                    const double rj = (double)(isR ? nb[j] : nb[2 * FT + j]);
                    const double pj = (double)(isR ? nb[FT + j] : nb[3 * FT + j]);
                    rlo = fmin(rlo, rj);
                    rhi = fmax(rhi, rj);
                    plo = fmin(plo, pj);
                    phi = fmax(phi, pj);
                }
                const double dr = fmax(1e-3 * fabs(rhi), 1e-3 * (rhi - rlo));
                const double dp = fmax(1e-3 * fabs(phi), 1e-3 * (phi - plo));
                const bool bad = !(rn > 0.0) || !(pn > 0.0) ||
                                 ((st.pos_range & 1) && (!(rn >= rlo - dr && rn <= rhi + dr) ||
                                                         !(pn >= plo - dp && pn <= phi + dp)));
                st.pflag[c] = bad ? 1 : 0;
                if (bad) {
                    // this tile and the tiles of the cell's face neighbours recompute the stage
                    st.dirty[tile] = 1;
                    for (int f = 0; f < 4; f++) {
                        const uint16_t sl = sls[f];
                        if (sl == 0xFFFF) break;
                        const int4 a = reinterpret_cast<const int4 *>(m.fgeo)[3 * (size_t)(fs + (sl >> 1))];
                        const int o = (sl & 1) ? a.x : a.y;
                        if (o >= 0 && (unsigned)(o - c0) >= (unsigned)nct) st.dirty[tile_of_cell(m, o)] = 1;
                    }
                }
            }
            st.Uout[c] = make_double4(Uo[0], Uo[1], Uo[2], Uo[3]);
            if (st.last) {
                if (st.err_partial) {
                    for (int q = 0; q < 4; q++) errp[q] += err_term(Uo[q] - Un[q], Un[q]);
                }
                if (st.dt_partial) {
                    // time step of the new state, for the next step's global dt
                    FState s;
                    state_from_U(Uo[0], Uo[1], Uo[2], Uo[3], s);
                    double den = 0.0, sdl2 = 0.0;
                    for (int f = 0; f < 4; f++) {
                        const uint16_t sl = sls[f];
                        if (sl == 0xFFFF) break;
                        const int j = sl >> 1;
                        den += (fabs(s.W[1] * sN[j] + s.W[2] * sN[FT + j]) + s.W[4]) * sN[2 * FT + j];
                        if constexpr (VISC) sdl2 += sN[2 * FT + j] * sN[2 * FT + j];
                    }
                    if constexpr (VISC) den += viscous_dt_term(cfg, s.W[0], s.W[3], sdl2, ar);
                    const double dtn = (den <= 0.0 || !isfinite(den)) ? 0.0 : cfg.cfl * ar / den;
                    dtmin = fmin(dtmin, dtn);
                }
            }
        }
    }
    // block reductions reuse LDS from its base (the allocation holds >= 4 * blockDim doubles)
    if constexpr (MODE == MODE_RESID) {
        if (st.dt_partial) {
            __syncthreads();
            double b = block_min(smem, dtmin);
            if (threadIdx.x == 0) st.dt_partial[tile] = b;
        }
    } else {
        if (st.last && st.err_partial) {
            __syncthreads();
            double out[4];
            block_sum4(smem, errp, out);
            if (threadIdx.x == 0)
                for (int q = 0; q < 4; q++) st.err_partial[4 * tile + q] = out[q];
            __syncthreads();
        }
        if (st.last && st.dt_partial) {
            __syncthreads();
            double b = block_min(smem, dtmin);
            if (threadIdx.x == 0) st.dt_partial[tile] = b;
        }
    }
    if constexpr (POS == 2) {
        if (threadIdx.x == 0) st.dirty[tile] = 0;  // read by every thread before the first barrier
    }
    } while (PERSIST && (v += gridDim.x) < span);
}

static inline int grid_of(const DevMesh &m, const DevCfg &c) {
    const int n = (c.grid_tiles > 0 && c.grid_tiles < m.ntiles) ? c.grid_tiles : m.ntiles;
    return tile_span(n, c.xcd_remap);
}

// Grid of a persistent launch: the workgroups the GPU holds at once for this kernel (occupancy
// at this block size and LDS, times the compute units), a multiple of 8, at most one per tile.
int persist_grid(const void *kernel, const DevMesh &m, const DevCfg &c, int block, size_t sm);
template <typename K>
static inline int grid_persist(K kernel, const DevMesh &m, const DevCfg &c, int block, size_t sm) {
    return persist_grid(reinterpret_cast<const void *>(kernel), m, c, block, sm);
}

// Persistent stage passes (fv_persist.hip): order 1 or 2, inviscid, with or without limiter
// freezing.
hipError_t launch_flux_persist(const DevMesh &m, const DevCfg &c, int flux, int order, int lim,
                               const double4 *Us, const double *grad, const StageArgs &st, int block,
                               size_t sm, hipStream_t s);

// Order-2 stages with limiter freezing (fv_freeze.hip): the FZ variants of the stage kernel.
hipError_t launch_flux_frozen(const DevMesh &m, const DevCfg &c, int flux, int lim, bool visc,
                              const double4 *Us, const double *grad, const StageArgs &st, int block,
                              size_t sm, hipStream_t s);

template <int FLUX, int ORDER, int LIM, bool VISC>
static void launch_flux_t(const DevMesh &m, const DevCfg &c, int mode, const double4 *Us,
                          const double *grad, const StageArgs &st, int block, size_t sm,
                          hipStream_t s) {
    if (mode == MODE_RESID)
        hipLaunchKernelGGL((flux_tile_kernel<FLUX, ORDER, LIM, MODE_RESID, VISC>), dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad, st);
    else if (ORDER == 2 && st.freeze)
        (void)launch_flux_frozen(m, c, FLUX, LIM, VISC, Us, grad, st, block, sm, s);
    else
        hipLaunchKernelGGL((flux_tile_kernel<FLUX, ORDER, LIM, MODE_STAGE, VISC>), dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad, st);
}

// Every flux and limiter of one (order, viscous) family of the flux pass.
template <int ORDER, bool VISC>
static hipError_t launch_flux_family(const DevMesh &m, const DevCfg &c, int flux, int lim, int mode,
                                     const double4 *Us, const double *grad, const StageArgs &st,
                                     int block, size_t sm, hipStream_t s) {
    auto by_lim = [&](auto fluxc) {
        constexpr int F = decltype(fluxc)::value;
        if constexpr (ORDER == 1) {
            launch_flux_t<F, 1, LIM_NONE, VISC>(m, c, mode, Us, grad, st, block, sm, s);
        } else {
            switch (lim) {
            case LIM_NONE: launch_flux_t<F, ORDER, LIM_NONE, VISC>(m, c, mode, Us, grad, st, block, sm, s); break;
            case LIM_MINMOD: launch_flux_t<F, ORDER, LIM_MINMOD, VISC>(m, c, mode, Us, grad, st, block, sm, s); break;
            case LIM_VANLEER: launch_flux_t<F, ORDER, LIM_VANLEER, VISC>(m, c, mode, Us, grad, st, block, sm, s); break;
            case LIM_VANALBADA: launch_flux_t<F, ORDER, LIM_VANALBADA, VISC>(m, c, mode, Us, grad, st, block, sm, s); break;
            default: launch_flux_t<F, ORDER, LIM_SUPERBEE, VISC>(m, c, mode, Us, grad, st, block, sm, s); break;
            }
        }
    };
    switch (flux) {
    case FLUX_ROE: by_lim(std::integral_constant<int, FLUX_ROE>{}); break;
    case FLUX_HLLC: by_lim(std::integral_constant<int, FLUX_HLLC>{}); break;
    case FLUX_LLF: by_lim(std::integral_constant<int, FLUX_LLF>{}); break;
    default: by_lim(std::integral_constant<int, FLUX_AUSM>{}); break;
    }
    return hipGetLastError();
}

}  // namespace cfd
