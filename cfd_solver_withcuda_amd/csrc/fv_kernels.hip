// Native HIP kernels of the explicit finite-volume path (gfx950, wave64, FP64).
//
// One workgroup owns one tile: a contiguous range of renumbered cells plus the list of faces
// touching them. Each pass
//   A) stages the tile's cell states in LDS (coalesced loads of the contiguous range),
//   B) evaluates every face of the tile once (neighbour cells outside the tile come from
//      global memory, i.e. L2) and writes the face result into LDS,
//   C) gathers, per cell and in the cell's own face order, the face results with their sign.
// The gather order is fixed, so results are deterministic and need no atomics; faces shared by
// two tiles are evaluated by both with identical inputs, so the scheme stays exactly
// conservative. Phase C of the flux pass also applies the Runge-Kutta stage update, so a
// stage is one gradient pass plus one flux pass.
#include <hip/hip_runtime.h>

#include "fv_device.h"
#include "fv_kernels.h"

// Compile-time tuning of the flux pass (defaults chosen by measurement on MI355X):
//   CFD_FLUX_WAVES_PER_EU   waves per SIMD the register allocation must allow (0: compiler's choice)
//   CFD_FLUX_PREFETCH_CELL  1: load the gather phase's cell inputs at kernel entry, 0: in the gather
#ifndef CFD_FLUX_WAVES_PER_EU
#define CFD_FLUX_WAVES_PER_EU 0
#endif
#ifndef CFD_FLUX_PREFETCH_CELL
#define CFD_FLUX_PREFETCH_CELL 1
#endif
#if CFD_FLUX_WAVES_PER_EU > 0
#define CFD_FLUX_ATTR __attribute__((amdgpu_waves_per_eu(CFD_FLUX_WAVES_PER_EU)))
#else
#define CFD_FLUX_ATTR
#endif

namespace cfd {

namespace {

__device__ __forceinline__ int tile_of_block(int b, int ntiles, int remap) {
    if (!remap) return b;
    // Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8); give each
    // XCD a contiguous run of tiles so that faces shared by neighbouring tiles hit its L2.
    const int per = (ntiles + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Tile of this workgroup: its XCD-aware position in the tile sequence, or in the list of tiles
// still marching for a batched mesh. -1 when the workgroup has no tile.
__device__ __forceinline__ int block_tile(const DevMesh &m, int remap) {
    const int n = m.tile_list ? *m.tile_count : m.ntiles;
    const int b = tile_of_block(blockIdx.x, n, remap);
    if (b >= n) return -1;
    return m.tile_list ? m.tile_list[b] : b;
}

// Prescribed (freestream / inlet) state of a tile: the configuration's, or on a batched mesh the
// tile's instance's. Kept in a local array (the kernel parameters stay read-only).
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
// Gradient pass: Green-Gauss (reference semantics, ghost faces use the cell value) or weighted
// least squares (ghost neighbours at the reflected centre with the boundary ghost state).
// grad[8c..8c+7] = (d/dx, d/dy) of (rho, u, v, p). One cell per thread (tile <= blockDim).
// ------------------------------------------------------------------------------------------
template <int GRAD>
__device__ __forceinline__ void grad_face(const DevMesh &m, const DevCfg &cfg,
                                          const double4 *__restrict__ Us, const double *sW,
                                          double *sX, int T, int FT, int c0, int nct, int e,
                                          int j, const FaceRec &f, const double *fsv) {
    const int lL = f.L - c0;
    double WL[5], WR[4];
    if ((unsigned)lL < (unsigned)nct) {
        for (int q = 0; q < 5; q++) WL[q] = sW[q * T + lL];
    } else {
        FState s;
        load_state(Us, f.L, s);
        for (int q = 0; q < 5; q++) WL[q] = s.W[q];
    }
    if (f.R < 0) {
        if constexpr (GRAD == GRAD_GG) {
            for (int q = 0; q < 4; q++) WR[q] = WL[q];
        } else {
            FState in, g;
            for (int q = 0; q < 5; q++) in.W[q] = WL[q];
            for (int q = 0; q < 4; q++) in.U[q] = 0.0;
            ghost_state(cfg.bc[-f.R], in, m.fnx[e], m.fny[e], fsv, g);
            for (int q = 0; q < 4; q++) WR[q] = g.W[q];
        }
    } else {
        const int lR = f.R - c0;
        if ((unsigned)lR < (unsigned)nct) {
            for (int q = 0; q < 4; q++) WR[q] = sW[q * T + lR];
        } else {
            FState s;
            load_state(Us, f.R, s);
            for (int q = 0; q < 4; q++) WR[q] = s.W[q];
        }
    }
    // per face: Green-Gauss, the face value times the scaled normal (8 values); WLSQ, the
    // difference of each variable (4 values; the gather phase reads each cell face's
    // least-squares vector itself)
    if constexpr (GRAD == GRAD_GG) {
        for (int q = 0; q < 4; q++) {
            const double phi_face = 0.5 * (WL[q] + WR[q]);
            sX[(2 * q) * FT + j] = phi_face * f.nx * f.dl;
            sX[(2 * q + 1) * FT + j] = phi_face * f.ny * f.dl;
        }
    } else {
        for (int q = 0; q < 4; q++) sX[q * FT + j] = WR[q] - WL[q];
    }
}

template <int GRAD>
__global__ void __launch_bounds__(256) grad_tile_kernel(DevMesh m, DevCfg cfg,
                                                        const double4 *__restrict__ Us,
                                                        double *__restrict__ grad) {
    extern __shared__ double smem[];
    const int tile = block_tile(m, cfg.xcd_remap);
    if (tile < 0) return;
    double fsv[4];
    tile_fs(m, cfg, tile, fsv);
    const int T = m.tile;
    const int FT = m.max_tile_faces;
    const int B = blockDim.x;
    const int c0 = m.tile_cstart[tile];
    const int nct = m.tile_cstart[tile + 1] - c0;
    const int fs = m.tile_fstart[tile], nft = m.tile_fstart[tile + 1] - fs;
    double *sW = smem;              // [5][T]
    double *sX = smem + 5 * T;      // [8][FT] Green-Gauss, [4][FT] WLSQ
    constexpr bool GG = (GRAD == GRAD_GG);
    const int k = threadIdx.x;
    const bool hasCell = k < nct;
    const int c = c0 + k;

    // ---- prefetch -----------------------------------------------------------------------
    double4 u4 = make_double4(1.0, 0.0, 0.0, 1.0);
    ushort4 sl4 = make_ushort4(0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF);
    double ia = 0.0;
    if (hasCell) {
        u4 = Us[c];
        sl4 = reinterpret_cast<const ushort4 *>(m.slots)[c];
        if constexpr (GG) ia = m.inv_area[c];
    }
    FaceRec f0{}, f1{};
    const bool h0 = k < nft, h1 = k + B < nft;
    if (h0) load_face_grad<GG>(m, fs + k, f0);
    if (h1) load_face_grad<GG>(m, fs + k + B, f1);
    // ---- A ------------------------------------------------------------------------------
    if (hasCell) {
        FState s;
        state_from_U(u4.x, u4.y, u4.z, u4.w, s);
        for (int q = 0; q < 5; q++) sW[q * T + k] = s.W[q];
    }
    __syncthreads();
    // ---- B ------------------------------------------------------------------------------
    if (h0) grad_face<GRAD>(m, cfg, Us, sW, sX, T, FT, c0, nct, fs + k, k, f0, fsv);
    if (h1) grad_face<GRAD>(m, cfg, Us, sW, sX, T, FT, c0, nct, fs + k + B, k + B, f1, fsv);
    for (int j = k + 2 * B; j < nft; j += B) {
        FaceRec f;
        load_face_grad<GG>(m, fs + j, f);
        grad_face<GRAD>(m, cfg, Us, sW, sX, T, FT, c0, nct, fs + j, j, f, fsv);
    }
    __syncthreads();
    // ---- C ------------------------------------------------------------------------------
    if (hasCell) {
        double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint16_t sls[4] = {sl4.x, sl4.y, sl4.z, sl4.w};
        // WLSQ: the least-squares vector of each of the cell's faces, all loads issued before
        // the sums
        double ga[4] = {0, 0, 0, 0}, gb[4] = {0, 0, 0, 0};
        if constexpr (!GG) {
            for (int f = 0; f < 4; f++) {
                const uint16_t sl = sls[f];
                if (sl == 0xFFFF) break;
                const double2 w = m.fw[2 * (size_t)(fs + (sl >> 1)) + (sl & 1)];
                ga[f] = w.x;
                gb[f] = w.y;
            }
        }
        for (int f = 0; f < 4; f++) {
            const uint16_t sl = sls[f];
            if (sl == 0xFFFF) break;
            const int j = sl >> 1;
            const bool isR = sl & 1;
            if constexpr (GG) {
                for (int q = 0; q < 8; q++) {
                    double v = sX[q * FT + j];
                    g[q] += isR ? -v : v;
                }
            } else {
                for (int q = 0; q < 4; q++) {
                    double d = sX[q * FT + j];
                    double dW = isR ? -d : d;
                    g[2 * q] += ga[f] * dW;
                    g[2 * q + 1] += gb[f] * dW;
                }
            }
        }
        if constexpr (GG)
            for (int q = 0; q < 8; q++) g[q] = g[q] * ia;
        double2 *out = reinterpret_cast<double2 *>(grad + 8 * (size_t)c);
        out[0] = make_double2(g[0], g[1]);
        out[1] = make_double2(g[2], g[3]);
        out[2] = make_double2(g[4], g[5]);
        out[3] = make_double2(g[6], g[7]);
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

template <int FLUX, int ORDER, int LIM, bool VISC>
__device__ __forceinline__ void flux_face(const DevMesh &m, const DevCfg &cfg,
                                          const double4 *__restrict__ Us,
                                          const double *__restrict__ grad, const double4 *Uexit,
                                          const double *sW, const double *sU, const double *sG,
                                          double *sF, bool glds, int T, int FT, int c0, int nct,
                                          int j, const FaceRec &f, int fz, float *psi, int e,
                                          const double *fsv) {
    const double nx = f.nx, ny = f.ny, dl = f.dl;
    const int lL = f.L - c0;
    const bool locL = (unsigned)lL < (unsigned)nct;
    FState SL, SR;
    // Tile cells come from LDS: primitives always, conserved values at first order only (the
    // second-order face states form their conserved values from the reconstructed primitives).
    if (locL) {
        for (int q = 0; q < 5; q++) SL.W[q] = sW[q * T + lL];
        if constexpr (ORDER == 1)
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
            if constexpr (ORDER == 1)
                for (int q = 0; q < 4; q++) SR.U[q] = sU[q * T + lR];
        } else {
            load_state(Us, f.R, SR);
        }
    }
    double F[4];
    if constexpr (ORDER == 2) {
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

template <int FLUX, int ORDER, int LIM, int MODE, bool VISC>
__global__ void __launch_bounds__(256) CFD_FLUX_ATTR flux_tile_kernel(DevMesh m, DevCfg cfg,
                                                        const double4 *__restrict__ Us,
                                                        const double *__restrict__ grad,
                                                        StageArgs st) {
    extern __shared__ double smem[];
    const int tile = block_tile(m, cfg.xcd_remap);
    if (tile < 0) return;
    double fsv[4];
    tile_fs(m, cfg, tile, fsv);
    const int T = m.tile;
    const int FT = m.max_tile_faces;
    const int B = blockDim.x;
    const int c0 = m.tile_cstart[tile];
    const int nct = m.tile_cstart[tile + 1] - c0;
    const int fs = m.tile_fstart[tile], nft = m.tile_fstart[tile + 1] - fs;
    const bool glds = (ORDER == 2) && cfg.grad_lds;
    const int fz = (ORDER == 2 && st.freeze) ? st.freeze[m.tile_inst ? m.tile_inst[tile] : 0] : 0;
    double *sW = smem;                          // [5][T] primitives (rho, u, v, p, a)
    double *sU = sW + 5 * T;                    // [4][T] conserved (first order only)
    double *sF = sU + (ORDER == 1 ? 4 * T : 0); // [4][FT] face fluxes
    double *sG = sF + 4 * FT;                   // [8][T] (second order, LDS-staged gradients)
    double *sN = sG + (glds ? 8 * T : 0);       // [3][FT] face normal and length (time steps)
    // phase C forms time steps from the face geometry: staged here from the face records of
    // phase B instead of read again from HBM
    const bool needN = (MODE == MODE_RESID) || (st.lts && st.stage == 0) || (st.last && st.dt_partial);
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
    if (h0) load_face<true, ORDER == 2>(m, fs + k, f0);
    if (h1) load_face<true, ORDER == 2>(m, fs + k + B, f1);

    // ---- A: stage the tile's states ----------------------------------------------------
    if (hasCell) {
        FState s;
        state_from_U(u4.x, u4.y, u4.z, u4.w, s);
        for (int q = 0; q < 5; q++) sW[q * T + k] = s.W[q];
        if constexpr (ORDER == 1)
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
        flux_face<FLUX, ORDER, LIM, VISC>(m, cfg, Us, grad, st.Uexit, sW, sU, sG, sF, glds, T, FT, c0, nct, k, f0, fz, st.psi, fs + k, fsv);
    }
    if (h1) {
        stageN(k + B, f1);
        flux_face<FLUX, ORDER, LIM, VISC>(m, cfg, Us, grad, st.Uexit, sW, sU, sG, sF, glds, T, FT, c0, nct, k + B, f1, fz, st.psi, fs + k + B, fsv);
    }
    for (int j = k + 2 * B; j < nft; j += B) {
        FaceRec f;
        load_face<true, ORDER == 2>(m, fs + j, f);
        stageN(j, f);
        flux_face<FLUX, ORDER, LIM, VISC>(m, cfg, Us, grad, st.Uexit, sW, sU, sG, sF, glds, T, FT, c0, nct, j, f, fz, st.psi, fs + j, fsv);
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
                double4 a4 = (st.stage == 0) ? make_double4(0, 0, 0, 0) : st.acc[c];
                double acc[4] = {a4.x, a4.y, a4.z, a4.w};
                for (int q = 0; q < 4; q++) {
                    switch (st.stage) {
                    case 0: acc[q] = r[q]; Uo[q] = Un[q] - 0.5 * h * r[q]; break;
                    case 1: acc[q] += 2.0 * r[q]; Uo[q] = Un[q] - 0.5 * h * r[q]; break;
                    case 2: acc[q] += 2.0 * r[q]; Uo[q] = Un[q] - h * r[q]; break;
                    default: Uo[q] = Un[q] - (h / 6.0) * (acc[q] + r[q]); break;
                    }
                }
                if (st.stage < 3) st.acc[c] = make_double4(acc[0], acc[1], acc[2], acc[3]);
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
}

// ------------------------------------------------------------------------------------------
// Small kernels
// ------------------------------------------------------------------------------------------
__global__ void fill_kernel(double *p, double v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sums the per-tile error partials and takes the square root (Estimate_Error). One workgroup of
// 1024 threads; thread t sums tiles t, t + 1024, ... (four independent loads in flight), then a
// fixed-order tree, so the result is deterministic.
__global__ void __launch_bounds__(1024) err_finalize_kernel(const double *partial, int ntiles,
                                                            double *err4) {
    __shared__ double red[4 * 1024];
    const int t = threadIdx.x, B = blockDim.x;
    double v[4] = {0, 0, 0, 0};
    int i = t;
    for (; i + 3 * B < ntiles; i += 4 * B) {
        double a[4][4];
        for (int u = 0; u < 4; u++)
            for (int q = 0; q < 4; q++) a[u][q] = partial[4 * (i + u * B) + q];
        for (int u = 0; u < 4; u++)
            for (int q = 0; q < 4; q++) v[q] += a[u][q];
    }
    for (; i < ntiles; i += B)
        for (int q = 0; q < 4; q++) v[q] += partial[4 * i + q];
    for (int q = 0; q < 4; q++) red[q * B + t] = v[q];
    __syncthreads();
    for (int s = B / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 4; q++) red[q * B + t] += red[q * B + t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < 4; q++) err4[q] = sqrt(red[q * B]);
}

// Ghost states of the boundary faces (Apply_Boundary_Conditions, src/Boundary_Conditions.cpp:12-61):
// ghost k of owner cell gcell[k] across a face of kind gkind[k] with unit normal gnrm[k].
__global__ void __launch_bounds__(256) ghost_kernel(int n, const int *__restrict__ gcell,
                                                    const double2 *__restrict__ gnrm,
                                                    const int *__restrict__ gkind, DevCfg cfg,
                                                    const double4 *__restrict__ Us,
                                                    const double4 *__restrict__ Uexit,
                                                    double4 *__restrict__ Ug) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = gcell[k];
    const double2 nv = gnrm[k];
    const int bc = cfg.bc[gkind[k]];
    FState in, g;
    load_state((bc == BC_EXIT && Uexit) ? Uexit : Us, c, in);
    ghost_state(bc, in, nv.x, nv.y, cfg.fs, g);
    Ug[k] = make_double4(g.U[0], g.U[1], g.U[2], g.U[3]);
}

// ---- the reference's second order: Roe dissipation on limited differences ------------------
// Evaluate_Cell_Net_Flux_2O with ROE_2O (src/Net_Flux.cpp:97-128, src/Roe_Scheme.cpp:257-447)
// and Second_Order_Limiter with the 2-argument MinMod (src/Limiters.cpp:58-63,97-200), cell by
// cell on the reference's per-cell arrays (renumbered). U_all holds the physical cells then the
// ghosts. Each face is evaluated from both sides, as the reference does.
__device__ __forceinline__ double sign_of(double a) { return (double)((a > 0.0) - (a < 0.0)); }

__global__ void __launch_bounds__(256) roe2o_kernel(int nc, const RefCells rc, const double4 *__restrict__ U_all,
                                                    double zeta, double zeta1, double cfl,
                                                    double4 *__restrict__ R, double *__restrict__ dt,
                                                    double *__restrict__ dt_partial) {
    __shared__ double red[256];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double dtc = 1.7976931348623157e308;
    if (c < nc) {
        FState L;
        {
            const double4 u = U_all[c];
            state_from_U(u.x, u.y, u.z, u.w, L);
        }
        const int nf = rc.nf[c];
        double net[4] = {0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < nf; f++) {
            const int j = rc.nbr[4 * c + f];
            const double nx = rc.nrm[8 * c + 2 * f], ny = rc.nrm[8 * c + 2 * f + 1], dl = rc.len[4 * c + f];
            FState Rs;
            {
                const double4 u = U_all[j];
                state_from_U(u.x, u.y, u.z, u.w, Rs);
            }
            double A[4];
            avg_flux(L, Rs, nx, ny, dl, A);
            const double Rho_L = L.W[0], u_L = L.W[1], v_L = L.W[2], C_L = L.W[4];
            const double Vmag_L = 0.5 * (u_L * u_L + v_L * v_L);
            const double H_L = ((C_L * C_L) / (kGamma - 1.0)) + Vmag_L;
            const double Rho_R = Rs.W[0], u_R = Rs.W[1], v_R = Rs.W[2], C_R = Rs.W[4];
            const double Vmag_R = 0.5 * (u_R * u_R + v_R * v_R);
            const double H_R = ((C_R * C_R) / (kGamma - 1.0)) + Vmag_R;
            const double sqrt_Rho_R = sqrt(Rho_R);
            const double sqrt_Rho_L = sqrt(Rho_L);
            const double Term1 = 1.0 / (sqrt_Rho_L + sqrt_Rho_R);
            const double Roe_Rho = sqrt(Rho_L * Rho_R);
            const double Roe_u = Term1 * (u_R * sqrt_Rho_R + u_L * sqrt_Rho_L);
            const double Roe_v = Term1 * (v_R * sqrt_Rho_R + v_L * sqrt_Rho_L);
            const double Roe_H = Term1 * (H_R * sqrt_Rho_R + H_L * sqrt_Rho_L);
            const double Roe_Vmag = 0.5 * (Roe_u * Roe_u + Roe_v * Roe_v);
            const double Un = nx * Roe_u + ny * Roe_v;
            const double Ut = -ny * Roe_u + nx * Roe_v;
            const double Roe_a = sqrt((kGamma - 1.0) * (Roe_H - Roe_Vmag));
            // limited differences d_U (Second_Order_Limiter): quads take the next cell across
            // the opposite face, other cells fall back to first order
            int N1, N2;
            double d1, d2;
            N1 = j;
            d1 = rc.dist[4 * c + f];
            if (nf != 4) {
                if (d1 <= 0.0) d1 = 1.0;
                N2 = N1;
                d2 = d1;
            } else if (f < 2) {
                N2 = (N1 >= nc) ? N1 : rc.nbr[4 * N1 + f];
                d2 = (N1 >= nc) ? d1 : rc.dist[4 * N1 + f];
            } else {
                N2 = rc.nbr[4 * c + f - 2];
                d2 = rc.dist[4 * c + f - 2];
            }
            const double4 uc4 = U_all[c], u14 = U_all[N1], u24 = U_all[N2];
            const double uc[4] = {uc4.x, uc4.y, uc4.z, uc4.w}, u1[4] = {u14.x, u14.y, u14.z, u14.w},
                         u2[4] = {u24.x, u24.y, u24.z, u24.w};
            double d_U[4];
            for (int k = 0; k < 4; k++) {
                const double Slope1 = zeta * (uc[k] - u1[k]) / d1;
                const double Slope2 = zeta1 * (u1[k] - u2[k]) / d2;
                const double a1 = fabs(Slope1), b1 = fabs(Slope2);
                const double phi = 0.5 * (sign_of(Slope1) + sign_of(Slope2)) * (b1 < a1 ? b1 : a1);
                const double dR = u1[k] + 0.5 * phi * d1;
                const double dL = uc[k] - 0.5 * phi * d1;
                d_U[k] = dR - dL;
            }
            const double du = d_U[1] / Roe_Rho - d_U[0] * Roe_u / (Roe_Rho * Roe_Rho);
            const double dv = d_U[2] / Roe_Rho - d_U[0] * Roe_v / (Roe_Rho * Roe_Rho);
            const double dE = d_U[3] - 0.5 * Roe_Rho * (2.0 * (Roe_u * du + Roe_v * dv) + du * du + dv * dv) -
                              0.5 * d_U[0] * Roe_Vmag;
            const double dP = (kGamma - 1.0) * dE;
            const double dRho = d_U[0];
            const double dUn = nx * du + ny * dv;
            const double dUt = -ny * du + nx * dv;
            double L1 = fabs(Un - Roe_a), L2 = fabs(Un), L3 = fabs(Un), L4 = fabs(Un + Roe_a);
            const double ef = 0.1 * Roe_a;
            if (L1 < ef) L1 = 0.5 * (L1 * L1 / ef + ef);
            if (L4 < ef) L4 = 0.5 * (L4 * L4 / ef + ef);
            const double R12 = Roe_u - nx * Roe_a, R13 = Roe_v - ny * Roe_a, R14 = Roe_H - Roe_a * Un;
            const double R22 = Roe_u, R23 = Roe_v, R24 = Roe_Vmag;
            const double R32 = -ny, R33 = nx, R34 = Ut;
            const double R42 = Roe_u + nx * Roe_a, R43 = Roe_v + ny * Roe_a, R44 = Roe_H + Roe_a * Un;
            const double Term2 = 1.0 / (Roe_a * Roe_a);
            const double a1 = 0.5 * Term2 * (dP - Roe_Rho * Roe_a * dUn);
            const double a2 = dRho - Term2 * dP;
            const double a3 = Roe_Rho * dUt;
            const double a4 = 0.5 * Term2 * (dP + Roe_Rho * Roe_a * dUn);
            double D[4];
            D[0] = 0.5 * (L1 * a1 * 1.0 + L2 * a2 * 1.0 + L3 * a3 * 0.0 + L4 * a4 * 1.0) * dl;
            D[1] = 0.5 * (L1 * a1 * R12 + L2 * a2 * R22 + L3 * a3 * R32 + L4 * a4 * R42) * dl;
            D[2] = 0.5 * (L1 * a1 * R13 + L2 * a2 * R23 + L3 * a3 * R33 + L4 * a4 * R43) * dl;
            D[3] = 0.5 * (L1 * a1 * R14 + L2 * a2 * R24 + L3 * a3 * R34 + L4 * a4 * R44) * dl;
            for (int k = 0; k < 4; k++) net[k] += A[k] - D[k];
        }
        R[c] = make_double4(net[0], net[1], net[2], net[3]);
        // Inviscid_Time_Step (src/Time_Step.cpp:293-329) of the cell's own state
        double denom = 0.0;
        for (int f = 0; f < nf; f++) {
            const double vn = L.W[1] * rc.nrm[8 * c + 2 * f] + L.W[2] * rc.nrm[8 * c + 2 * f + 1];
            denom += (fabs(vn) + L.W[4]) * rc.len[4 * c + f];
        }
        dtc = (denom <= 0.0 || !isfinite(denom)) ? 0.0 : cfl * rc.area[c] / denom;
        dt[c] = dtc;
    }
    if (dt_partial) {
        const double b = block_min(red, dtc);
        if (threadIdx.x == 0) dt_partial[blockIdx.x] = b;
    }
}

// ---- batched meshes ------------------------------------------------------------------------
// After a step of a batch of independent instances: each marching instance (status 1) sums its
// tiles' error partials [inst_tile[i], inst_tile[i+1]) in a fixed order (Estimate_Error of that
// instance), counts the step and stops when Error[0] <= tol (status 0) or Error[0] is not finite
// (status -1). One workgroup of 256 threads per instance.
__device__ __forceinline__ void freeze_step(double e0, int &fz, double &best, int &since,
                                            double freeze_tol, int freeze_stall) {
    // Limiter freezing: a step that recorded psi (1) switches to frozen (2); otherwise freezing
    // starts (record on the next step) once Error[0] <= freeze_tol, or once Error[0] has not
    // dropped below 0.9 x its best value for freeze_stall steps (a limiter-switching stall).
    // A frozen march whose Error[0] exceeds 10 x the best value seen up to the recording step
    // unfreezes.
    if (fz == 1) {
        fz = 2;
        best = fmin(best, e0);
    } else if (fz == 2) {
        // the frozen factors no longer fit the state (it moved away from where they were
        // recorded): return to the evaluated limiter and start the search again
        if (!(e0 <= 10.0 * best)) {
            fz = 0;
            best = e0;
            since = 0;
        }
    } else if (fz == 0 && (freeze_tol > 0.0 || freeze_stall > 0)) {
        if (e0 < 0.9 * best) {
            best = e0;
            since = 0;
        } else {
            since += 1;
        }
        if (e0 <= freeze_tol || (freeze_stall > 0 && since >= freeze_stall)) fz = 1;
    }
}

__global__ void __launch_bounds__(1) freeze_update_kernel(const double *__restrict__ err4, int *fz,
                                                          double *best, int *since, double freeze_tol,
                                                          int freeze_stall) {
    int f = *fz, sn = *since;
    double b = *best;
    freeze_step(err4[0], f, b, sn, freeze_tol, freeze_stall);
    *fz = f;
    *best = b;
    *since = sn;
}

__global__ void __launch_bounds__(256) batch_status_kernel(const int *__restrict__ inst_tile,
                                                           const double *__restrict__ err_partial,
                                                           double tol, int *status, int *steps,
                                                           double *err4, int *freeze, double *best,
                                                           int *since, double freeze_tol,
                                                           int freeze_stall) {
    __shared__ double red[4 * 256];
    const int i = blockIdx.x, t = threadIdx.x;
    if (status[i] != 1) return;
    const int t0 = inst_tile[i], t1 = inst_tile[i + 1];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = t0 + t; j < t1; j += 256)
        for (int q = 0; q < 4; q++) v[q] += err_partial[4 * j + q];
    for (int q = 0; q < 4; q++) red[q * 256 + t] = v[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 4; q++) red[q * 256 + t] += red[q * 256 + t + s];
        __syncthreads();
    }
    if (t == 0) {
        double e[4];
        for (int q = 0; q < 4; q++) {
            e[q] = sqrt(red[q * 256]);
            err4[4 * i + q] = e[q];
        }
        steps[i] += 1;
        // Error[1..3] are infinite wherever a momentum component is exactly zero (the reference's
        // relative norm, src/Error_Estimate_Update.cpp:10-30); only Error[0] decides the march
        if (!isfinite(e[0])) status[i] = -1;
        else if (e[0] <= tol) status[i] = 0;
        if (freeze) {
            int f = freeze[i], sn = since[i];
            double b = best[i];
            freeze_step(e[0], f, b, sn, freeze_tol, freeze_stall);
            freeze[i] = f;
            best[i] = b;
            since[i] = sn;
        }
    }
}

// Offsets of the marching instances' tiles in the compacted tile list, and its length. One
// workgroup of 1024 threads scans the instances in chunks of 1024.
__global__ void __launch_bounds__(1024) batch_list_scan_kernel(int ninst, const int *__restrict__ inst_tile,
                                                               const int *__restrict__ status,
                                                               int *__restrict__ inst_off,
                                                               int *__restrict__ tile_count) {
    __shared__ int buf[1024];
    const int t = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < ninst; base += 1024) {
        const int i = base + t;
        const int v = (i < ninst && status[i] == 1) ? inst_tile[i + 1] - inst_tile[i] : 0;
        buf[t] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int add = (t >= d) ? buf[t - d] : 0;
            __syncthreads();
            buf[t] += add;
            __syncthreads();
        }
        if (i < ninst) inst_off[i] = carry + buf[t] - v;
        carry += buf[1023];
        __syncthreads();
    }
    if (t == 0) *tile_count = carry;
}

// Writes the compacted list: tile j of marching instance i goes to inst_off[i] + j - inst_tile[i].
__global__ void __launch_bounds__(256) batch_list_fill_kernel(int ntiles, const int *__restrict__ tile_inst,
                                                              const int *__restrict__ inst_tile,
                                                              const int *__restrict__ status,
                                                              const int *__restrict__ inst_off,
                                                              int *__restrict__ tile_list) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntiles) return;
    const int i = tile_inst[j];
    if (status[i] != 1) return;
    tile_list[inst_off[i] + j - inst_tile[i]] = j;
}

// Minimum of n per-tile partial minima (fmin ignores NaN, as the reference's `<` scan does).
__global__ void __launch_bounds__(1024) min_reduce_kernel(const double *partial, int n, double *out) {
    __shared__ double red[1024];
    const int t = threadIdx.x, B = blockDim.x;
    double v = 1.7976931348623157e308;
    int i = t;
    for (; i + 3 * B < n; i += 4 * B) {
        const double a0 = partial[i], a1 = partial[i + B], a2 = partial[i + 2 * B], a3 = partial[i + 3 * B];
        v = fmin(v, fmin(fmin(a0, a1), fmin(a2, a3)));
    }
    for (; i < n; i += B) v = fmin(v, partial[i]);
    red[t] = v;
    __syncthreads();
    for (int s = B / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = fmin(red[t], red[t + s]);
        __syncthreads();
    }
    if (t == 0) *out = red[0];
}

// The reference's Runge_Kutta_Method / Explicit_Method updates (src/Numerical_Method.cpp:26-103,
// 342-457) with the stage's global dt (device scalar) and, in stage 0 with LTS, the cell's dt.
// kind 0: U1 = Un - dt/A R; 1: U2 = 1/4 (U1 + 3 Un - dt/A R); 2: dU = 2/3 (U2 - Un - dt/A R),
// Uout = Un + dU; 3 (forward Euler): dU = -dt/A R, Uout = Un + dU. Kinds 2 and 3 also emit the
// error partials of dU.
__global__ void __launch_bounds__(256) ref_update_kernel(int nc, int kind, int lts,
                                                         const double *__restrict__ inv_area,
                                                         const double4 *__restrict__ Un,
                                                         const double4 *__restrict__ Us,
                                                         const double4 *__restrict__ R,
                                                         const double *__restrict__ dt_cell,
                                                         const double *__restrict__ dt_global,
                                                         double4 *__restrict__ Uout,
                                                         double *__restrict__ err_partial) {
    __shared__ double red[4 * 256];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double e[4] = {0, 0, 0, 0};
    if (c < nc) {
        const double dtg = *dt_global;
        const double ia = inv_area[c];
        const double4 n4 = Un[c], s4 = Us[c], r4 = R[c];
        const double un[4] = {n4.x, n4.y, n4.z, n4.w};
        const double us[4] = {s4.x, s4.y, s4.z, s4.w};
        const double r[4] = {r4.x, r4.y, r4.z, r4.w};
        double o[4];
        for (int q = 0; q < 4; q++) {
            switch (kind) {
            case 0: {
                const double d = lts ? dt_cell[c] : dtg;
                o[q] = un[q] - d * ia * (r[q] - 0.0);
                break;
            }
            case 1:
                o[q] = 0.25 * (us[q] + 3.0 * un[q] - dtg * ia * (r[q] - 0.0));
                break;
            case 2: {
                const double dU = (2.0 / 3.0) * (us[q] - un[q] - dtg * ia * (r[q] - 0.0));
                e[q] = err_term(dU, un[q]);
                o[q] = un[q] + dU;
                break;
            }
            default: {
                const double dU = -dtg * ia * r[q];
                e[q] = err_term(dU, un[q]);
                o[q] = un[q] + dU;
                break;
            }
            }
        }
        Uout[c] = make_double4(o[0], o[1], o[2], o[3]);
    }
    if (kind >= 2) {
        double out[4];
        block_sum4(red, e, out);
        if (threadIdx.x == 0)
            for (int q = 0; q < 4; q++) err_partial[4 * blockIdx.x + q] = out[q];
    }
}

// ------------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------------
size_t grad_smem_bytes(const DevMesh &m, int grad_type) {
    return sizeof(double) * (5 * (size_t)m.tile + (grad_type == GRAD_GG ? 8 : 4) * (size_t)m.max_tile_faces);
}

size_t flux_smem_bytes(const DevMesh &m, int order, int grad_lds) {
    size_t d = 5 * (size_t)m.tile + (order == 1 ? 4 * (size_t)m.tile : 0) + 7 * (size_t)m.max_tile_faces;
    if (order == 2 && grad_lds) d += 8 * (size_t)m.tile;
    // the block reductions of phase C need 4 doubles per thread (blockDim <= 256)
    if (d < 4 * 256) d = 4 * 256;
    return sizeof(double) * d;
}

static int grid_of(const DevMesh &m, const DevCfg &c) {
    return c.xcd_remap ? 8 * ((m.ntiles + 7) / 8) : m.ntiles;
}

hipError_t launch_grad(const DevMesh &m, const DevCfg &c, int grad_type, const double4 *Us,
                       double *grad, int block, hipStream_t s) {
    const size_t sm = grad_smem_bytes(m, grad_type);
    if (grad_type == GRAD_GG)
        hipLaunchKernelGGL(grad_tile_kernel<GRAD_GG>, dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad);
    else
        hipLaunchKernelGGL(grad_tile_kernel<GRAD_WLSQ>, dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad);
    return hipGetLastError();
}

template <int FLUX, int ORDER, int LIM>
static void launch_flux_t(const DevMesh &m, const DevCfg &c, int mode, const double4 *Us,
                          const double *grad, const StageArgs &st, int block, size_t sm,
                          hipStream_t s) {
    if constexpr (ORDER == 2) {
        if (c.viscous) {
            if (mode == MODE_RESID)
                hipLaunchKernelGGL((flux_tile_kernel<FLUX, ORDER, LIM, MODE_RESID, true>), dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad, st);
            else
                hipLaunchKernelGGL((flux_tile_kernel<FLUX, ORDER, LIM, MODE_STAGE, true>), dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad, st);
            return;
        }
    }
    if (mode == MODE_RESID)
        hipLaunchKernelGGL((flux_tile_kernel<FLUX, ORDER, LIM, MODE_RESID, false>), dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad, st);
    else
        hipLaunchKernelGGL((flux_tile_kernel<FLUX, ORDER, LIM, MODE_STAGE, false>), dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad, st);
}

template <int FLUX>
static void launch_flux_f(const DevMesh &m, const DevCfg &c, int order, int lim, int mode,
                          const double4 *Us, const double *grad, const StageArgs &st, int block,
                          size_t sm, hipStream_t s) {
    if (order == 1) {
        launch_flux_t<FLUX, 1, LIM_NONE>(m, c, mode, Us, grad, st, block, sm, s);
        return;
    }
    switch (lim) {
    case LIM_NONE: launch_flux_t<FLUX, 2, LIM_NONE>(m, c, mode, Us, grad, st, block, sm, s); break;
    case LIM_MINMOD: launch_flux_t<FLUX, 2, LIM_MINMOD>(m, c, mode, Us, grad, st, block, sm, s); break;
    case LIM_VANLEER: launch_flux_t<FLUX, 2, LIM_VANLEER>(m, c, mode, Us, grad, st, block, sm, s); break;
    default: launch_flux_t<FLUX, 2, LIM_SUPERBEE>(m, c, mode, Us, grad, st, block, sm, s); break;
    }
}

hipError_t launch_flux(const DevMesh &m, const DevCfg &c, int flux, int order, int lim, int mode,
                       const double4 *Us, const double *grad, const StageArgs &st, int block,
                       hipStream_t s) {
    const size_t sm = flux_smem_bytes(m, order, c.grad_lds);
    switch (flux) {
    case FLUX_ROE: launch_flux_f<FLUX_ROE>(m, c, order, lim, mode, Us, grad, st, block, sm, s); break;
    case FLUX_HLLC: launch_flux_f<FLUX_HLLC>(m, c, order, lim, mode, Us, grad, st, block, sm, s); break;
    case FLUX_LLF: launch_flux_f<FLUX_LLF>(m, c, order, lim, mode, Us, grad, st, block, sm, s); break;
    default: launch_flux_f<FLUX_AUSM>(m, c, order, lim, mode, Us, grad, st, block, sm, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_fill(double *p, double v, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_err_finalize(const double *partial, int n, double *err4, hipStream_t s) {
    hipLaunchKernelGGL(err_finalize_kernel, dim3(1), dim3(1024), 0, s, partial, n, err4);
    return hipGetLastError();
}

hipError_t launch_ghosts(int n, const int *gcell, const double2 *gnrm, const int *gkind,
                         const DevCfg &c, const double4 *Us, const double4 *Uexit, double4 *Ug,
                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ghost_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, gcell, gnrm, gkind, c, Us, Uexit, Ug);
    return hipGetLastError();
}

hipError_t launch_batch_status(int ninst, const int *inst_tile, const double *err_partial, double tol,
                               int *status, int *steps, double *err4, int *freeze, double *best,
                               int *since, double freeze_tol, int freeze_stall, hipStream_t s) {
    if (ninst <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_status_kernel, dim3(ninst), dim3(256), 0, s, inst_tile, err_partial, tol, status, steps, err4,
                       freeze, best, since, freeze_tol, freeze_stall);
    return hipGetLastError();
}

hipError_t launch_freeze_update(const double *err4, int *freeze, double *best, int *since,
                                double freeze_tol, int freeze_stall, hipStream_t s) {
    hipLaunchKernelGGL(freeze_update_kernel, dim3(1), dim3(1), 0, s, err4, freeze, best, since, freeze_tol, freeze_stall);
    return hipGetLastError();
}

hipError_t launch_batch_list(int ninst, int ntiles, const int *inst_tile, const int *tile_inst,
                             const int *status, int *inst_off, int *tile_list, int *tile_count,
                             hipStream_t s) {
    hipLaunchKernelGGL(batch_list_scan_kernel, dim3(1), dim3(1024), 0, s, ninst, inst_tile, status, inst_off, tile_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ntiles <= 0) return e;
    hipLaunchKernelGGL(batch_list_fill_kernel, dim3((ntiles + 255) / 256), dim3(256), 0, s, ntiles, tile_inst, inst_tile, status, inst_off, tile_list);
    return hipGetLastError();
}

hipError_t launch_roe2o(int nc, const RefCells &rc, const double4 *U_all, double zeta, double zeta1,
                        double cfl, double4 *R, double *dt, double *dt_partial, hipStream_t s) {
    if (nc <= 0) return hipSuccess;
    hipLaunchKernelGGL(roe2o_kernel, dim3((nc + 255) / 256), dim3(256), 0, s, nc, rc, U_all, zeta, zeta1, cfl, R, dt, dt_partial);
    return hipGetLastError();
}

hipError_t launch_min_reduce(const double *partial, int n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(min_reduce_kernel, dim3(1), dim3(1024), 0, s, partial, n, out);
    return hipGetLastError();
}

hipError_t launch_ref_update(int nc, int kind, int lts, const double *inv_area, const double4 *Un,
                             const double4 *Us, const double4 *R, const double *dt_cell,
                             const double *dt_global, double4 *Uout, double *err_partial,
                             hipStream_t s) {
    if (nc <= 0) return hipSuccess;
    hipLaunchKernelGGL(ref_update_kernel, dim3((nc + 255) / 256), dim3(256), 0, s, nc, kind, lts,
                       inv_area, Un, Us, R, dt_cell, dt_global, Uout, err_partial);
    return hipGetLastError();
}

}  // namespace cfd
