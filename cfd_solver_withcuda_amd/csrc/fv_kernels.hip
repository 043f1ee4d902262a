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

#include <algorithm>

#include "fv_flux.h"

namespace cfd {

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

// PERSIST: persistent workgroups walking the tiles (as flux_tile_kernel).
template <int GRAD, bool PERSIST = false>
__global__ void __launch_bounds__(256) grad_tile_kernel(DevMesh m, DevCfg cfg,
                                                        const double4 *__restrict__ Us,
                                                        double *__restrict__ grad) {
    extern __shared__ double smem[];
    const int ntr = tiles_to_run(m);
    const int span = PERSIST ? tile_span(ntr, cfg.xcd_remap) : 0;
    int v = blockIdx.x;
    do {
    const int tile = virtual_tile(m, v, ntr, cfg.xcd_remap);
    if (tile < 0) continue;
    if (PERSIST && v != (int)blockIdx.x) __syncthreads();  // the previous tile's LDS reads are done
    double fsv[5];
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
    } while (PERSIST && (v += gridDim.x) < span);
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
    const double fsv[5] = {cfg.fs[0], cfg.fs[1], cfg.fs[2], cfg.fs[3], (cfg.p_exit > 0.0) ? cfg.p_exit : cfg.fs[3]};
    ghost_state(bc, in, nv.x, nv.y, fsv, g);
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
    size_t d = 5 * (size_t)m.tile + (order != 2 ? 4 * (size_t)m.tile : 0) + 7 * (size_t)m.max_tile_faces;
    if (order >= 2 && grad_lds) d += 8 * (size_t)m.tile;
    // the block reductions of phase C need 4 doubles per thread (blockDim <= 256)
    if (d < 4 * 256) d = 4 * 256;
    return sizeof(double) * d;
}

hipError_t launch_grad(const DevMesh &m, const DevCfg &c, int grad_type, const double4 *Us,
                       double *grad, int block, hipStream_t s) {
    const size_t sm = grad_smem_bytes(m, grad_type);
    if (c.persist) {
        if (grad_type == GRAD_GG) {
            auto k = grad_tile_kernel<GRAD_GG, true>;
            hipLaunchKernelGGL(k, dim3(grid_persist(k, m, c, block, sm)), dim3(block), sm, s, m, c, Us, grad);
        } else {
            auto k = grad_tile_kernel<GRAD_WLSQ, true>;
            hipLaunchKernelGGL(k, dim3(grid_persist(k, m, c, block, sm)), dim3(block), sm, s, m, c, Us, grad);
        }
    } else if (grad_type == GRAD_GG) {
        hipLaunchKernelGGL(grad_tile_kernel<GRAD_GG>, dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad);
    } else {
        hipLaunchKernelGGL(grad_tile_kernel<GRAD_WLSQ>, dim3(grid_of(m, c)), dim3(block), sm, s, m, c, Us, grad);
    }
    return hipGetLastError();
}

int persist_grid(const void *kernel, const DevMesh &m, const DevCfg &c, int block, size_t sm) {
    const int full = grid_of(m, c);
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, sm) != hipSuccess || per_cu < 1)
        return full;  // no occupancy figure: one workgroup per tile
    const int g = std::max(8, (per_cu * cus) / 8 * 8);
    return std::min(g, full);
}

// LDS of one flux launch: the face-geometry copy (3 doubles per face) only when the launch forms
// time steps (the kernel's needN), at least the block reductions' 4 doubles per thread.
size_t flux_launch_smem(const DevMesh &m, int order, int grad_lds, int block, bool geom) {
    size_t d = 5 * (size_t)m.tile + (order != 2 ? 4 * (size_t)m.tile : 0) +
               (geom ? 7 : 4) * (size_t)m.max_tile_faces;
    if (order >= 2 && grad_lds) d += 8 * (size_t)m.tile;
    if (d < 4 * (size_t)block) d = 4 * (size_t)block;
    return sizeof(double) * d;
}

hipError_t launch_flux(const DevMesh &m, const DevCfg &c, int flux, int order, int lim, int mode,
                       const double4 *Us, const double *grad, const StageArgs &st, int block,
                       hipStream_t s) {
    const bool geom = (mode == MODE_RESID) || (st.lts && st.stage == 0) || (st.last && st.dt_partial);
    const size_t sm = flux_launch_smem(m, order, c.grad_lds, block, geom);
    if (c.persist) {
        if (mode != MODE_STAGE || order > 2 || c.viscous) return hipErrorInvalidValue;
        return launch_flux_persist(m, c, flux, order, lim, Us, grad, st, block, sm, s);
    }
    if (order == 3)
        return c.viscous ? launch_flux_weno_visc(m, c, flux, lim, mode, Us, grad, st, block, sm, s)
                         : launch_flux_weno(m, c, flux, lim, mode, Us, grad, st, block, sm, s);
    if (order == 2 && c.viscous) return launch_flux_visc(m, c, flux, lim, mode, Us, grad, st, block, sm, s);
    if (order == 1) return launch_flux_family<1, false>(m, c, flux, lim, mode, Us, grad, st, block, sm, s);
    return launch_flux_family<2, false>(m, c, flux, lim, mode, Us, grad, st, block, sm, s);
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

// dst = src over the cells of the tiles still marching (one workgroup per list position).
__global__ void __launch_bounds__(256) batch_commit_kernel(const int *__restrict__ tile_list,
                                                           const int *__restrict__ tile_count,
                                                           const int *__restrict__ tile_cstart,
                                                           const double4 *__restrict__ src,
                                                           double4 *__restrict__ dst) {
    if ((int)blockIdx.x >= *tile_count) return;
    const int t = tile_list[blockIdx.x];
    for (int c = tile_cstart[t] + threadIdx.x; c < tile_cstart[t + 1]; c += blockDim.x) dst[c] = src[c];
}

hipError_t launch_batch_commit(const DevMesh &m, int grid_tiles, const double4 *src, double4 *dst, hipStream_t s) {
    if (grid_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_commit_kernel, dim3(grid_tiles), dim3(128), 0, s, m.tile_list, m.tile_count,
                       m.tile_cstart, src, dst);
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
