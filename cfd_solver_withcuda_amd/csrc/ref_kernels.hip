// Reference-layout kernels: the launch_* entry points of rameshkolluru43/CFD_Solver_withCUDA's
// CUDA_KERNELS/ (Advanced_Flux_Schemes_Cuda.h, Limiter_Cuda_Kernels.h,
// Reconstruction_Schemes_Cuda.h, Boundary_Conditions_Cuda_Kernels.h) with their array layouts:
//   U[N*4] conserved, P[N*6] = (rho, u, v, P, T, a), neighbours[N*4], normals[N*8],
//   face lengths[N*4], per-face outputs [(N*4 + face)*4 + var].
// One thread per cell (per list entry for BCs). A neighbour outside [0, N) marks a boundary face,
// which the flux kernels skip, as the reference's do.
//
// Deviations from the kernel text, each because the text is degenerate or mis-paired:
//   - limiters: the text builds slope_backward = -slope_forward, so r = -1 and phi = 0 always
//     (Limiter_Cuda_Kernels.cu:306-307); here r = (U_i - U_opp)/d_opp / ((U_j - U_i)/d_f) with
//     U_opp the neighbour across the opposite face (f+2)%4;
//   - MUSCL and 2nd-order Roe take the neighbour's limiter value from the neighbour's face that
//     points back to the cell (the text reads the neighbour's entry at the same face index).
#include <hip/hip_runtime.h>

#include "../../include/cfd_hip.h"
#include "../../include/cfd_hip_ref.h"
#include "ref_kernels.h"

namespace cfd {
namespace {

constexpr double kRGas = 287.0;  // R_GC (include/definitions.h:66)

// apply_entropy_fix_device (Roe_Flux_Cuda_Kernels.cu:51-57)
__device__ __forceinline__ double ref_entropy_fix(double lambda, double a) {
    double delta = 0.1 * a;
    if (fabs(lambda) < delta) return 0.5 * (lambda * lambda / delta + delta);
    return fabs(lambda);
}

// compute_roe_averages_device (Roe_Flux_Cuda_Kernels.cu:28-46)
__device__ __forceinline__ void ref_roe_avg(double rho_L, double u_L, double v_L, double H_L,
                                            double rho_R, double u_R, double v_R, double H_R,
                                            double gamma, double &ru, double &rv, double &rH,
                                            double &ra) {
    double sL = sqrt(rho_L), sR = sqrt(rho_R);
    double term = 1.0 / (sL + sR);
    ru = term * (u_L * sL + u_R * sR);
    rv = term * (v_L * sL + v_R * sR);
    rH = term * (H_L * sL + H_R * sR);
    double vm = 0.5 * (ru * ru + rv * rv);
    ra = sqrt(fmax((gamma - 1.0) * (rH - vm), 1e-14));
}

// Roe flux 0.5 (F_L + F_R - |A| dU) from left/right primitive + total energy states.
__device__ __forceinline__ void ref_roe_flux(double rho_L, double u_L, double v_L, double P_L,
                                             double E_L, double H_L, double rho_R, double u_R,
                                             double v_R, double P_R, double E_R, double H_R,
                                             double nx, double ny, double gamma, double area,
                                             double *out) {
    double roe_u, roe_v, roe_H, roe_a;
    ref_roe_avg(rho_L, u_L, v_L, H_L, rho_R, u_R, v_R, H_R, gamma, roe_u, roe_v, roe_H, roe_a);
    double roe_rho = sqrt(rho_L * rho_R);
    double Un = nx * roe_u + ny * roe_v;
    double Ut = -ny * roe_u + nx * roe_v;
    double du = u_R - u_L, dv = v_R - v_L, dP = P_R - P_L, drho = rho_R - rho_L;
    double dUn = nx * du + ny * dv;
    double dUt = -ny * du + nx * dv;
    double a2_inv = 1.0 / (roe_a * roe_a);
    double al1 = 0.5 * a2_inv * (dP - roe_rho * roe_a * dUn);
    double al2 = drho - a2_inv * dP;
    double al3 = roe_rho * dUt;
    double al4 = 0.5 * a2_inv * (dP + roe_rho * roe_a * dUn);
    double l1 = ref_entropy_fix(Un - roe_a, roe_a), l2 = fabs(Un), l3 = fabs(Un);
    double l4 = ref_entropy_fix(Un + roe_a, roe_a);
    double R12 = roe_u - nx * roe_a, R13 = roe_v - ny * roe_a, R14 = roe_H - roe_a * Un;
    double R24 = 0.5 * (roe_u * roe_u + roe_v * roe_v);
    double R42 = roe_u + nx * roe_a, R43 = roe_v + ny * roe_a, R44 = roe_H + roe_a * Un;
    double t1 = l1 * al1, t2 = l2 * al2, t3 = l3 * al3, t4 = l4 * al4;
    double D1 = 0.5 * (t1 * 1.0 + t2 * 1.0 + t3 * 0.0 + t4 * 1.0);
    double D2 = 0.5 * (t1 * R12 + t2 * roe_u + t3 * (-ny) + t4 * R42);
    double D3 = 0.5 * (t1 * R13 + t2 * roe_v + t3 * nx + t4 * R43);
    double D4 = 0.5 * (t1 * R14 + t2 * R24 + t3 * Ut + t4 * R44);
    double Un_L = nx * u_L + ny * v_L, Un_R = nx * u_R + ny * v_R;
    double FL0 = rho_L * Un_L, FL1 = rho_L * u_L * Un_L + P_L * nx, FL2 = rho_L * v_L * Un_L + P_L * ny;
    double FL3 = (E_L + P_L) * Un_L;
    double FR0 = rho_R * Un_R, FR1 = rho_R * u_R * Un_R + P_R * nx, FR2 = rho_R * v_R * Un_R + P_R * ny;
    double FR3 = (E_R + P_R) * Un_R;
    out[0] = 0.5 * (FL0 + FR0 - D1) * area;
    out[1] = 0.5 * (FL1 + FR1 - D2) * area;
    out[2] = 0.5 * (FL2 + FR2 - D3) * area;
    out[3] = 0.5 * (FL3 + FR3 - D4) * area;
}

__device__ __forceinline__ bool unit_normal(const double *normals, int cell, int face, double &nx,
                                            double &ny) {
    nx = normals[cell * 8 + face * 2 + 0];
    ny = normals[cell * 8 + face * 2 + 1];
    double mag = sqrt(nx * nx + ny * ny);
    if (mag < 1e-14) return false;
    nx /= mag;
    ny /= mag;
    return true;
}

// face of cell j whose neighbour is cell i (-1 if none)
__device__ __forceinline__ int back_face(const int *nb, int j, int i) {
    for (int g = 0; g < 4; g++)
        if (nb[j * 4 + g] == i) return g;
    return -1;
}

}  // namespace

// ---- first-order fluxes ------------------------------------------------------------------
__global__ void ref_roe_flux_kernel(const double *U, const double *P, const int *nb,
                                    const double *normals, const double *areas, double *out,
                                    int N, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        if (j < 0 || j >= N) continue;
        double nx, ny;
        if (!unit_normal(normals, i, f, nx, ny)) continue;
        const double area = areas[i * 4 + f];
        const double rL = P[i * 6], uL = P[i * 6 + 1], vL = P[i * 6 + 2], pL = P[i * 6 + 3], aL = P[i * 6 + 5];
        if (rL <= 0.0 || pL <= 0.0 || aL <= 0.0) continue;
        const double HL = (aL * aL) / (gamma - 1.0) + 0.5 * (uL * uL + vL * vL);
        const double rR = P[j * 6], uR = P[j * 6 + 1], vR = P[j * 6 + 2], pR = P[j * 6 + 3], aR = P[j * 6 + 5];
        if (rR <= 0.0 || pR <= 0.0 || aR <= 0.0) continue;
        const double HR = (aR * aR) / (gamma - 1.0) + 0.5 * (uR * uR + vR * vR);
        ref_roe_flux(rL, uL, vL, pL, U[i * 4 + 3], HL, rR, uR, vR, pR, U[j * 4 + 3], HR, nx, ny,
                     gamma, area, &out[(i * 4 + f) * 4]);
    }
}

// HLLC as the kernel text writes it (HLLC_LLF_Flux_Cuda_Kernels.cu:25-208): sqrt(rho)-weighted
// wave-speed estimate, the star state through rho (S_K - u_n) / (S_K - S*) and E / rho. The
// native path's hllc_flux is the same flux with shared reciprocals; this keeps the text's
// rounding for the reference-layout entry point.
__device__ __forceinline__ void ref_hllc_text(const FState &L, const FState &R, double nx, double ny,
                                              double dl, double *F) {
    double rho_L = L.W[0], u_L = L.W[1], v_L = L.W[2], P_L = L.W[3], a_L = L.W[4];
    double rho_R = R.W[0], u_R = R.W[1], v_R = R.W[2], P_R = R.W[3], a_R = R.W[4];
    double un_L = nx * u_L + ny * v_L;
    double un_R = nx * u_R + ny * v_R;
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    double a_roe = (a_L * sqrt_rho_L + a_R * sqrt_rho_R) / (sqrt_rho_L + sqrt_rho_R);
    double u_roe = (u_L * sqrt_rho_L + u_R * sqrt_rho_R) / (sqrt_rho_L + sqrt_rho_R);
    double v_roe = (v_L * sqrt_rho_L + v_R * sqrt_rho_R) / (sqrt_rho_L + sqrt_rho_R);
    double un_roe = nx * u_roe + ny * v_roe;
    double S_L = fmin(un_L - a_L, un_roe - a_roe);
    double S_R = fmax(un_R + a_R, un_roe + a_roe);
    double numer = P_R - P_L + rho_L * un_L * (S_L - un_L) - rho_R * un_R * (S_R - un_R);
    double denom = rho_L * (S_L - un_L) - rho_R * (S_R - un_R);
    double S_star = (fabs(denom) > 1e-14) ? numer / denom : 0.5 * (un_L + un_R);
    double f[4];
    if (S_L >= 0.0) {
        f[0] = rho_L * un_L;
        f[1] = rho_L * u_L * un_L + P_L * nx;
        f[2] = rho_L * v_L * un_L + P_L * ny;
        f[3] = (L.U[3] + P_L) * un_L;
    } else if (S_R <= 0.0) {
        f[0] = rho_R * un_R;
        f[1] = rho_R * u_R * un_R + P_R * nx;
        f[2] = rho_R * v_R * un_R + P_R * ny;
        f[3] = (R.U[3] + P_R) * un_R;
    } else {
        const bool left = (S_star >= 0.0);
        const FState &K = left ? L : R;
        double S_K = left ? S_L : S_R;
        double rho = K.W[0], u = K.W[1], v = K.W[2], P = K.W[3], E = K.U[3];
        double un = nx * u + ny * v;
        double factor = rho * (S_K - un) / (S_K - S_star);
        double Us0 = factor;
        double Us1 = factor * (u + (S_star - un) * nx);
        double Us2 = factor * (v + (S_star - un) * ny);
        double Us3 = factor * (E / rho + (S_star - un) * (S_star + P / (rho * (S_K - un))));
        f[0] = rho * un + S_K * (Us0 - K.U[0]);
        f[1] = (rho * u * un + P * nx) + S_K * (Us1 - K.U[1]);
        f[2] = (rho * v * un + P * ny) + S_K * (Us2 - K.U[2]);
        f[3] = ((E + P) * un) + S_K * (Us3 - K.U[3]);
    }
    for (int q = 0; q < 4; q++) F[q] = f[q] * dl;
}

__global__ void ref_hllc_flux_kernel(const double *U, const double *P, const int *nb,
                                     const double *normals, const double *areas, double *out,
                                     int N, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        if (j < 0 || j >= N) continue;
        double nx, ny;
        if (!unit_normal(normals, i, f, nx, ny)) continue;
        const double area = areas[i * 4 + f];
        FState L, R;
        L.W[0] = P[i * 6]; L.W[1] = P[i * 6 + 1]; L.W[2] = P[i * 6 + 2]; L.W[3] = P[i * 6 + 3]; L.W[4] = P[i * 6 + 5];
        if (L.W[0] <= 0.0 || L.W[3] <= 0.0 || L.W[4] <= 0.0) continue;
        R.W[0] = P[j * 6]; R.W[1] = P[j * 6 + 1]; R.W[2] = P[j * 6 + 2]; R.W[3] = P[j * 6 + 3]; R.W[4] = P[j * 6 + 5];
        if (R.W[0] <= 0.0 || R.W[3] <= 0.0 || R.W[4] <= 0.0) continue;
        for (int q = 0; q < 4; q++) {
            L.U[q] = U[i * 4 + q];
            R.U[q] = U[j * 4 + q];
        }
        ref_hllc_text(L, R, nx, ny, area, &out[(i * 4 + f) * 4]);
    }
}

__global__ void ref_llf_flux_kernel(const double *U, const double *P, const int *nb,
                                    const double *normals, const double *areas, double *out,
                                    int N, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    (void)gamma;
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        if (j < 0 || j >= N) continue;
        double nx, ny;
        if (!unit_normal(normals, i, f, nx, ny)) continue;
        const double area = areas[i * 4 + f];
        const double rL = P[i * 6], uL = P[i * 6 + 1], vL = P[i * 6 + 2], pL = P[i * 6 + 3], aL = P[i * 6 + 5];
        if (rL <= 0.0 || pL <= 0.0 || aL <= 0.0) continue;
        const double rR = P[j * 6], uR = P[j * 6 + 1], vR = P[j * 6 + 2], pR = P[j * 6 + 3], aR = P[j * 6 + 5];
        if (rR <= 0.0 || pR <= 0.0 || aR <= 0.0) continue;
        const double unL = nx * uL + ny * vL, unR = nx * uR + ny * vR;
        const double lmax = fmax(fabs(unL) + aL, fabs(unR) + aR);
        double FL[4] = {rL * unL, rL * uL * unL + pL * nx, rL * vL * unL + pL * ny, (U[i * 4 + 3] + pL) * unL};
        double FR[4] = {rR * unR, rR * uR * unR + pR * nx, rR * vR * unR + pR * ny, (U[j * 4 + 3] + pR) * unR};
        for (int q = 0; q < 4; q++)
            out[(i * 4 + f) * 4 + q] = 0.5 * (FL[q] + FR[q] - lmax * (U[j * 4 + q] - U[i * 4 + q])) * area;
    }
}

// ---- second-order Roe with limited reconstruction ------------------------------------------
__global__ void ref_roe2_flux_kernel(const double *U, const double *P, const int *nb,
                                     const double *normals, const double *areas,
                                     const double *phi, const double *ctr, const double *fctr,
                                     double *out, int N, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    (void)P;
    const double xc = ctr[i * 2], yc = ctr[i * 2 + 1];
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        if (j < 0 || j >= N) continue;
        double nx, ny;
        if (!unit_normal(normals, i, f, nx, ny)) continue;
        const double area = areas[i * 4 + f];
        const double xf = fctr[(i * 4 + f) * 2], yf = fctr[(i * 4 + f) * 2 + 1];
        const double dL = sqrt((xf - xc) * (xf - xc) + (yf - yc) * (yf - yc));
        const double xr = ctr[j * 2], yr = ctr[j * 2 + 1];
        const double dR = sqrt((xf - xr) * (xf - xr) + (yf - yr) * (yf - yr));
        const double wL = (dL + dR > 0.0) ? dL / (dL + dR) : 0.5;
        const double wR = 1.0 - wL;
        const int g = back_face(nb, j, i);
        double UL[4], UR[4];
        for (int q = 0; q < 4; q++) {
            const double d = U[j * 4 + q] - U[i * 4 + q];
            const double pL = phi[(i * 4 + f) * 4 + q];
            const double pR = (g >= 0) ? phi[(j * 4 + g) * 4 + q] : 0.0;
            UL[q] = U[i * 4 + q] + pL * d * wL;
            UR[q] = U[j * 4 + q] - pR * d * wR;
        }
        const double rho_L = fmax(UL[0], 1e-14);
        const double u_L = UL[1] / rho_L, v_L = UL[2] / rho_L, E_L = UL[3];
        const double P_L = fmax((gamma - 1.0) * (E_L - 0.5 * rho_L * (u_L * u_L + v_L * v_L)), 1e-14);
        const double H_L = (E_L + P_L) / rho_L;
        const double rho_R = fmax(UR[0], 1e-14);
        const double u_R = UR[1] / rho_R, v_R = UR[2] / rho_R, E_R = UR[3];
        const double P_R = fmax((gamma - 1.0) * (E_R - 0.5 * rho_R * (u_R * u_R + v_R * v_R)), 1e-14);
        const double H_R = (E_R + P_R) / rho_R;
        ref_roe_flux(rho_L, u_L, v_L, P_L, E_L, H_L, rho_R, u_R, v_R, P_R, E_R, H_R, nx, ny, gamma,
                     area, &out[(i * 4 + f) * 4]);
    }
}

// ---- limiters ------------------------------------------------------------------------------
__device__ __forceinline__ double ref_phi(int type, double a, double b) {
    // a = backward slope, b = forward slope; phi(r), r = a/b (Limiter_Cuda_Kernels.cu:89-157)
    if (fabs(b) < 1e-12) return 0.0;
    const double r = a / b;
    switch (type) {
    case REF_LIM_VANLEER:
        return (r <= 0.0) ? 0.0 : 2.0 * r / (1.0 + r);
    case REF_LIM_SUPERBEE:
        return (r <= 0.0) ? 0.0 : fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0));
    case REF_LIM_VANALBADA:
        return (r <= 0.0) ? 0.0 : (r * r + r) / (r * r + 1.0);
    default:  // minmod: Phi(r) (include/definitions.h:74-75)
        return (r >= 1.0) ? 1.0 : ((r < 1.0) && (r > 0.0)) ? r : 0.0;
    }
}

__global__ void ref_limiter_kernel(const double *U, const int *nb, const double *dist, double *phi,
                                   int N, double zeta, int type, int three_arg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        const int o = nb[i * 4 + (f + 2) % 4];
        for (int q = 0; q < 4; q++) {
            double v = 0.0;
            if (j >= 0 && j < N) {
                const double fwd = zeta * (U[j * 4 + q] - U[i * 4 + q]) / dist[i * 4 + f];
                const double bwd = (o >= 0 && o < N) ? zeta * (U[i * 4 + q] - U[o * 4 + q]) / dist[i * 4 + (f + 2) % 4] : 0.0;
                v = ref_phi(type, bwd, fwd);
                if (three_arg && type == REF_LIM_MINMOD && fabs(fwd) >= 1e-12) {
                    // minmod(r, 1, (1 + r) / 2), the 3-argument form
                    const double r = bwd / fwd;
                    const double c = 0.5 * (1.0 + r);
                    v = (r > 0.0) ? fmin(fmin(r, 1.0), c) : 0.0;
                }
            }
            phi[(i * 4 + f) * 4 + q] = v;
        }
    }
}

// venkatakrishnan_limiter_device (Limiter_Cuda_Kernels.cu:162-179)
__global__ void ref_venkat_kernel(const double *U, const int *nb, const double *dist,
                                  const double *vol, double *phi, int N, double K) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double dj = sqrt(vol[i]);
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        const int o = nb[i * 4 + (f + 2) % 4];
        for (int q = 0; q < 4; q++) {
            double v = 1.0;
            if (j >= 0 && j < N) {
                const double dp = U[j * 4 + q] - U[i * 4 + q];
                const double dm = (o >= 0 && o < N) ? U[i * 4 + q] - U[o * 4 + q] : 0.0;
                if (fabs(dp) >= 1e-12) {
                    const double eps2 = (K * dj) * (K * dj) * (K * dj);
                    const double num = (dp * dp + eps2) * dm + 2.0 * dm * dm * dp;
                    const double den = dp * dp + 2.0 * dm * dm + dp * dm + eps2;
                    v = (fabs(den) < 1e-12) ? 1.0 : num / den;
                }
            }
            phi[(i * 4 + f) * 4 + q] = v;
        }
    }
    (void)dist;
}

// barth_jespersen_limiter_kernel (Limiter_Cuda_Kernels.cu:443-489)
__global__ void ref_bj_kernel(const double *U, const int *nb, const double *dist,
                              const double *unlim, double *phi, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int q = 0; q < 4; q++) {
        const double uc = U[i * 4 + q];
        double umax = uc, umin = uc;
        for (int f = 0; f < 4; f++) {
            const int j = nb[i * 4 + f];
            if (j >= 0 && j < N) {
                umax = fmax(umax, U[j * 4 + q]);
                umin = fmin(umin, U[j * 4 + q]);
            }
        }
        for (int f = 0; f < 4; f++) {
            const int j = nb[i * 4 + f];
            double v = 0.0;
            if (j >= 0) {
                const double du = unlim[(i * 4 + f) * 4 + q] * dist[i * 4 + f];
                if (fabs(du) < 1e-12) {
                    v = 1.0;
                } else {
                    double p = 1.0;
                    if (du > 0.0) p = fmin(1.0, (umax - uc) / du);
                    else if (du < 0.0) p = fmin(1.0, (umin - uc) / du);
                    v = fmax(0.0, p);
                }
            }
            phi[(i * 4 + f) * 4 + q] = v;
        }
    }
}

// muscl_reconstruction_kernel (MUSCL_WENO_Reconstruction_Cuda_Kernels.cu:29-83)
__global__ void ref_muscl_kernel(const double *U, const int *nb, const double *dist,
                                 const double *phi, double *QL, double *QR, int N, double kappa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int f = 0; f < 4; f++) {
        const int j = nb[i * 4 + f];
        if (j < 0 || j >= N) {
            for (int q = 0; q < 4; q++) {
                QL[(i * 4 + f) * 4 + q] = U[i * 4 + q];
                QR[(i * 4 + f) * 4 + q] = U[i * 4 + q];
            }
            continue;
        }
        const double d = dist[i * 4 + f];
        const int g = back_face(nb, j, i);
        for (int q = 0; q < 4; q++) {
            const double Ui = U[i * 4 + q], Uj = U[j * 4 + q];
            const double pL = phi[(i * 4 + f) * 4 + q];
            const double pR = (g >= 0) ? phi[(j * 4 + g) * 4 + q] : 0.0;
            QL[(i * 4 + f) * 4 + q] = Ui + 0.5 * kappa * pL * ((Uj - Ui) / d) * d;
            QR[(i * 4 + f) * 4 + q] = Uj + 0.5 * kappa * pR * ((Ui - Uj) / d) * d;
        }
    }
}

// ---- in-place boundary conditions (Boundary_Conditions_Cuda_Kernels.cu:60-324) ------------
__device__ __forceinline__ void bc_prim(const double *U, double &rho, double &u, double &v,
                                       double &p) {
    rho = U[0];
    if (rho < 1e-12) rho = 1e-12;
    u = U[1] / rho;
    v = U[2] / rho;
    p = kGm1 * (U[3] - 0.5 * rho * (u * u + v * v));
    if (p < 1e-12) p = 1e-12;
}

__device__ __forceinline__ void bc_store(double *U, double rho, double u, double v, double p) {
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = p / kGm1 + 0.5 * rho * (u * u + v * v);
}

__global__ void ref_bc_kernel(int kind, double *Ucells, const int *list, const double *normals,
                              RefBcData bc, int n, double gamma) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int cell = list[idx];
    double *U = &Ucells[cell * 4];
    double rho, u, v, p;
    switch (kind) {
    case REF_BC_SUBSONIC_INLET:
        bc_prim(U, rho, u, v, p);
        bc_store(U, bc.rho, u, v, bc.p);
        break;
    case REF_BC_SUPERSONIC_INLET:
        bc_store(U, bc.rho, bc.u, bc.v, bc.p);
        break;
    case REF_BC_SUBSONIC_EXIT:
        bc_prim(U, rho, u, v, p);
        bc_store(U, rho, u, v, bc.p);
        break;
    case REF_BC_SUPERSONIC_EXIT:
        break;  // zero-order extrapolation: nothing to change
    case REF_BC_VISCOUS_WALL:
        U[1] = -U[1];
        U[2] = -U[2];
        break;
    case REF_BC_INVISCID_WALL:
    case REF_BC_SYMMETRY: {
        rho = U[0];
        if (rho < 1e-12) rho = 1e-12;
        u = U[1] / rho;
        v = U[2] / rho;
        const double nx = normals[cell * 8], ny = normals[cell * 8 + 1];
        const double vn = u * nx + v * ny;
        U[1] = rho * (u - 2.0 * vn * nx);
        U[2] = rho * (v - 2.0 * vn * ny);
        break;
    }
    default: {  // far field, Riemann invariants
        bc_prim(U, rho, u, v, p);
        const double nx = normals[cell * 8], ny = normals[cell * 8 + 1];
        const double ai = sqrt(kGamma * p / rho), af = sqrt(kGamma * bc.p / bc.rho);
        const double vni = u * nx + v * ny, vnf = bc.u * nx + bc.v * ny;
        const double Rp = vni + 2.0 * ai / kGm1, Rm = vnf - 2.0 * af / kGm1;
        const double vnb = 0.5 * (Rp + Rm);
        double ab = 0.25 * kGm1 * (Rp - Rm);
        if (ab < 1e-12) ab = 1e-12;
        double rb, ub, vb, pb;
        if (vnb >= 0.0) {
            const double s = p / pow(rho, kGamma);
            rb = pow(ab * ab / (kGamma * s), 1.0 / kGm1);
            ub = u + (vnb - vni) * nx;
            vb = v + (vnb - vni) * ny;
        } else {
            const double s = bc.p / pow(bc.rho, kGamma);
            rb = pow(ab * ab / (kGamma * s), 1.0 / kGm1);
            ub = bc.u + (vnb - vnf) * nx;
            vb = bc.v + (vnb - vnf) * ny;
        }
        pb = rb * ab * ab / kGamma;
        bc_store(U, rb, ub, vb, pb);
        break;
    }
    }
    (void)gamma;
    (void)kRGas;
}

// ---- Runge-Kutta stages on the reference's 4-variable arrays -------------------------------
// runge_kutta_stage_kernel (CUDA_KERNELS/Cuda_Kernels.h:4-33): the reference's delta form,
// stages 1..3; R_visc may be null (inviscid).
__global__ void ref_rk3_stage_kernel(const double *U, const double *R, const double *Rv,
                                     double *Urk, const double *inv_area, double dt, int N,
                                     int stage) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double ia = inv_area[i];
    for (int q = 0; q < 4; q++) {
        const double r = R[i * 4 + q] - (Rv ? Rv[i * 4 + q] : 0.0);
        const double u = U[i * 4 + q];
        if (stage == 1) Urk[i * 4 + q] = u - dt * ia * r;
        else if (stage == 2) Urk[i * 4 + q] = 0.25 * (Urk[i * 4 + q] + 3.0 * u - dt * ia * r);
        else if (stage == 3) Urk[i * 4 + q] = (2.0 / 3.0) * (Urk[i * 4 + q] - u - dt * ia * r);
    }
}

// classical RK4 stage s = 1..4 (stage s of the native integrator): Ustage = stage state out,
// K = accumulator; stage 4 writes U^{n+1} into Ustage.
__global__ void ref_rk4_stage_kernel(const double *U, const double *R, double *Ustage, double *K,
                                     const double *inv_area, double dt, int N, int stage) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double h = dt * inv_area[i];
    for (int q = 0; q < 4; q++) {
        const int e = i * 4 + q;
        const double r = R[e], un = U[e];
        switch (stage) {
        case 1: K[e] = r; Ustage[e] = un - 0.5 * h * r; break;
        case 2: K[e] += 2.0 * r; Ustage[e] = un - 0.5 * h * r; break;
        case 3: K[e] += 2.0 * r; Ustage[e] = un - h * r; break;
        default: Ustage[e] = un - (h / 6.0) * (K[e] + r); break;
        }
    }
}

}  // namespace cfd

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
namespace {
inline int bs(int b) { return (b <= 0) ? 256 : b; }
inline int grid(int n, int b) { return (n + b - 1) / b; }
inline int ret() { return (int)hipGetLastError(); }
}  // namespace

using namespace cfd;

extern "C" {

int launch_roe_flux(const double *U, const double *P, const int *nb, const double *nrm,
                    const double *ar, double *out, int N, double gamma, int block, void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !P || !nb || !nrm || !ar || !out || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_roe_flux_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, P, nb, nrm, ar, out, N, gamma);
    return ret();
}

int launch_roe_flux_2nd_order(const double *U, const double *P, const int *nb, const double *nrm,
                              const double *ar, const double *phi, const double *ctr,
                              const double *fctr, double *out, int N, double gamma, int block,
                              void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !nb || !nrm || !ar || !phi || !ctr || !fctr || !out || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_roe2_flux_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, P, nb, nrm, ar, phi, ctr, fctr, out, N, gamma);
    return ret();
}

int launch_hllc_flux(const double *U, const double *P, const int *nb, const double *nrm,
                     const double *ar, double *out, int N, double gamma, int block, void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !P || !nb || !nrm || !ar || !out || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_hllc_flux_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, P, nb, nrm, ar, out, N, gamma);
    return ret();
}

int launch_llf_flux(const double *U, const double *P, const int *nb, const double *nrm,
                    const double *ar, double *out, int N, double gamma, int block, void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !P || !nb || !nrm || !ar || !out || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_llf_flux_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, P, nb, nrm, ar, out, N, gamma);
    return ret();
}

static int limiter(const double *U, const int *nb, const double *d, double *phi, int N, double zeta,
                   int type, int three, int block, void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !nb || !d || !phi || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_limiter_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, nb, d, phi, N, zeta, type, three);
    return ret();
}

int launch_minmod_limiter(const double *U, const int *nb, const double *d, double *phi, int N,
                          double zeta, int use_three_arg, int block, void *stream) {
    return limiter(U, nb, d, phi, N, zeta, REF_LIM_MINMOD, use_three_arg ? 1 : 0, block, stream);
}
int launch_vanleer_limiter(const double *U, const int *nb, const double *d, double *phi, int N,
                           double zeta, int block, void *stream) {
    return limiter(U, nb, d, phi, N, zeta, REF_LIM_VANLEER, 0, block, stream);
}
int launch_superbee_limiter(const double *U, const int *nb, const double *d, double *phi, int N,
                            double zeta, int block, void *stream) {
    return limiter(U, nb, d, phi, N, zeta, REF_LIM_SUPERBEE, 0, block, stream);
}
int launch_vanalbada_limiter(const double *U, const int *nb, const double *d, double *phi, int N,
                             double zeta, int block, void *stream) {
    return limiter(U, nb, d, phi, N, zeta, REF_LIM_VANALBADA, 0, block, stream);
}

int launch_venkatakrishnan_limiter(const double *U, const int *nb, const double *d,
                                   const double *vol, double *phi, int N, double K, int block,
                                   void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !nb || !d || !vol || !phi || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_venkat_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, nb, d, vol, phi, N, K);
    return ret();
}

int launch_barth_jespersen_limiter(const double *U, const int *nb, const double *d,
                                   const double *unlim, double *phi, int N, int block,
                                   void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !nb || !d || !unlim || !phi || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_bj_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, nb, d, unlim, phi, N);
    return ret();
}

int launch_generic_limiter(const double *U, const int *nb, const double *d, double *phi, int N,
                           ::LimiterType type, double zeta, int block, void *stream) {
    // generic_limiter_kernel (Limiter_Cuda_Kernels.cu:492-544) dispatches 0..3, others -> minmod
    int t = ((int)type >= 0 && (int)type <= 3) ? (int)type : REF_LIM_MINMOD;
    return limiter(U, nb, d, phi, N, zeta, t, 0, block, stream);
}

const char *get_limiter_name(::LimiterType t) {
    switch (t) {
    case LIMITER_MINMOD: return "MinMod";
    case LIMITER_VANLEER: return "Van Leer";
    case LIMITER_SUPERBEE: return "Superbee";
    case LIMITER_VANALBADA: return "Van Albada";
    case LIMITER_VENKATAKRISHNAN: return "Venkatakrishnan";
    case LIMITER_BARTH_JESPERSEN: return "Barth-Jespersen";
    default: return "Unknown";
    }
}

int launch_muscl_reconstruction(const double *U, const int *nb, const double *d, const double *phi,
                                double *QL, double *QR, int N, double kappa, int block,
                                void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !nb || !d || !phi || !QL || !QR || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_muscl_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, nb, d, phi, QL, QR, N, kappa);
    return ret();
}

static int bc(int kind, double *U, const int *list, const double *nrm, RefBcData data, int n,
              double gamma, int block, void *stream) {
    if (n == 0) return 0;
    if (n < 0 || !U || !list || block < 0 || block > 1024) return CFD_E_ARG;
    if ((kind == REF_BC_INVISCID_WALL || kind == REF_BC_SYMMETRY || kind == REF_BC_FARFIELD) && !nrm) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_bc_kernel, dim3(grid(n, block)), dim3(block), 0, (hipStream_t)stream, kind, U, list, nrm, data, n, gamma);
    return ret();
}

int launch_subsonic_inlet_bc(double *U, const int *list, const double *nrm,
                             const InletCondition_GPU *d, int n, double gamma, int block,
                             void *stream) {
    if (!d) return CFD_E_ARG;
    return bc(REF_BC_SUBSONIC_INLET, U, list, nrm, RefBcData{d->Rho, d->u, d->v, d->P}, n, gamma, block, stream);
}
int launch_supersonic_inlet_bc(double *U, const int *list, const InletCondition_GPU *d, int n,
                               double gamma, int block, void *stream) {
    if (!d) return CFD_E_ARG;
    return bc(REF_BC_SUPERSONIC_INLET, U, list, nullptr, RefBcData{d->Rho, d->u, d->v, d->P}, n, gamma, block, stream);
}
int launch_subsonic_exit_bc(double *U, const int *list, const double *nrm,
                            const ExitCondition_GPU *d, int n, double gamma, int block,
                            void *stream) {
    if (!d) return CFD_E_ARG;
    return bc(REF_BC_SUBSONIC_EXIT, U, list, nrm, RefBcData{0.0, 0.0, 0.0, d->P}, n, gamma, block, stream);
}
int launch_supersonic_exit_bc(double *U, const int *list, int n, int block, void *stream) {
    return bc(REF_BC_SUPERSONIC_EXIT, U, list, nullptr, RefBcData{0, 0, 0, 0}, n, 1.4, block, stream);
}
int launch_viscous_wall_bc(double *U, const int *list, int n, int block, void *stream) {
    return bc(REF_BC_VISCOUS_WALL, U, list, nullptr, RefBcData{0, 0, 0, 0}, n, 1.4, block, stream);
}
int launch_inviscid_wall_bc(double *U, const int *list, const double *nrm, int n, int block,
                            void *stream) {
    return bc(REF_BC_INVISCID_WALL, U, list, nrm, RefBcData{0, 0, 0, 0}, n, 1.4, block, stream);
}
int launch_symmetry_bc(double *U, const int *list, const double *nrm, int n, int block,
                       void *stream) {
    return bc(REF_BC_SYMMETRY, U, list, nrm, RefBcData{0, 0, 0, 0}, n, 1.4, block, stream);
}
int launch_farfield_bc(double *U, const int *list, const double *nrm,
                       const FarfieldCondition_GPU *d, int n, double gamma, int block,
                       void *stream) {
    if (!d) return CFD_E_ARG;
    return bc(REF_BC_FARFIELD, U, list, nrm, RefBcData{d->Rho_inf, d->u_inf, d->v_inf, d->P_inf}, n, gamma, block, stream);
}

int launch_tvd_rk3_stage(const double *U, const double *R, const double *Rv, double *Urk,
                         const double *ia, double dt, int N, int stage, int block, void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !R || !Urk || !ia || stage < 1 || stage > 3 || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_rk3_stage_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, R, Rv, Urk, ia, dt, N, stage);
    return ret();
}

int launch_rk4_stage(const double *U, const double *R, double *Us, double *K, const double *ia,
                     double dt, int N, int stage, int block, void *stream) {
    if (N == 0) return 0;
    if (N < 0 || !U || !R || !Us || !K || !ia || stage < 1 || stage > 4 || block < 0 || block > 1024) return CFD_E_ARG;
    block = bs(block);
    hipLaunchKernelGGL(ref_rk4_stage_kernel, dim3(grid(N, block)), dim3(block), 0, (hipStream_t)stream, U, R, Us, K, ia, dt, N, stage);
    return ret();
}

}  // extern "C"
