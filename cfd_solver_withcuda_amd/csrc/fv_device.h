// Device-side FP64 physics of the explicit finite-volume path: primitive recovery, ghost
// states, face fluxes (Roe + entropy fix, HLLC, LLF, AUSM), limited slopes.
//
// Every function keeps the operation order of the reference expression it restates, so that
// with floating-point contraction disabled (-ffp-contract=off) each kernel reproduces the CPU
// restatement bit for bit. References are to rameshkolluru43/CFD_Solver_withCUDA.
#pragma once
#include <hip/hip_runtime.h>

#include "cfd_types.h"

namespace cfd {

constexpr double kGamma = 1.4;
constexpr double kGm1 = (kGamma - 1.0);
constexpr double kGamma1 = (kGamma / kGm1);  // include/definitions.h:65 (gamma / gamma_M_1)
constexpr double kInvGm1 = 1.0 / kGm1;
// Regularisation of the van Albada limiter (squared difference of the non-dimensional primitives
// below which slopes are not limited).
constexpr double kVanAlbadaEps = 1e-4;

// Face state: primitives (rho, u, v, p, a) and conserved (rho, rho u, rho v, rho E).
struct FState {
    double W[5];
    double U[4];
};

// Calculate_Primitive_Variables (src/Basic_Functions.cpp:85-142): velocity flush |v|<=1e-10.
__device__ __forceinline__ void state_from_U(double U0, double U1, double U2, double U3, FState &s) {
    double Rho = U0;
    double inv = 1.0 / Rho;
    double v1 = U1 * inv;
    double v2 = U2 * inv;
    if (fabs(v1) <= 1e-10) v1 = 0.0;
    if (fabs(v2) <= 1e-10) v2 = 0.0;
    double vmag = (v1 * v1 + v2 * v2);
    double P = kGm1 * (U3 - 0.5 * Rho * vmag);
    s.W[0] = Rho;
    s.W[1] = v1;
    s.W[2] = v2;
    s.W[3] = P;
    s.W[4] = sqrt(kGamma * P * inv);
    s.U[0] = U0;
    s.U[1] = U1;
    s.U[2] = U2;
    s.U[3] = U3;
}

// Density and pressure of a conserved state, with state_from_U's arithmetic.
__device__ __forceinline__ void rho_p_of(const double *U, double &rho, double &p) {
    double Rho = U[0];
    double inv = 1.0 / Rho;
    double v1 = U[1] * inv;
    double v2 = U[2] * inv;
    if (fabs(v1) <= 1e-10) v1 = 0.0;
    if (fabs(v2) <= 1e-10) v2 = 0.0;
    double vmag = (v1 * v1 + v2 * v2);
    rho = Rho;
    p = kGm1 * (U[3] - 0.5 * Rho * vmag);
}

// Calculate_Computational_Variables case 2 (src/Basic_Functions.cpp:171-206).
__device__ __forceinline__ void cons_from_prim(double rho, double u, double v, double p, double *U) {
    double vmag = 0.5 * (u * u + v * v);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = ((p / kGm1) + rho * vmag);
}

// Conserved ghost state of the subsonic outflow and inflow boundaries (types 9 and 10) from the
// interior primitives (r, u, v, p, sound speed c), the outward normal and the prescribed state
// (f0..f3 = rho, u, v, p; f4 = exit static pressure). Called rather than inlined: its power
// functions would otherwise raise the register count of every flux and gradient kernel (the
// step mesh's stage kernel from 108 to 156 VGPRs, 4 to 3 waves per SIMD), where these
// boundaries are a small share of the faces.
__device__ __noinline__ double4 ghost_subsonic(int bc, double r, double u, double v, double p, double c,
                                                     double n1, double n2, double f0, double f1, double f2,
                                                     double f3, double f4) {
    double Ug[4];
    if (bc == BC_SUBSONIC_EXIT) {
        // back-pressure exit (Subsonic_Exit_Boundary_Condition, src/Exit_Boundary_Conditions.cpp:8-66):
        // the face state takes the prescribed static pressure and the outgoing characteristics of
        // the interior (reference density and sound speed: the interior cell's), and the ghost is
        // 2 U_face - U_cell
        const double pb = f4;
        const double t = (p - pb) / (r * c * c);
        const double rf = r - t * r;
        const double uf = u + n1 * t * c;
        const double vf = v + n2 * t * c;
        double Uf[4], Uc[4];
        cons_from_prim(rf, uf, vf, pb, Uf);
        cons_from_prim(r, u, v, p, Uc);
        for (int q = 0; q < 4; q++) Ug[q] = 2.0 * Uf[q] - Uc[q];
        double rg, pg;
        rho_p_of(Ug, rg, pg);
        if (!(rg > 0.0) || !(pg > 0.0))
            for (int q = 0; q < 4; q++) Ug[q] = Uf[q];  // a non-physical extrapolation: the face state
    } else {
        // total-condition inlet: total pressure, total density and flow direction of the
        // prescribed state, static pressure from the interior (isentropic; the reference's
        // subsonic inlet prescribes rho and p instead)
        const double q2 = f1 * f1 + f2 * f2;
        const double f = 1.0 + 0.5 * kGm1 * (q2 / (kGamma * f3 / f0));
        const double p0 = f3 * pow(f, kGamma / kGm1);
        const double r0 = f0 * pow(f, 1.0 / kGm1);
        const double qinf = sqrt(q2);
        const double dx = (qinf > 0.0) ? f1 / qinf : -n1;
        const double dy = (qinf > 0.0) ? f2 / qinf : -n2;
        const double pi = fmin(p, p0);
        const double rb = r0 * pow(pi / p0, 1.0 / kGamma);
        const double qb = sqrt(fmax((2.0 / kGm1) * (kGamma * p0 / r0 - kGamma * pi / rb), 0.0));
        cons_from_prim(rb, qb * dx, qb * dy, pi, Ug);
    }
    return make_double4(Ug[0], Ug[1], Ug[2], Ug[3]);
}

// Ghost state across a boundary face with outward unit normal (n1, n2); fs = the prescribed state
// (rho, u, v, p) and the exit's static pressure.
__device__ __forceinline__ void ghost_state(int bc, const FState &in, double n1, double n2,
                                            const double *fs, FState &g) {
    double Ug[4];
    switch (bc) {
    case BC_INLET_REF: {  // Supersonic_Inlet_Condition (src/Inlet_Boundary_Conditions.cpp:41-74)
        double u = in.W[1], v = in.W[2];
        double vmag = (u * u + v * v) * 0.5;
        Ug[0] = fs[0];
        Ug[1] = fs[0] * u;
        Ug[2] = fs[0] * v;
        Ug[3] = (fs[3] / kGm1) + fs[0] * vmag;
        break;
    }
    case BC_SUBSONIC_INLET_REF: {  // Subsonic_Inlet_Condition (:7-39), energy as written there
        double u = in.W[1], v = in.W[2];
        double Vm = 0.5 * (u * u + v * v);
        Ug[0] = fs[0];
        Ug[1] = fs[0] * u;
        Ug[2] = fs[0] * v;
        Ug[3] = (fs[3] / kGm1) + fs[0] * Vm * Vm;
        break;
    }
    case BC_WALL:
    case BC_SYMMETRY: {  // InViscid_Wall / Symmetry (src/Wall_Boundary_Conditions.cpp:8-124)
        double u = in.W[1], v = in.W[2];
        double vdotn = (u * n1 + v * n2);
        double v1g = u - 2.0 * vdotn * n1;
        double v2g = v - 2.0 * vdotn * n2;
        double vmag = 0.5 * (v1g * v1g + v2g * v2g);
        Ug[0] = in.W[0];
        Ug[1] = in.W[0] * v1g;
        Ug[2] = in.W[0] * v2g;
        Ug[3] = ((in.W[3] / kGm1) + in.W[0] * vmag);
        break;
    }
    case BC_WALL_NOSLIP: {  // Viscous_Wall_Boundary_Condition (src/Wall_Boundary_Conditions.cpp:55-76):
        // momentum negated, density and energy kept; formed from the primitives, as the face
        // states hold them
        double u = in.W[1], v = in.W[2];
        double vmag = 0.5 * (u * u + v * v);
        Ug[0] = in.W[0];
        Ug[1] = in.W[0] * (-u);
        Ug[2] = in.W[0] * (-v);
        Ug[3] = ((in.W[3] / kGm1) + in.W[0] * vmag);
        break;
    }
    case BC_FREESTREAM:
        cons_from_prim(fs[0], fs[1], fs[2], fs[3], Ug);
        break;
    case BC_FARFIELD: {  // characteristic far field (Riemann invariants normal to the face)
        double ri = in.W[0], ui = in.W[1], vi = in.W[2], pi = in.W[3], ci = in.W[4];
        double rf = fs[0], uf = fs[1], vf = fs[2], pf = fs[3];
        double cf = sqrt(kGamma * pf / rf);
        double vni = ui * n1 + vi * n2, vnf = uf * n1 + vf * n2;
        double rb, ub, vb, pb;
        if (fabs(vni) >= ci) {
            if (vni < 0.0) { rb = rf; ub = uf; vb = vf; pb = pf; }
            else { rb = ri; ub = ui; vb = vi; pb = pi; }
        } else {
            double Rp = vni + 2.0 * ci / kGm1;
            double Rm = vnf - 2.0 * cf / kGm1;
            double vnb = 0.5 * (Rp + Rm);
            double cb = 0.25 * kGm1 * (Rp - Rm);
            double s, ut, vt;
            if (vnb < 0.0) {
                s = pf / pow(rf, kGamma);
                ut = uf - vnf * n1;
                vt = vf - vnf * n2;
            } else {
                s = pi / pow(ri, kGamma);
                ut = ui - vni * n1;
                vt = vi - vni * n2;
            }
            rb = pow(cb * cb / (kGamma * s), 1.0 / kGm1);
            pb = rb * cb * cb / kGamma;
            ub = ut + vnb * n1;
            vb = vt + vnb * n2;
        }
        cons_from_prim(rb, ub, vb, pb, Ug);
        break;
    }
    case BC_SUBSONIC_EXIT:
    case BC_SUBSONIC_INLET: {
        if (bc == BC_SUBSONIC_EXIT && in.W[1] * n1 + in.W[2] * n2 >= in.W[4]) {
            g = in;  // supersonic outflow copies the interior state
            return;
        }
        const double4 G = ghost_subsonic(bc, in.W[0], in.W[1], in.W[2], in.W[3], in.W[4], n1, n2, fs[0], fs[1],
                                         fs[2], fs[3], fs[4]);
        Ug[0] = G.x;
        Ug[1] = G.y;
        Ug[2] = G.z;
        Ug[3] = G.w;
        break;
    }
    case BC_EXIT:  // Supersonic_Exit_Boundary_Condition (src/Exit_Boundary_Conditions.cpp:68-81)
    default:
        g = in;
        return;
    }
    state_from_U(Ug[0], Ug[1], Ug[2], Ug[3], g);
}

// Calculate_Face_Average_Flux (src/Average_Interface_Flux.cpp:20-82): the mean of the two sides'
// Euler fluxes through the face, times its length (same operations in the same order).
__device__ __forceinline__ void avg_flux(const FState &L, const FState &R, double nx, double ny,
                                         double dl, double *A) {
    const double rl = L.W[0], ul = L.W[1], vl = L.W[2], pl = L.W[3];
    const double rr = R.W[0], ur = R.W[1], vr = R.W[2], pr = R.W[3];
    const double qnl = (ul * nx + vl * ny);
    const double qnr = (ur * nx + vr * ny);
    const double kel = 0.5 * (ul * ul + vl * vl);
    const double ker = 0.5 * (ur * ur + vr * vr);
    const double qnl_dl = qnl * dl;
    const double qnr_dl = qnr * dl;
    A[0] = 0.5 * (rl * qnl_dl + rr * qnr_dl);
    A[1] = 0.5 * ((rl * ul * qnl_dl + pl * nx * dl) + (rr * ur * qnr_dl + pr * nx * dl));
    A[2] = 0.5 * ((rl * vl * qnl_dl + pl * ny * dl) + (rr * vr * qnr_dl + pr * ny * dl));
    A[3] = 0.5 * (((kGamma1 * pl + rl * kel) * qnl_dl) + ((kGamma1 * pr + rr * ker) * qnr_dl));
}

// ROE dissipation (src/Roe_Scheme.cpp:6-256), Harten entropy fix delta = 0.1 a~:
// D = 1/2 sum_k |lambda_k| alpha_k r_k dl over the four waves of the Roe-averaged state (same
// operations in the same order as the reference, with this file's own names).
__device__ __forceinline__ void roe_diss(const FState &L, const FState &R, double nx, double ny,
                                         double dl, double *D) {
    D[0] = D[1] = D[2] = D[3] = 0.0;
    const double nmag = sqrt(nx * nx + ny * ny);
    if (nmag < 1e-12) return;
    const double rl = L.W[0], ul = L.W[1], vl = L.W[2], pl = L.W[3], cl = L.W[4];
    if (rl <= 0.0 || pl <= 0.0 || cl <= 0.0) return;
    const double rr = R.W[0], ur = R.W[1], vr = R.W[2], pr = R.W[3], cr = R.W[4];
    if (rr <= 0.0 || pr <= 0.0 || cr <= 0.0) return;
    const double kel = 0.5 * (ul * ul + vl * vl);
    const double hl = ((cl * cl) / (kGamma - 1.0)) + kel;
    const double ker = 0.5 * (ur * ur + vr * vr);
    const double hr = ((cr * cr) / (kGamma - 1.0)) + ker;
    const double wl = sqrt(rl);  // sqrt(rho) weights of the Roe average
    const double wr = sqrt(rr);
    const double inv_w = 1.0 / (wl + wr);
    const double rho_t = sqrt(rl * rr);
    const double u_t = inv_w * (ur * wr + ul * wl);
    const double v_t = inv_w * (vr * wr + vl * wl);
    const double h_t = inv_w * (hr * wr + hl * wl);
    const double ke_t = 0.5 * (u_t * u_t + v_t * v_t);
    double c_t = sqrt((kGamma - 1.0) * (h_t - ke_t));
    if (c_t <= 0.0) {
        c_t = 0.5 * (cl + cr);
        if (c_t <= 0.0) return;
    }
    const double qn = nx * u_t + ny * v_t;   // normal and tangential Roe velocities
    const double qt = -ny * u_t + nx * v_t;
    const double du = (ur - ul), dv = (vr - vl), dp = (pr - pl), drho = (rr - rl);
    const double dqn = nx * du + ny * dv;
    const double dqt = -ny * du + nx * dv;
    // right eigenvectors (mass component 1, 1, 0, 1): acoustic -, entropy, shear, acoustic +
    const double e1u = u_t - nx * c_t, e1v = v_t - ny * c_t, e1e = h_t - c_t * qn;
    const double e2u = u_t, e2v = v_t, e2e = ke_t;
    const double e3u = -ny, e3v = nx, e3e = qt;
    const double e4u = u_t + nx * c_t, e4v = v_t + ny * c_t, e4e = h_t + c_t * qn;
    const double inv_c2 = 1.0 / (c_t * c_t);
    const double s1 = 0.5 * inv_c2 * (dp - rho_t * c_t * dqn);  // wave strengths
    const double s2 = drho - inv_c2 * dp;
    const double s3 = rho_t * dqt;
    const double s4 = 0.5 * inv_c2 * (dp + rho_t * c_t * dqn);
    double l1 = fabs(qn - c_t), l2 = fabs(qn), l3 = fabs(qn), l4 = fabs(qn + c_t);
    const double ef = 0.1 * c_t;
    if (l1 < ef) l1 = 0.5 * (l1 * l1 / ef + ef);
    if (l4 < ef) l4 = 0.5 * (l4 * l4 / ef + ef);
    const double t1 = l1 * s1, t2 = l2 * s2, t3 = l3 * s3, t4 = l4 * s4;
    D[0] = 0.5 * (t1 * 1.0 + t2 * 1.0 + t3 * 0.0 + t4 * 1.0) * dl;
    D[1] = 0.5 * (t1 * e1u + t2 * e2u + t3 * e3u + t4 * e4u) * dl;
    D[2] = 0.5 * (t1 * e1v + t2 * e2v + t3 * e3v + t4 * e4v) * dl;
    D[3] = 0.5 * (t1 * e1e + t2 * e2e + t3 * e3e + t4 * e4e) * dl;
}

// The larger of three values, first-listed winning ties (Maximum, src/Utility_Files.cpp:5-13).
__device__ __forceinline__ double max3_ref(double a, double b, double c) {
    if ((a >= b) && (a >= c)) return a;
    if ((b >= a) && (b >= c)) return b;
    return c;
}

// LLF dissipation (src/LLF.cpp:8-83): 1/2 of the largest |qn -+ c|, |qn| over both sides, times
// the jump of the conserved state.
__device__ __forceinline__ void llf_diss(const FState &L, const FState &R, double nx, double ny,
                                         double dl, double *D) {
    const double qnl = (L.W[1] * nx + L.W[2] * ny);
    const double qnr = (R.W[1] * nx + R.W[2] * ny);
    const double cl = L.W[4], cr = R.W[4];
    const double sl = max3_ref(fabs(qnl - cl) * dl, fabs(qnl + cl) * dl, fabs(qnl) * dl);
    const double sr = max3_ref(fabs(qnr - cr) * dl, fabs(qnr + cr) * dl, fabs(qnr) * dl);
    const double smax = (sl >= sr) ? sl : sr;
    for (int k = 0; k < 4; k++) D[k] = 0.5 * smax * (R.U[k] - L.U[k]);
}

// HLLC (CUDA_KERNELS/HLLC_LLF_Flux_Cuda_Kernels.cu:25-208), times the face length.
__device__ __forceinline__ void hllc_flux(const FState &L, const FState &R, double nx, double ny,
                                          double dl, double *F) {
    double rho_L = L.W[0], u_L = L.W[1], v_L = L.W[2], P_L = L.W[3], a_L = L.W[4];
    double rho_R = R.W[0], u_R = R.W[1], v_R = R.W[2], P_R = R.W[3], a_R = R.W[4];
    double un_L = nx * u_L + ny * v_L;
    double un_R = nx * u_R + ny * v_R;
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    double inv_s = 1.0 / (sqrt_rho_L + sqrt_rho_R);
    double a_roe = (a_L * sqrt_rho_L + a_R * sqrt_rho_R) * inv_s;
    double u_roe = (u_L * sqrt_rho_L + u_R * sqrt_rho_R) * inv_s;
    double v_roe = (v_L * sqrt_rho_L + v_R * sqrt_rho_R) * inv_s;
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
        // star state with one reciprocal: U*_3 = (E (S_K - un) + (S* - un)(rho (S_K - un) S* + P)) / (S_K - S*)
        double inv_k = 1.0 / (S_K - S_star);
        double mK = rho * (S_K - un);
        double factor = mK * inv_k;
        double Us0 = factor;
        double Us1 = factor * (u + (S_star - un) * nx);
        double Us2 = factor * (v + (S_star - un) * ny);
        double Us3 = (E * (S_K - un) + (S_star - un) * (mK * S_star + P)) * inv_k;
        f[0] = rho * un + S_K * (Us0 - K.U[0]);
        f[1] = (rho * u * un + P * nx) + S_K * (Us1 - K.U[1]);
        f[2] = (rho * v * un + P * ny) + S_K * (Us2 - K.U[2]);
        f[3] = ((E + P) * un) + S_K * (Us3 - K.U[3]);
    }
    for (int i = 0; i < 4; i++) F[i] = f[i] * dl;
}

// AUSM (src/Ausm_Flux.cpp:36-160), from the conserved variables.
__device__ __forceinline__ void ausm_flux(const double *UL, const double *UR, double nx, double ny,
                                          double face_area, double *F) {
    double rho_L = UL[0], u_L = UL[1] / rho_L, v_L = UL[2] / rho_L, rhoE_L = UL[3];
    double p_L = (kGamma - 1.0) * (rhoE_L - 0.5 * rho_L * (u_L * u_L + v_L * v_L));
    double H_L = (rhoE_L + p_L) / rho_L;
    double rho_R = UR[0], u_R = UR[1] / rho_R, v_R = UR[2] / rho_R, rhoE_R = UR[3];
    double p_R = (kGamma - 1.0) * (rhoE_R - 0.5 * rho_R * (u_R * u_R + v_R * v_R));
    double H_R = (rhoE_R + p_R) / rho_R;
    double Vn_L = u_L * nx + v_L * ny;
    double Vn_R = u_R * nx + v_R * ny;
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    double inv_sqrt_sum = 1.0 / (sqrt_rho_L + sqrt_rho_R);
    double u_roe = (sqrt_rho_L * u_L + sqrt_rho_R * u_R) * inv_sqrt_sum;
    double v_roe = (sqrt_rho_L * v_L + sqrt_rho_R * v_R) * inv_sqrt_sum;
    double H_roe = (sqrt_rho_L * H_L + sqrt_rho_R * H_R) * inv_sqrt_sum;
    double a_roe = sqrt((kGamma - 1.0) * (H_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe)));
    double M_L = Vn_L / a_roe;
    double M_R = Vn_R / a_roe;
    double M_plus = (fabs(M_L) >= 1.0) ? ((M_L > 0.0) ? M_L : 0.0) : 0.25 * (M_L + 1.0) * (M_L + 1.0);
    double M_minus = (fabs(M_R) >= 1.0) ? ((M_R < 0.0) ? M_R : 0.0) : -0.25 * (M_R - 1.0) * (M_R - 1.0);
    double M_interface = M_plus + M_minus;
    double P_plus = (fabs(M_L) >= 1.0) ? ((M_L > 0.0) ? 1.0 : 0.0)
                                       : 0.25 * (M_L + 1.0) * (M_L + 1.0) * (2.0 - M_L);
    double P_minus = (fabs(M_R) >= 1.0) ? ((M_R < 0.0) ? 1.0 : 0.0)
                                        : 0.25 * (M_R - 1.0) * (M_R - 1.0) * (2.0 + M_R);
    double p_interface = P_plus * p_L + P_minus * p_R;
    const bool up = (M_interface >= 0.0);
    double mass_flux = M_interface * a_roe * (up ? rho_L : rho_R);
    double uu = up ? u_L : u_R, vv = up ? v_L : v_R, HH = up ? H_L : H_R;
    F[0] = mass_flux * face_area;
    F[1] = (mass_flux * uu + p_interface * nx) * face_area;
    F[2] = (mass_flux * vv + p_interface * ny) * face_area;
    F[3] = mass_flux * HH * face_area;
}

template <int FLUX>
__device__ __forceinline__ void face_flux(bool boundary, int roe_bdiss, const FState &L,
                                          const FState &R, double nx, double ny, double dl,
                                          double *F) {
    if constexpr (FLUX == FLUX_ROE || FLUX == FLUX_LLF) {
        double A[4], D[4];
        avg_flux(L, R, nx, ny, dl, A);
        if constexpr (FLUX == FLUX_ROE) {
            if (boundary && !roe_bdiss) {
                D[0] = D[1] = D[2] = D[3] = 0.0;
            } else {
                roe_diss(L, R, nx, ny, dl, D);
            }
        } else {
            llf_diss(L, R, nx, ny, dl, D);
        }
        for (int k = 0; k < 4; k++) F[k] = A[k] - D[k];
    } else if constexpr (FLUX == FLUX_HLLC) {
        hllc_flux(L, R, nx, ny, dl, F);
    } else {
        ausm_flux(L.U, R.U, nx, ny, dl, F);
    }
}

// phi(r) * dplus for r = dminus / dplus, without the division (limiter functions of
// CUDA_KERNELS/Limiter_Cuda_Kernels.cu:89-157).
template <int LIM>
__device__ __forceinline__ double limited_slope(double a, double b) {
    if constexpr (LIM == LIM_NONE) {
        return 0.5 * (a + b);
    } else if constexpr (LIM == LIM_MINMOD) {
        if (a * b <= 0.0) return 0.0;
        return (b > 0.0) ? fmin(a, b) : fmax(a, b);
    } else if constexpr (LIM == LIM_VANLEER) {
        if (a * b <= 0.0) return 0.0;
        return 2.0 * a * b / (a + b);
    } else if constexpr (LIM == LIM_VANALBADA) {
        // van Albada, phi(r) = (r^2 + r) / (r^2 + 1) (Limiter_Cuda_Kernels.cu, vanalbada_limiter_device),
        // in the differentiable form (a (b^2 + e) + b (a^2 + e)) / (a^2 + b^2 + 2 e): differences
        // well above sqrt(e) are limited as phi(r) b, smaller ones take the central slope
        return (a * (b * b + kVanAlbadaEps) + b * (a * a + kVanAlbadaEps)) / (a * a + b * b + 2.0 * kVanAlbadaEps);
    } else {
        if (a * b <= 0.0) return 0.0;
        double aa = fabs(a), bb = fabs(b);
        double s = fmax(fmin(2.0 * aa, bb), fmin(aa, 2.0 * bb));
        return (b > 0.0) ? s : -s;
    }
}

// MUSCL face state on one side: W_f = W_c + sgn/2 * slope(dminus, dplus) with
// dplus = sgn (W_other - W_c), dminus = 2 grad(W_c).d - dplus; falls back to the cell state when
// the reconstructed density or pressure is not positive.
// Limiter freezing: the limited slope equals psi * c with c = grad W . d the unlimited central
// slope; freeze = 1 records psi[k] = slope / c (0 where c = 0), rounded to FP32; freeze = 2 uses
// slope = psi[k] * c (smooth in the state, so steady marches do not stall on limiter switching).
// FP32 factors halve the frozen march's extra reads.
template <int LIM>
__device__ __forceinline__ void reconstruct_side(const FState &c, const FState &o, const double *g,
                                                 double dx, double dy, double sgn, int freeze,
                                                 float *psi, FState &out) {
    double Wf[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double dplus = sgn * (o.W[k] - c.W[k]);
        const double cc = g[2 * k] * dx + g[2 * k + 1] * dy;
        double slope;
        if (freeze == 2) {
            slope = (double)psi[k] * cc;
        } else {
            const double dminus = 2.0 * cc - dplus;
            slope = limited_slope<LIM>(dminus, dplus);
            if (freeze == 1) psi[k] = (float)((cc != 0.0) ? slope / cc : 0.0);
        }
        Wf[k] = c.W[k] + sgn * 0.5 * slope;
    }
    if (!(Wf[0] > 0.0) || !(Wf[3] > 0.0)) {
        // first order on this side: the cell's primitives, conserved values formed from them
        for (int k = 0; k < 5; k++) out.W[k] = c.W[k];
        cons_from_prim(c.W[0], c.W[1], c.W[2], c.W[3], out.U);
        return;
    }
    out.W[0] = Wf[0];
    out.W[1] = Wf[1];
    out.W[2] = Wf[2];
    out.W[3] = Wf[3];
    out.W[4] = sqrt(kGamma * Wf[3] * (1.0 / Wf[0]));
    // conserved values of the face state, with p / (gamma - 1) as a product by the reciprocal
    const double vmag = 0.5 * (Wf[1] * Wf[1] + Wf[2] * Wf[2]);
    out.U[0] = Wf[0];
    out.U[1] = Wf[0] * Wf[1];
    out.U[2] = Wf[0] * Wf[2];
    out.U[3] = Wf[3] * kInvGm1 + Wf[0] * vmag;
}

// WENO_Reconstruction (src/WENO2D.cpp:9-79): fifth-order WENO-JS (epsilon 1e-6, p = 2) of one
// variable on the stencil a..e centred on c; shift 0 gives the value at the edge between c and d,
// shift 1 the value at the edge between b and c. Squares are products (the reference's
// pow(x, 2)), so device and host agree bit for bit.
__device__ __forceinline__ double weno5(double a, double b, double c, double d, double e, int shift) {
    if (!isfinite(a) || !isfinite(b) || !isfinite(c) || !isfinite(d) || !isfinite(e)) return c;
    const double eps = 1e-6;
    double d0, d1, d2, v0, v1, v2;
    if (shift == 0) {
        d0 = 3.0 / 10.0;
        d1 = 6.0 / 10.0;
        d2 = 1.0 / 10.0;
        v2 = (2.0 * a - 7.0 * b + 11.0 * c) / 6.0;
        v1 = (-1.0 * b + 5.0 * c + 2.0 * d) / 6.0;
        v0 = (2.0 * c + 5.0 * d - 1.0 * e) / 6.0;
    } else {
        d0 = 1.0 / 10.0;
        d1 = 6.0 / 10.0;
        d2 = 3.0 / 10.0;
        v2 = (-1.0 * a + 5.0 * b + 2.0 * c) / 6.0;
        v1 = (2.0 * b + 5.0 * c - 1.0 * d) / 6.0;
        v0 = (11.0 * c - 7.0 * d + 2.0 * e) / 6.0;
    }
    const double s2a = a - 2.0 * b + c, s2b = a - 4.0 * b + 3.0 * c;
    const double s1a = b - 2.0 * c + d, s1b = b - d;
    const double s0a = c - 2.0 * d + e, s0b = 3.0 * c - 4.0 * d + e;
    const double b2 = (13.0 / 12.0) * (s2a * s2a) + (1.0 / 4.0) * (s2b * s2b);
    const double b1 = (13.0 / 12.0) * (s1a * s1a) + (1.0 / 4.0) * (s1b * s1b);
    const double b0 = (13.0 / 12.0) * (s0a * s0a) + (1.0 / 4.0) * (s0b * s0b);
    const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
    const double a0 = d0 / (e0 * e0);
    const double a1 = d1 / (e1 * e1);
    const double a2 = d2 / (e2 * e2);
    const double sum = a0 + a1 + a2;
    double U;
    if (sum < eps) {
        U = (v0 + v1 + v2) / 3.0;
    } else {
        const double w0 = a0 / sum, w1 = a1 / sum, w2 = a2 / sum;
        U = w0 * v0 + w1 * v1 + w2 * v2;
    }
    return isfinite(U) ? U : c;
}

// Face state from WENO-reconstructed conserved values (Calculate_Primitive_Variables on U_L /
// U_R, src/WENO2D.cpp:482-489); a non-positive density or pressure falls back to the cell state.
__device__ __forceinline__ void weno_face_state(const double *U, const FState &cell, FState &out) {
    state_from_U(U[0], U[1], U[2], U[3], out);
    if (!(out.W[0] > 0.0) || !(out.W[3] > 0.0)) out = cell;
}

// ------------------------------------------------------------------------------------------
// Laminar viscous terms (2-D Navier-Stokes, Stokes hypothesis).
// ------------------------------------------------------------------------------------------
// Sutherland's law constants of CUDA_KERNELS/Viscous_Flux_Cuda_Kernels.cu:6-7 (air, SI units).
constexpr double kSuthC1 = 1.45793265452e-06;
constexpr double kSuthS = 110.4;

// Dynamic viscosity at temperature T: the constant cfg.mu, or Sutherland's law
// C1 T^1.5 / (T + S) (calculate_viscosity, Viscous_Flux_Cuda_Kernels.cu:12-14) with T^1.5 formed
// as T sqrt(T), which is correctly rounded on every target.
__device__ __forceinline__ double viscosity(const DevCfg &cfg, double T) {
    if (cfg.viscous == 2) return kSuthC1 * (T * sqrt(T)) / (T + kSuthS);
    return cfg.mu;
}

// Face gradient of one variable: the mean of the two cells' gradients with its component along
// the centre-to-centre direction t replaced by the difference quotient (phi_R - phi_L) / |d|, so
// that neighbouring values are coupled directly (no odd-even decoupling on the averaged stencil).
__device__ __forceinline__ void face_gradient(double axL, double ayL, double axR, double ayR,
                                              double phiL, double phiR, double inv_dist, double tx,
                                              double ty, double &gx, double &gy) {
    const double ax = 0.5 * (axL + axR);
    const double ay = 0.5 * (ayL + ayR);
    const double corr = (phiR - phiL) * inv_dist - (ax * tx + ay * ty);
    gx = ax + corr * tx;
    gy = ay + corr * ty;
}

// Viscous flux through a face times its length, from the cell-centre primitives WL, WR (WR the
// ghost state on a boundary face) and the cells' gradients gL, gR of (rho, u, v, p) (gR = gL on a
// boundary face); (dx, dy) = x_R - x_L. Stress and heat flux as Viscous_Flux_on_Face
// (src/Grid_Refine_Functions.cpp:418-479): tau_xx = 2/3 mu (2 u_x - v_y), tau_xy = mu (u_y + v_x),
// tau_yy = 2/3 mu (2 v_y - u_x), q = k grad T, face velocity and temperature the two cells' mean;
// F_v = (0, tau_xx nx + tau_xy ny, tau_xy nx + tau_yy ny, (tau_x. u + q_x) nx + (tau_y. u + q_y) ny) dl.
// Temperatures are T = p / (rho r_gas), grad T = (grad p - (p / rho) grad rho) / (rho r_gas).
__device__ __forceinline__ void viscous_flux(const DevCfg &cfg, const double *WL, const double *WR,
                                             const double *gL, const double *gR, double nx,
                                             double ny, double dl, double dx, double dy,
                                             double *Fv) {
    const double rgL = WL[0] * cfg.r_gas, rgR = WR[0] * cfg.r_gas;
    const double TL = WL[3] / rgL, TR = WR[3] / rgR;
    const double qL = WL[3] / WL[0], qR = WR[3] / WR[0];
    const double TxL = (gL[6] - qL * gL[0]) / rgL, TyL = (gL[7] - qL * gL[1]) / rgL;
    const double TxR = (gR[6] - qR * gR[0]) / rgR, TyR = (gR[7] - qR * gR[1]) / rgR;
    const double dist = sqrt(dx * dx + dy * dy);
    const double inv_dist = 1.0 / dist;
    const double tx = dx * inv_dist, ty = dy * inv_dist;
    double ux, uy, vx, vy, Tx, Ty;
    face_gradient(gL[2], gL[3], gR[2], gR[3], WL[1], WR[1], inv_dist, tx, ty, ux, uy);
    face_gradient(gL[4], gL[5], gR[4], gR[5], WL[2], WR[2], inv_dist, tx, ty, vx, vy);
    face_gradient(TxL, TyL, TxR, TyR, TL, TR, inv_dist, tx, ty, Tx, Ty);
    const double mu = viscosity(cfg, 0.5 * (TL + TR));
    const double k = mu * cfg.kfac;
    const double t11 = (2.0 / 3.0) * mu * (2.0 * ux - vy);
    const double t12 = mu * (uy + vx);
    const double t22 = (2.0 / 3.0) * mu * (2.0 * vy - ux);
    const double uf = 0.5 * (WL[1] + WR[1]), vf = 0.5 * (WL[2] + WR[2]);
    const double Qx = k * Tx, Qy = k * Ty;
    Fv[0] = 0.0;
    Fv[1] = (t11 * nx + t12 * ny) * dl;
    Fv[2] = (t12 * nx + t22 * ny) * dl;
    Fv[3] = ((t11 * uf + t12 * vf + Qx) * nx * dl + (t12 * uf + t22 * vf + Qy) * ny * dl);
}

// Viscous part of a cell's time-step denominator, Viscous_Time_Step_1
// (src/Time_Step.cpp:103-109) with the cell's own density and viscosity:
// 4 (1/A) max(4 / (3 rho), gamma / rho) (mu / Pr) sum_f dl_f^2; dt = CFL A / (inviscid + this).
__device__ __forceinline__ double viscous_dt_term(const DevCfg &cfg, double rho, double p,
                                                  double sdl2, double area) {
    const double mu = viscosity(cfg, p / (rho * cfg.r_gas));
    return 4.0 * ((fmax(4.0 / (3.0 * rho), kGamma / rho) * mu * cfg.inv_pr * sdl2) / area);
}

}  // namespace cfd
