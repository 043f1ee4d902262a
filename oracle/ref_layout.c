/* CPU restatement of the reference-layout (launch_*) kernels, cell by cell.
 *
 * TEST INFRASTRUCTURE ONLY (see cfd_oracle.h). Layouts follow CUDA_KERNELS/ of
 * rameshkolluru43/CFD_Solver_withCUDA: U[N*4], P[N*6] = (rho,u,v,P,T,a), nb[N*4], nrm[N*8],
 * areas[N*4], outputs [(N*4+f)*4+var]. Formulas follow:
 *   Roe_Flux_Cuda_Kernels.cu:28-57,77-410, HLLC_LLF_Flux_Cuda_Kernels.cu:25-301,
 *   Limiter_Cuda_Kernels.cu:89-489, MUSCL_WENO_Reconstruction_Cuda_Kernels.cu:29-83,
 *   Boundary_Conditions_Cuda_Kernels.cu:30-324, Cuda_Kernels.h:4-33,
 * with the corrections documented in the HIP file: limiter slope ratio from the opposite face,
 * neighbour limiter values taken from the face pointing back to the cell.
 * Parity pins (tests/test_oracle_ref_layout.py): HLLC and LLF equal the native face fluxes; the
 * Roe kernel text halves the dissipation (0.5 (F_L + F_R - D) with D = 0.5 sum |lambda| alpha R,
 * Roe_Flux_Cuda_Kernels.cu:197-238) and equals the mean of the native Roe and central fluxes;
 * limiter functions at r in {-1,0,0.5,1,2,3}; limiters, MUSCL and Barth-Jespersen on linear
 * data; Venkatakrishnan's limits; boundary-state invariants; RK stage orders 3 and 4; the
 * second-order Roe flux's first-order and uniform-state limits.
 */
#include <math.h>
#include <stddef.h>

#define G 1.4
#define GM1 (G - 1.0)

static double entropy_fix(double lambda, double a) {
    double delta = 0.1 * a;
    if (fabs(lambda) < delta) return 0.5 * (lambda * lambda / delta + delta);
    return fabs(lambda);
}

static void roe_avg(double rho_L, double u_L, double v_L, double H_L, double rho_R, double u_R,
                    double v_R, double H_R, double gamma, double *ru, double *rv, double *rH,
                    double *ra) {
    double sL = sqrt(rho_L), sR = sqrt(rho_R);
    double term = 1.0 / (sL + sR);
    *ru = term * (u_L * sL + u_R * sR);
    *rv = term * (v_L * sL + v_R * sR);
    *rH = term * (H_L * sL + H_R * sR);
    double vm = 0.5 * (*ru * *ru + *rv * *rv);
    *ra = sqrt(fmax((gamma - 1.0) * (*rH - vm), 1e-14));
}

static void roe_flux(double rho_L, double u_L, double v_L, double P_L, double E_L, double H_L,
                     double rho_R, double u_R, double v_R, double P_R, double E_R, double H_R,
                     double nx, double ny, double gamma, double area, double *out) {
    double roe_u, roe_v, roe_H, roe_a;
    roe_avg(rho_L, u_L, v_L, H_L, rho_R, u_R, v_R, H_R, gamma, &roe_u, &roe_v, &roe_H, &roe_a);
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
    double l1 = entropy_fix(Un - roe_a, roe_a), l2 = fabs(Un), l3 = fabs(Un);
    double l4 = entropy_fix(Un + roe_a, roe_a);
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

static int unit_normal(const double *nrm, int i, int f, double *nx, double *ny) {
    *nx = nrm[i * 8 + f * 2];
    *ny = nrm[i * 8 + f * 2 + 1];
    double mag = sqrt(*nx * *nx + *ny * *ny);
    if (mag < 1e-14) return 0;
    *nx /= mag;
    *ny /= mag;
    return 1;
}

static int back_face(const int *nb, int j, int i) {
    for (int g = 0; g < 4; g++)
        if (nb[j * 4 + g] == i) return g;
    return -1;
}

/* scheme: 0 Roe, 1 HLLC, 2 LLF */
void rl_flux(int scheme, const double *U, const double *P, const int *nb, const double *nrm,
             const double *ar, double *out, int N, double gamma) {
    for (int i = 0; i < N; i++)
        for (int f = 0; f < 4; f++) {
            int j = nb[i * 4 + f];
            if (j < 0 || j >= N) continue;
            double nx, ny;
            if (!unit_normal(nrm, i, f, &nx, &ny)) continue;
            double area = ar[i * 4 + f];
            double rL = P[i * 6], uL = P[i * 6 + 1], vL = P[i * 6 + 2], pL = P[i * 6 + 3], aL = P[i * 6 + 5];
            if (rL <= 0.0 || pL <= 0.0 || aL <= 0.0) continue;
            double rR = P[j * 6], uR = P[j * 6 + 1], vR = P[j * 6 + 2], pR = P[j * 6 + 3], aR = P[j * 6 + 5];
            if (rR <= 0.0 || pR <= 0.0 || aR <= 0.0) continue;
            double *o = &out[(i * 4 + f) * 4];
            if (scheme == 0) {
                double HL = (aL * aL) / (gamma - 1.0) + 0.5 * (uL * uL + vL * vL);
                double HR = (aR * aR) / (gamma - 1.0) + 0.5 * (uR * uR + vR * vR);
                roe_flux(rL, uL, vL, pL, U[i * 4 + 3], HL, rR, uR, vR, pR, U[j * 4 + 3], HR, nx, ny, gamma, area, o);
            } else if (scheme == 2) {
                double unL = nx * uL + ny * vL, unR = nx * uR + ny * vR;
                double lmax = fmax(fabs(unL) + aL, fabs(unR) + aR);
                double FL[4] = {rL * unL, rL * uL * unL + pL * nx, rL * vL * unL + pL * ny, (U[i * 4 + 3] + pL) * unL};
                double FR[4] = {rR * unR, rR * uR * unR + pR * nx, rR * vR * unR + pR * ny, (U[j * 4 + 3] + pR) * unR};
                for (int q = 0; q < 4; q++)
                    o[q] = 0.5 * (FL[q] + FR[q] - lmax * (U[j * 4 + q] - U[i * 4 + q])) * area;
            } else {
                double un_L = nx * uL + ny * vL, un_R = nx * uR + ny * vR;
                double sqL = sqrt(rL), sqR = sqrt(rR);
                double a_roe = (aL * sqL + aR * sqR) / (sqL + sqR);
                double u_roe = (uL * sqL + uR * sqR) / (sqL + sqR);
                double v_roe = (vL * sqL + vR * sqR) / (sqL + sqR);
                double un_roe = nx * u_roe + ny * v_roe;
                double S_L = fmin(un_L - aL, un_roe - a_roe);
                double S_R = fmax(un_R + aR, un_roe + a_roe);
                double numer = pR - pL + rL * un_L * (S_L - un_L) - rR * un_R * (S_R - un_R);
                double denom = rL * (S_L - un_L) - rR * (S_R - un_R);
                double S_star = (fabs(denom) > 1e-14) ? numer / denom : 0.5 * (un_L + un_R);
                double fl[4];
                if (S_L >= 0.0) {
                    fl[0] = rL * un_L; fl[1] = rL * uL * un_L + pL * nx; fl[2] = rL * vL * un_L + pL * ny;
                    fl[3] = (U[i * 4 + 3] + pL) * un_L;
                } else if (S_R <= 0.0) {
                    fl[0] = rR * un_R; fl[1] = rR * uR * un_R + pR * nx; fl[2] = rR * vR * un_R + pR * ny;
                    fl[3] = (U[j * 4 + 3] + pR) * un_R;
                } else {
                    int left = S_star >= 0.0;
                    int c = left ? i : j;
                    double S_K = left ? S_L : S_R;
                    double rho = P[c * 6], u = P[c * 6 + 1], v = P[c * 6 + 2], Pp = P[c * 6 + 3], E = U[c * 4 + 3];
                    double un = nx * u + ny * v;
                    double factor = rho * (S_K - un) / (S_K - S_star);
                    double Us0 = factor;
                    double Us1 = factor * (u + (S_star - un) * nx);
                    double Us2 = factor * (v + (S_star - un) * ny);
                    double Us3 = factor * (E / rho + (S_star - un) * (S_star + Pp / (rho * (S_K - un))));
                    fl[0] = rho * un + S_K * (Us0 - U[c * 4]);
                    fl[1] = (rho * u * un + Pp * nx) + S_K * (Us1 - U[c * 4 + 1]);
                    fl[2] = (rho * v * un + Pp * ny) + S_K * (Us2 - U[c * 4 + 2]);
                    fl[3] = ((E + Pp) * un) + S_K * (Us3 - U[c * 4 + 3]);
                }
                for (int q = 0; q < 4; q++) o[q] = fl[q] * area;
            }
        }
}

void rl_roe2(const double *U, const int *nb, const double *nrm, const double *ar,
             const double *phi, const double *ctr, const double *fctr, double *out, int N,
             double gamma) {
    for (int i = 0; i < N; i++) {
        double xc = ctr[i * 2], yc = ctr[i * 2 + 1];
        for (int f = 0; f < 4; f++) {
            int j = nb[i * 4 + f];
            if (j < 0 || j >= N) continue;
            double nx, ny;
            if (!unit_normal(nrm, i, f, &nx, &ny)) continue;
            double area = ar[i * 4 + f];
            double xf = fctr[(i * 4 + f) * 2], yf = fctr[(i * 4 + f) * 2 + 1];
            double dL = sqrt((xf - xc) * (xf - xc) + (yf - yc) * (yf - yc));
            double xr = ctr[j * 2], yr = ctr[j * 2 + 1];
            double dR = sqrt((xf - xr) * (xf - xr) + (yf - yr) * (yf - yr));
            double wL = (dL + dR > 0.0) ? dL / (dL + dR) : 0.5;
            double wR = 1.0 - wL;
            int g = back_face(nb, j, i);
            double UL[4], UR[4];
            for (int q = 0; q < 4; q++) {
                double d = U[j * 4 + q] - U[i * 4 + q];
                double pL = phi[(i * 4 + f) * 4 + q];
                double pR = (g >= 0) ? phi[(j * 4 + g) * 4 + q] : 0.0;
                UL[q] = U[i * 4 + q] + pL * d * wL;
                UR[q] = U[j * 4 + q] - pR * d * wR;
            }
            double rho_L = fmax(UL[0], 1e-14);
            double u_L = UL[1] / rho_L, v_L = UL[2] / rho_L, E_L = UL[3];
            double P_L = fmax((gamma - 1.0) * (E_L - 0.5 * rho_L * (u_L * u_L + v_L * v_L)), 1e-14);
            double H_L = (E_L + P_L) / rho_L;
            double rho_R = fmax(UR[0], 1e-14);
            double u_R = UR[1] / rho_R, v_R = UR[2] / rho_R, E_R = UR[3];
            double P_R = fmax((gamma - 1.0) * (E_R - 0.5 * rho_R * (u_R * u_R + v_R * v_R)), 1e-14);
            double H_R = (E_R + P_R) / rho_R;
            roe_flux(rho_L, u_L, v_L, P_L, E_L, H_L, rho_R, u_R, v_R, P_R, E_R, H_R, nx, ny, gamma, area,
                     &out[(i * 4 + f) * 4]);
        }
    }
}

/* phi(r) of the limiter functions, a = backward slope, b = forward slope; type 0 minmod
 * (Phi macro, include/definitions.h:74-75), 1 Van Leer, 2 superbee, 3 Van Albada */
double rl_phi(int type, double a, double b) {
    if (fabs(b) < 1e-12) return 0.0;
    double r = a / b;
    switch (type) {
    case 1: return (r <= 0.0) ? 0.0 : 2.0 * r / (1.0 + r);
    case 2: return (r <= 0.0) ? 0.0 : fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0));
    case 3: return (r <= 0.0) ? 0.0 : (r * r + r) / (r * r + 1.0);
    default: return (r >= 1.0) ? 1.0 : ((r < 1.0) && (r > 0.0)) ? r : 0.0;
    }
}

void rl_limiter(const double *U, const int *nb, const double *dist, double *phi, int N, double zeta,
                int type, int three) {
    for (int i = 0; i < N; i++)
        for (int f = 0; f < 4; f++) {
            int j = nb[i * 4 + f], o = nb[i * 4 + (f + 2) % 4];
            for (int q = 0; q < 4; q++) {
                double v = 0.0;
                if (j >= 0 && j < N) {
                    double fwd = zeta * (U[j * 4 + q] - U[i * 4 + q]) / dist[i * 4 + f];
                    double bwd = (o >= 0 && o < N) ? zeta * (U[i * 4 + q] - U[o * 4 + q]) / dist[i * 4 + (f + 2) % 4] : 0.0;
                    v = rl_phi(type, bwd, fwd);
                    if (three && type == 0 && fabs(fwd) >= 1e-12) {
                        double r = bwd / fwd;
                        double c = 0.5 * (1.0 + r);
                        v = (r > 0.0) ? fmin(fmin(r, 1.0), c) : 0.0;
                    }
                }
                phi[(i * 4 + f) * 4 + q] = v;
            }
        }
}

void rl_venkat(const double *U, const int *nb, const double *vol, double *phi, int N, double K) {
    for (int i = 0; i < N; i++) {
        double dj = sqrt(vol[i]);
        for (int f = 0; f < 4; f++) {
            int j = nb[i * 4 + f], o = nb[i * 4 + (f + 2) % 4];
            for (int q = 0; q < 4; q++) {
                double v = 1.0;
                if (j >= 0 && j < N) {
                    double dp = U[j * 4 + q] - U[i * 4 + q];
                    double dm = (o >= 0 && o < N) ? U[i * 4 + q] - U[o * 4 + q] : 0.0;
                    if (fabs(dp) >= 1e-12) {
                        double eps2 = (K * dj) * (K * dj) * (K * dj);
                        double num = (dp * dp + eps2) * dm + 2.0 * dm * dm * dp;
                        double den = dp * dp + 2.0 * dm * dm + dp * dm + eps2;
                        v = (fabs(den) < 1e-12) ? 1.0 : num / den;
                    }
                }
                phi[(i * 4 + f) * 4 + q] = v;
            }
        }
    }
}

void rl_bj(const double *U, const int *nb, const double *dist, const double *unlim, double *phi, int N) {
    for (int i = 0; i < N; i++)
        for (int q = 0; q < 4; q++) {
            double uc = U[i * 4 + q], umax = uc, umin = uc;
            for (int f = 0; f < 4; f++) {
                int j = nb[i * 4 + f];
                if (j >= 0 && j < N) {
                    umax = fmax(umax, U[j * 4 + q]);
                    umin = fmin(umin, U[j * 4 + q]);
                }
            }
            for (int f = 0; f < 4; f++) {
                int j = nb[i * 4 + f];
                double v = 0.0;
                if (j >= 0) {
                    double du = unlim[(i * 4 + f) * 4 + q] * dist[i * 4 + f];
                    if (fabs(du) < 1e-12) v = 1.0;
                    else {
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

void rl_muscl(const double *U, const int *nb, const double *dist, const double *phi, double *QL,
              double *QR, int N, double kappa) {
    for (int i = 0; i < N; i++)
        for (int f = 0; f < 4; f++) {
            int j = nb[i * 4 + f];
            if (j < 0 || j >= N) {
                for (int q = 0; q < 4; q++) {
                    QL[(i * 4 + f) * 4 + q] = U[i * 4 + q];
                    QR[(i * 4 + f) * 4 + q] = U[i * 4 + q];
                }
                continue;
            }
            double d = dist[i * 4 + f];
            int g = back_face(nb, j, i);
            for (int q = 0; q < 4; q++) {
                double Ui = U[i * 4 + q], Uj = U[j * 4 + q];
                double pL = phi[(i * 4 + f) * 4 + q];
                double pR = (g >= 0) ? phi[(j * 4 + g) * 4 + q] : 0.0;
                QL[(i * 4 + f) * 4 + q] = Ui + 0.5 * kappa * pL * ((Uj - Ui) / d) * d;
                QR[(i * 4 + f) * 4 + q] = Uj + 0.5 * kappa * pR * ((Ui - Uj) / d) * d;
            }
        }
}

static void bc_prim(const double *U, double *rho, double *u, double *v, double *p) {
    *rho = U[0];
    if (*rho < 1e-12) *rho = 1e-12;
    *u = U[1] / *rho;
    *v = U[2] / *rho;
    *p = GM1 * (U[3] - 0.5 * *rho * (*u * *u + *v * *v));
    if (*p < 1e-12) *p = 1e-12;
}

static void bc_store(double *U, double rho, double u, double v, double p) {
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = p / GM1 + 0.5 * rho * (u * u + v * v);
}

/* kind: 0 subsonic inlet, 1 supersonic inlet, 2 subsonic exit, 3 supersonic exit, 4 viscous
 * wall, 5 inviscid wall, 6 symmetry, 7 far field; bc = (rho, u, v, p) */
void rl_bc(int kind, double *Uc, const int *list, const double *nrm, const double *bc, int n) {
    for (int idx = 0; idx < n; idx++) {
        int cell = list[idx];
        double *U = &Uc[cell * 4];
        double rho, u, v, p;
        switch (kind) {
        case 0: bc_prim(U, &rho, &u, &v, &p); bc_store(U, bc[0], u, v, bc[3]); break;
        case 1: bc_store(U, bc[0], bc[1], bc[2], bc[3]); break;
        case 2: bc_prim(U, &rho, &u, &v, &p); bc_store(U, rho, u, v, bc[3]); break;
        case 3: break;
        case 4: U[1] = -U[1]; U[2] = -U[2]; break;
        case 5:
        case 6: {
            rho = U[0];
            if (rho < 1e-12) rho = 1e-12;
            u = U[1] / rho;
            v = U[2] / rho;
            double nx = nrm[cell * 8], ny = nrm[cell * 8 + 1];
            double vn = u * nx + v * ny;
            U[1] = rho * (u - 2.0 * vn * nx);
            U[2] = rho * (v - 2.0 * vn * ny);
            break;
        }
        default: {
            bc_prim(U, &rho, &u, &v, &p);
            double nx = nrm[cell * 8], ny = nrm[cell * 8 + 1];
            double ai = sqrt(G * p / rho), af = sqrt(G * bc[3] / bc[0]);
            double vni = u * nx + v * ny, vnf = bc[1] * nx + bc[2] * ny;
            double Rp = vni + 2.0 * ai / GM1, Rm = vnf - 2.0 * af / GM1;
            double vnb = 0.5 * (Rp + Rm);
            double ab = 0.25 * GM1 * (Rp - Rm);
            if (ab < 1e-12) ab = 1e-12;
            double rb, ub, vb, pb;
            if (vnb >= 0.0) {
                double s = p / pow(rho, G);
                rb = pow(ab * ab / (G * s), 1.0 / GM1);
                ub = u + (vnb - vni) * nx;
                vb = v + (vnb - vni) * ny;
            } else {
                double s = bc[3] / pow(bc[0], G);
                rb = pow(ab * ab / (G * s), 1.0 / GM1);
                ub = bc[1] + (vnb - vnf) * nx;
                vb = bc[2] + (vnb - vnf) * ny;
            }
            pb = rb * ab * ab / G;
            bc_store(U, rb, ub, vb, pb);
            break;
        }
        }
    }
}

void rl_rk3_stage(const double *U, const double *R, const double *Rv, double *Urk,
                  const double *ia, double dt, int N, int stage) {
    for (int i = 0; i < N; i++)
        for (int q = 0; q < 4; q++) {
            double r = R[i * 4 + q] - (Rv ? Rv[i * 4 + q] : 0.0);
            double u = U[i * 4 + q];
            if (stage == 1) Urk[i * 4 + q] = u - dt * ia[i] * r;
            else if (stage == 2) Urk[i * 4 + q] = 0.25 * (Urk[i * 4 + q] + 3.0 * u - dt * ia[i] * r);
            else if (stage == 3) Urk[i * 4 + q] = (2.0 / 3.0) * (Urk[i * 4 + q] - u - dt * ia[i] * r);
        }
}

void rl_rk4_stage(const double *U, const double *R, double *Us, double *K, const double *ia,
                  double dt, int N, int stage) {
    for (int i = 0; i < N; i++) {
        double h = dt * ia[i];
        for (int q = 0; q < 4; q++) {
            int e = i * 4 + q;
            switch (stage) {
            case 1: K[e] = R[e]; Us[e] = U[e] - 0.5 * h * R[e]; break;
            case 2: K[e] += 2.0 * R[e]; Us[e] = U[e] - 0.5 * h * R[e]; break;
            case 3: K[e] += 2.0 * R[e]; Us[e] = U[e] - h * R[e]; break;
            default: Us[e] = U[e] - (h / 6.0) * (K[e] + R[e]); break;
            }
        }
    }
}
