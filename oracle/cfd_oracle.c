/* CPU FP64 oracle for the explicit finite-volume residual-and-update path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP implementation. Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the product path.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (build_oracle() in oracle/oracle.py). Contraction is off
 * so that expressions round exactly as the reference's g++ x86-64 build does.
 *
 * Parity pins (tests/test_oracle_golden.py against the reference, tests/test_oracle_math.py
 * against closed forms and invariants):
 *   om_build/om_export ........ bit-exact vs the reference build (oracle/_ref) and its golden
 *                               fixtures in tests/golden (Sod, jittered mixed meshes, the
 *                               reference's own Ramp_15o_52_18 grid); closed forms (shoelace,
 *                               unit normals, sum_f n dl = 0, neighbour symmetry).
 *   or_prim ................... bit-exact vs fixtures; closed form round trip.
 *   on_ghost_state (ref BCs) .. bit-exact vs the reference's ghost arrays (Apply_Boundary_Conditions).
 *   on_residual Roe/LLF order 1 vs Evaluate_Cell_Net_Flux_1O (<=1e-13 rel; the face loop
 *                               reassociates the two-sided sums); D = 0 for equal states.
 *   on_refstep ................ reference iteration vs Inviscid_Solver steps (Sod, <=1e-12).
 *   or_gg_gradient ............ bit-exact vs Calculate_Gradient_At_Cell_Center.
 *   or_ausm_cell .............. bit-exact vs Ausm_Flux.
 *   or_roe2o_residual ......... bit-exact vs Evaluate_Cell_Net_Flux_2O (ROE_2O).
 *   HLLC (no CPU reference) ... consistency F(W,W) = F(W).n, supersonic upwinding, stationary
 *                               contact, rotational invariance, antisymmetry; Sod, Lax and
 *                               Toro test 3 tubes against Toro's exact Riemann solution.
 *   Roe / LLF / AUSM .......... the same flux properties (where the scheme has them).
 *   limiters .................. closed forms of phi(r) at r in {-1,0,0.25,0.5,1,2,3,10}, TVD
 *                               region, symmetry phi(r)/r = phi(1/r).
 *   MUSCL (on_residual order 2) both sides reconstruct the same state for linear fields, so the
 *                               dissipation of Roe, LLF and HLLC vanishes; MUSCL beats first
 *                               order on the shock tubes; freestream preservation.
 *   WLSQ / GG ................. exactness for linear fields (WLSQ on jittered mixed meshes, GG on
 *                               uniform quads); brute-force normal equations on a tiny mesh.
 *   or_rk_stage ............... convergence order 3 (TVD-RK3) and 4 (RK4) on a linear
 *                               oscillator; RK4 and RK3 agree on Sod.
 *   far field / freestream .... Riemann invariants, entropy and tangential velocity of the
 *                               upwind side; supersonic limits; freestream is a fixed point.
 */
#include "cfd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants formed as the reference forms them (include/definitions.h:62-70). */
#define GAMMA 1.4
#define GM1 (GAMMA - 1.0)
#define GAMMA1 (GAMMA / GM1)
#define R_GC 287.0
#define EPS_BBOX 1e-6

enum { KIND_INTERNAL = 0, KIND_LEFT = 1, KIND_RIGHT = 2, KIND_TOP = 3, KIND_BOTTOM = 4, KIND_WALL = 5 };

struct OMesh {
    int np;
    int nc, ng;
    int *nf, *nodes, *nbr, *kind; /* [4*nc] */
    double *vtx;                  /* [12*nc] sorted vertices */
    double *nrm, *len, *dist;     /* [8*nc], [4*nc], [4*nc] */
    double *ctr, *area, *inv_area; /* [3*(nc+ng)], [nc+ng], [nc+ng] */
    int *inl, *exl, *wal;
    int nin, nex, nwall;
    int *gcell, *gface, *gkind; /* per ghost: owner cell, owner face slot, face kind */
    /* face list for the face loop */
    int nfaces;
    int *fL, *fR, *fslot;  /* per face: L cell, R cell (or ghost id), slot in L */
    int *cface, *csign;    /* [4*nc]: face of slot, +1 if the cell is L else -1 */
};

/* ------------------------------------------------------------------------------------------
 * Mesh: Read_VTK_Grid + Form_Cells restated.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    int a, b, cell, slot;
} EdgeRec;

static int edge_cmp(const void *x, const void *y) {
    const EdgeRec *p = (const EdgeRec *)x, *q = (const EdgeRec *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
    return p->slot < q->slot ? -1 : (p->slot > q->slot);
}

/* Sort_Points_AntiClockWise(V, idx) (src/Grid_Computations.cpp:1092-1142): order by atan2
 * about the vertex average, with |dAngle| < 1e-9 compared by z. Sequences of <= 16 go through
 * std::sort's insertion sort, i.e. a stable insertion sort with this comparator. */
static int ccw_less(const double *A, const double *B, double cx, double cy) {
    double aA = atan2(A[1] - cy, A[0] - cx);
    double aB = atan2(B[1] - cy, B[0] - cx);
    if (fabs(aA - aB) < 1e-9) return A[2] < B[2];
    return aA < aB;
}

static void sort_ccw(int n, double *v /*3n*/, int *idx) {
    double x = 0.0, y = 0.0, z = 0.0;
    for (int i = 0; i < n; i++) {
        x += v[3 * i];
        y += v[3 * i + 1];
        z += v[3 * i + 2];
    }
    double cx = x / n, cy = y / n;
    (void)z;
    int sorted = 1;
    for (int i = 1; i < n; i++) {
        double ap = atan2(v[3 * (i - 1) + 1] - cy, v[3 * (i - 1)] - cx);
        double ac = atan2(v[3 * i + 1] - cy, v[3 * i] - cx);
        if (ap > ac) {
            sorted = 0;
            break;
        }
    }
    if (sorted) return;
    for (int i = 1; i < n; i++) {
        double tv[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        int ti = idx[i];
        int j = i;
        while (j > 0 && ccw_less(tv, &v[3 * (j - 1)], cx, cy)) {
            v[3 * j] = v[3 * (j - 1)];
            v[3 * j + 1] = v[3 * (j - 1) + 1];
            v[3 * j + 2] = v[3 * (j - 1) + 2];
            idx[j] = idx[j - 1];
            j--;
        }
        v[3 * j] = tv[0];
        v[3 * j + 1] = tv[1];
        v[3 * j + 2] = tv[2];
        idx[j] = ti;
    }
}

OMesh *om_build(int npts, const double *pts, int ncells, const int *cells4) {
    OMesh *m = (OMesh *)calloc(1, sizeof(OMesh));
    const int nc = ncells;
    m->np = npts;
    m->nc = nc;
    m->nf = (int *)calloc(nc, sizeof(int));
    m->nodes = (int *)malloc(sizeof(int) * 4 * nc);
    m->nbr = (int *)malloc(sizeof(int) * 4 * nc);
    m->kind = (int *)calloc(4 * (size_t)nc, sizeof(int));
    m->vtx = (double *)calloc(12 * (size_t)nc, sizeof(double));
    m->nrm = (double *)calloc(8 * (size_t)nc, sizeof(double));
    m->len = (double *)calloc(4 * (size_t)nc, sizeof(double));
    m->dist = (double *)calloc(4 * (size_t)nc, sizeof(double));
    double *area_p = (double *)malloc(sizeof(double) * nc);
    double *inv_p = (double *)malloc(sizeof(double) * nc);
    double *ctr_p = (double *)malloc(sizeof(double) * 3 * nc);
    for (int i = 0; i < 4 * nc; i++) m->nodes[i] = m->nbr[i] = -1;

    /* Identify_Cells: vertices, CCW sort, Construct_Cell (src/Grid_Computations.cpp:523-565) */
    for (int c = 0; c < nc; c++) {
        int n = cells4[4 * c + 3] < 0 ? 3 : 4;
        m->nf[c] = n;
        double *v = &m->vtx[12 * c];
        int *id = &m->nodes[4 * c];
        for (int i = 0; i < n; i++) {
            id[i] = cells4[4 * c + i];
            v[3 * i] = pts[3 * id[i]];
            v[3 * i + 1] = pts[3 * id[i] + 1];
            v[3 * i + 2] = pts[3 * id[i] + 2];
        }
        sort_ccw(n, v, id);
        /* Compute_Centroid (:957-980): sequential sums / n */
        double x = 0.0, y = 0.0, z = 0.0;
        for (int i = 0; i < n; i++) {
            x += v[3 * i];
            y += v[3 * i + 1];
            z += v[3 * i + 2];
        }
        ctr_p[3 * c] = x / n;
        ctr_p[3 * c + 1] = y / n;
        ctr_p[3 * c + 2] = z / n;
        /* Construct_Face (:691-752) */
        for (int f = 0; f < n; f++) {
            const double *a = &v[3 * f], *b = &v[3 * ((f + 1) % n)];
            double dy = b[1] - a[1];
            double dx = b[0] - a[0];
            double dL = sqrt(dx * dx + dy * dy);
            m->len[4 * c + f] = dL;
            m->nrm[8 * c + 2 * f] = dy / dL;
            m->nrm[8 * c + 2 * f + 1] = -dx / dL;
        }
        /* shoelace (:541-557) */
        double twice = 0.0;
        for (int i = 0; i < n; i++) {
            int j = (i + 1) % n;
            twice += (v[3 * i] * v[3 * j + 1] - v[3 * j] * v[3 * i + 1]);
        }
        area_p[c] = 0.5 * fabs(twice);
        inv_p[c] = 1.0 / area_p[c];
    }

    /* mapFacesToCells + fillFaceOrderedNeighbours (src/Read_Gmsh_File.cpp:325-402) */
    size_t ne = 0;
    for (int c = 0; c < nc; c++) ne += m->nf[c];
    EdgeRec *E = (EdgeRec *)malloc(sizeof(EdgeRec) * ne);
    size_t k = 0;
    for (int c = 0; c < nc; c++) {
        int n = m->nf[c];
        for (int f = 0; f < n; f++) {
            int a = m->nodes[4 * c + f], b = m->nodes[4 * c + (f + 1) % n];
            if (a > b) {
                int t = a;
                a = b;
                b = t;
            }
            E[k].a = a;
            E[k].b = b;
            E[k].cell = c;
            E[k].slot = f;
            k++;
        }
    }
    qsort(E, ne, sizeof(EdgeRec), edge_cmp);
    for (size_t i = 0; i < ne;) {
        size_t j = i;
        while (j < ne && E[j].a == E[i].a && E[j].b == E[i].b) j++;
        /* distinct cells sharing the edge (a std::set in the reference) */
        int distinct = 0, c0 = -1, c1 = -1;
        for (size_t q = i; q < j; q++) {
            if (E[q].cell != c0 && E[q].cell != c1) {
                if (distinct == 0) c0 = E[q].cell;
                else if (distinct == 1) c1 = E[q].cell;
                distinct++;
            }
        }
        if (distinct == 2) {
            for (size_t q = i; q < j; q++) {
                int other = (E[q].cell == c0) ? c1 : c0;
                m->nbr[4 * E[q].cell + E[q].slot] = other;
            }
        }
        i = j;
    }
    free(E);
    int ghost = nc;
    int *hasb = (int *)calloc(nc, sizeof(int));
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m->nf[c]; f++)
            if (m->nbr[4 * c + f] < 0) {
                m->nbr[4 * c + f] = ghost++;
                hasb[c] = 1;
            }
    const int ng = ghost - nc;
    m->ng = ng;

    /* BoundingBox (:441-472) over all points */
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    for (int i = 0; i < npts; i++) {
        double x = pts[3 * i], y = pts[3 * i + 1];
        if (i == 0) {
            xmin = xmax = x;
            ymin = ymax = y;
        } else {
            if (x < xmin) xmin = x;
            if (x > xmax) xmax = x;
            if (y < ymin) ymin = y;
            if (y > ymax) ymax = y;
        }
    }
    /* Classify_Domain_Boundaries + Get_Boundary_Type (:490-577): every edge of a cell that
     * has a boundary face is classified, interior edges of such cells included. */
    for (int c = 0; c < nc; c++) {
        int n = m->nf[c];
        for (int e = 0; e < n; e++) {
            const double *p1 = &m->vtx[12 * c + 3 * e], *p2 = &m->vtx[12 * c + 3 * ((e + 1) % n)];
            int kd = KIND_INTERNAL;
            if (hasb[c]) {
                if (fabs(p1[0] - xmin) < EPS_BBOX && fabs(p2[0] - xmin) < EPS_BBOX) kd = KIND_LEFT;
                else if (fabs(p1[0] - xmax) < EPS_BBOX && fabs(p2[0] - xmax) < EPS_BBOX) kd = KIND_RIGHT;
                else if (fabs(p1[1] - ymin) < EPS_BBOX && fabs(p2[1] - ymin) < EPS_BBOX) kd = KIND_BOTTOM;
                else if (fabs(p1[1] - ymax) < EPS_BBOX && fabs(p2[1] - ymax) < EPS_BBOX) kd = KIND_TOP;
                else kd = KIND_WALL;
            }
            m->kind[4 * c + e] = kd;
        }
    }
    free(hasb);
    /* Create_Boundary_Cells_Lists (:579-621) */
    m->inl = (int *)malloc(sizeof(int) * 3 * (ng + 1));
    m->exl = (int *)malloc(sizeof(int) * 3 * (ng + 1));
    m->wal = (int *)malloc(sizeof(int) * 3 * (ng + 1));
    m->gcell = (int *)malloc(sizeof(int) * (ng + 1));
    m->gface = (int *)malloc(sizeof(int) * (ng + 1));
    m->gkind = (int *)malloc(sizeof(int) * (ng + 1));
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m->nf[c]; f++) {
            int g = m->nbr[4 * c + f];
            if (g < nc) continue;
            int kd = m->kind[4 * c + f];
            m->gcell[g - nc] = c;
            m->gface[g - nc] = f;
            m->gkind[g - nc] = kd;
            int *L = NULL, *cnt = NULL;
            if (kd == KIND_LEFT) { L = m->inl; cnt = &m->nin; }
            else if (kd == KIND_RIGHT) { L = m->exl; cnt = &m->nex; }
            else if (kd == KIND_TOP || kd == KIND_BOTTOM || kd == KIND_WALL) { L = m->wal; cnt = &m->nwall; }
            if (L) {
                L[3 * *cnt] = c;
                L[3 * *cnt + 1] = f;
                L[3 * *cnt + 2] = g;
                (*cnt)++;
            }
        }
    /* Construct_Ghost_Cells (:394-442, :589-629): centre reflected through the face midpoint */
    const int nt = nc + ng;
    m->ctr = (double *)calloc(3 * (size_t)nt, sizeof(double));
    m->area = (double *)calloc(nt, sizeof(double));
    m->inv_area = (double *)calloc(nt, sizeof(double));
    memcpy(m->ctr, ctr_p, sizeof(double) * 3 * nc);
    memcpy(m->area, area_p, sizeof(double) * nc);
    memcpy(m->inv_area, inv_p, sizeof(double) * nc);
    free(ctr_p);
    free(area_p);
    free(inv_p);
    for (int gi = 0; gi < ng; gi++) {
        int c = m->gcell[gi], f = m->gface[gi], g = nc + gi;
        int nv = m->nf[c];
        const double *a = &m->vtx[12 * c + 3 * (f % nv)], *b = &m->vtx[12 * c + 3 * ((f + 1) % nv)];
        for (int q = 0; q < 3; q++) {
            double mid = 0.5 * (a[q] + b[q]);
            m->ctr[3 * g + q] = m->ctr[3 * c + q] + 2.0 * (mid - m->ctr[3 * c + q]);
        }
        m->area[g] = m->area[c];
        m->inv_area[g] = m->inv_area[c];
    }
    /* Calculate_Cell_Center_Distances (:802-827, Distance_Between_Points :863-885) */
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m->nf[c]; f++) {
            int j = m->nbr[4 * c + f];
            double D = 0.0;
            for (int q = 0; q < 3; q++) {
                double t = m->ctr[3 * j + q] - m->ctr[3 * c + q];
                D += t * t;
            }
            m->dist[4 * c + f] = sqrt(D);
        }
    /* face list for the face loop: interior faces once (from the lower-numbered cell),
     * boundary faces once, in (cell, slot) order */
    m->fL = (int *)malloc(sizeof(int) * 4 * nc);
    m->fR = (int *)malloc(sizeof(int) * 4 * nc);
    m->fslot = (int *)malloc(sizeof(int) * 4 * nc);
    m->cface = (int *)malloc(sizeof(int) * 4 * nc);
    m->csign = (int *)calloc(4 * (size_t)nc, sizeof(int));
    for (int i = 0; i < 4 * nc; i++) m->cface[i] = -1;
    int nfc = 0;
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m->nf[c]; f++) {
            int j = m->nbr[4 * c + f];
            if (j < nc && j < c) continue;
            m->fL[nfc] = c;
            m->fR[nfc] = j;
            m->fslot[nfc] = f;
            m->cface[4 * c + f] = nfc;
            m->csign[4 * c + f] = 1;
            if (j < nc) {
                for (int s = 0; s < m->nf[j]; s++)
                    if (m->nbr[4 * j + s] == c) {
                        m->cface[4 * j + s] = nfc;
                        m->csign[4 * j + s] = -1;
                    }
            }
            nfc++;
        }
    m->nfaces = nfc;
    return m;
}

void om_free(OMesh *m) {
    if (!m) return;
    free(m->nf); free(m->nodes); free(m->nbr); free(m->kind); free(m->vtx);
    free(m->nrm); free(m->len); free(m->dist); free(m->ctr); free(m->area); free(m->inv_area);
    free(m->inl); free(m->exl); free(m->wal); free(m->gcell); free(m->gface); free(m->gkind);
    free(m->fL); free(m->fR); free(m->fslot); free(m->cface); free(m->csign);
    free(m);
}

int om_nfaces(const OMesh *m) { return m->nfaces; }

void om_counts(const OMesh *m, int *nphys, int *nghost, int *n_in, int *n_ex, int *n_wall) {
    *nphys = m->nc;
    *nghost = m->ng;
    *n_in = m->nin;
    *n_ex = m->nex;
    *n_wall = m->nwall;
}

void om_export(const OMesh *m, int *nf, int *nodes4, int *nbr4, double *nrm8, double *len4,
               double *ctr3_all, double *area_all, double *inv_area_all, double *dist4, int *kind4,
               int *inl, int *exl, int *wal) {
    const int nc = m->nc, nt = m->nc + m->ng;
    for (int c = 0; c < nc; c++) {
        nf[c] = m->nf[c];
        for (int f = 0; f < m->nf[c]; f++) {
            nodes4[4 * c + f] = m->nodes[4 * c + f];
            nbr4[4 * c + f] = m->nbr[4 * c + f];
            nrm8[8 * c + 2 * f] = m->nrm[8 * c + 2 * f];
            nrm8[8 * c + 2 * f + 1] = m->nrm[8 * c + 2 * f + 1];
            len4[4 * c + f] = m->len[4 * c + f];
            dist4[4 * c + f] = m->dist[4 * c + f];
            kind4[4 * c + f] = m->kind[4 * c + f];
        }
    }
    memcpy(ctr3_all, m->ctr, sizeof(double) * 3 * nt);
    memcpy(area_all, m->area, sizeof(double) * nt);
    memcpy(inv_area_all, m->inv_area, sizeof(double) * nt);
    memcpy(inl, m->inl, sizeof(int) * 3 * m->nin);
    memcpy(exl, m->exl, sizeof(int) * 3 * m->nex);
    memcpy(wal, m->wal, sizeof(int) * 3 * m->nwall);
}

/* ------------------------------------------------------------------------------------------
 * Primitive recovery (src/Basic_Functions.cpp:85-142) and conserved state
 * ---------------------------------------------------------------------------------------- */
void or_prim(const double *U, double *P) {
    double Rho = U[0];
    double inv = 1.0 / Rho;
    double v1 = U[1] * inv;
    double v2 = U[2] * inv;
    if (fabs(v1) <= 1e-10) v1 = 0.0;
    if (fabs(v2) <= 1e-10) v2 = 0.0;
    double vmag = (v1 * v1 + v2 * v2);
    double Pr = GM1 * (U[3] - 0.5 * Rho * vmag);
    double T = (Pr * inv / R_GC);
    double C = sqrt(GAMMA * Pr * inv);
    double M = sqrt(vmag) / C;
    P[0] = Rho;
    P[1] = v1;
    P[2] = v2;
    P[3] = T;
    P[4] = Pr;
    P[5] = C;
    P[6] = U[3];
    P[7] = M;
    P[8] = 0.0;
    P[9] = 0.0;
    P[10] = Pr * pow((1.0 + 0.5 * GM1 * M * M), GAMMA / GM1);
}

/* A face state: primitives (rho, u, v, p, a) and conserved (rho, rho u, rho v, rho E). */
typedef struct {
    double W[5];
    double U[4];
} FState;

static void state_from_U(const double *U, FState *s) {
    double P[11];
    or_prim(U, P);
    s->W[0] = P[0];
    s->W[1] = P[1];
    s->W[2] = P[2];
    s->W[3] = P[4];
    s->W[4] = P[5];
    for (int k = 0; k < 4; k++) s->U[k] = U[k];
}

/* conserved variables from (p, V, rho): Calculate_Computational_Variables case 2
 * (src/Basic_Functions.cpp:171-206) */
static void cons_from_prim(double rho, double u, double v, double p, double *U) {
    double vmag = 0.5 * (u * u + v * v);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = ((p / GM1) + rho * vmag);
}

/* Density and pressure of a conserved state, with the arithmetic of Calculate_Primitive_Variables
 * (src/Basic_Functions.cpp:85-142). */
static void rho_p_of(const double *U, double *rho, double *p) {
    double Rho = U[0];
    double inv = 1.0 / Rho;
    double v1 = U[1] * inv;
    double v2 = U[2] * inv;
    if (fabs(v1) <= 1e-10) v1 = 0.0;
    if (fabs(v2) <= 1e-10) v2 = 0.0;
    double vmag = (v1 * v1 + v2 * v2);
    *rho = Rho;
    *p = GM1 * (U[3] - 0.5 * Rho * vmag);
}

/* Ghost state from the interior state across a boundary face with outward unit normal n. */
static void ghost_fstate(int bc, const FState *in, double n1, double n2, const double *fs,
                         FState *g) {
    double Ug[4];
    switch (bc) {
    case OBC_INLET_REF: { /* Supersonic_Inlet_Condition (src/Inlet_Boundary_Conditions.cpp:41-74) */
        double u = in->W[1], v = in->W[2];
        double vmag = (u * u + v * v) * 0.5;
        Ug[0] = fs[0];
        Ug[1] = fs[0] * u;
        Ug[2] = fs[0] * v;
        Ug[3] = (fs[3] / GM1) + fs[0] * vmag;
        state_from_U(Ug, g);
        return;
    }
    case OBC_SUBSONIC_INLET_REF: { /* Subsonic_Inlet_Condition (:7-39), energy as written there */
        double u = in->W[1], v = in->W[2];
        double Vm = 0.5 * (u * u + v * v);
        Ug[0] = fs[0];
        Ug[1] = fs[0] * u;
        Ug[2] = fs[0] * v;
        Ug[3] = (fs[3] / GM1) + fs[0] * Vm * Vm;
        state_from_U(Ug, g);
        return;
    }
    case OBC_WALL_NOSLIP: { /* Viscous_Wall_Boundary_Condition (src/Wall_Boundary_Conditions.cpp:55-76):
                           * momentum negated, density and energy kept, formed from primitives */
        double u = in->W[1], v = in->W[2];
        double vmag = 0.5 * (u * u + v * v);
        Ug[0] = in->W[0];
        Ug[1] = in->W[0] * (-u);
        Ug[2] = in->W[0] * (-v);
        Ug[3] = ((in->W[3] / GM1) + in->W[0] * vmag);
        state_from_U(Ug, g);
        return;
    }
    case OBC_SUBSONIC_EXIT: { /* back-pressure exit (Subsonic_Exit_Boundary_Condition,
                               * src/Exit_Boundary_Conditions.cpp:8-66) with the interior cell's
                               * density and sound speed as the reference state and the prescribed
                               * pressure fs[4]; supersonic outflow copies the interior state */
        const double r = in->W[0], u = in->W[1], v = in->W[2], p = in->W[3], c = in->W[4];
        if (u * n1 + v * n2 >= c) {
            *g = *in;
            return;
        }
        const double pb = fs[4];
        const double t = (p - pb) / (r * c * c);
        const double rf = r - t * r;
        const double uf = u + n1 * t * c;
        const double vf = v + n2 * t * c;
        double Uf[4], Uc[4];
        cons_from_prim(rf, uf, vf, pb, Uf);
        cons_from_prim(r, u, v, p, Uc);
        for (int q = 0; q < 4; q++) Ug[q] = 2.0 * Uf[q] - Uc[q];
        double rg, pg;
        rho_p_of(Ug, &rg, &pg);
        if (!(rg > 0.0) || !(pg > 0.0))
            for (int q = 0; q < 4; q++) Ug[q] = Uf[q]; /* non-physical extrapolation: the face state */
        state_from_U(Ug, g);
        return;
    }
    case OBC_SUBSONIC_INLET: { /* total pressure and density from the prescribed state's
                                * isentropic stagnation values, direction of its velocity, static
                                * pressure from the interior (capped at the total pressure) */
        const double rinf = fs[0], uinf = fs[1], vinf = fs[2], pinf = fs[3];
        const double q2 = uinf * uinf + vinf * vinf;
        const double f = 1.0 + 0.5 * GM1 * (q2 / (GAMMA * pinf / rinf));
        const double p0 = pinf * pow(f, GAMMA / GM1);
        const double r0 = rinf * pow(f, 1.0 / GM1);
        const double qinf = sqrt(q2);
        const double dx = (qinf > 0.0) ? uinf / qinf : -n1;
        const double dy = (qinf > 0.0) ? vinf / qinf : -n2;
        const double p = fmin(in->W[3], p0);
        const double rb = r0 * pow(p / p0, 1.0 / GAMMA);
        const double qb = sqrt(fmax((2.0 / GM1) * (GAMMA * p0 / r0 - GAMMA * p / rb), 0.0));
        cons_from_prim(rb, qb * dx, qb * dy, p, Ug);
        state_from_U(Ug, g);
        return;
    }
    case OBC_EXIT: /* Supersonic_Exit_Boundary_Condition (src/Exit_Boundary_Conditions.cpp:68-81) */
        *g = *in;
        return;
    case OBC_WALL:     /* InViscid_Wall_Boundary_Condition (src/Wall_Boundary_Conditions.cpp:79-124) */
    case OBC_SYMMETRY: { /* Symmetry_Boundary_Condition (:8-52) */
        double u = in->W[1], v = in->W[2];
        double vdotn = (u * n1 + v * n2);
        double v1g = u - 2.0 * vdotn * n1;
        double v2g = v - 2.0 * vdotn * n2;
        double vmag = 0.5 * (v1g * v1g + v2g * v2g);
        Ug[0] = in->W[0];
        Ug[1] = in->W[0] * v1g;
        Ug[2] = in->W[0] * v2g;
        Ug[3] = ((in->W[3] / GM1) + in->W[0] * vmag);
        state_from_U(Ug, g);
        return;
    }
    case OBC_FREESTREAM:
        cons_from_prim(fs[0], fs[1], fs[2], fs[3], Ug);
        state_from_U(Ug, g);
        return;
    case OBC_FARFIELD: { /* characteristic far field (Riemann invariants normal to the face) */
        double ri = in->W[0], ui = in->W[1], vi = in->W[2], pi = in->W[3], ci = in->W[4];
        double rf = fs[0], uf = fs[1], vf = fs[2], pf = fs[3];
        double cf = sqrt(GAMMA * pf / rf);
        double vni = ui * n1 + vi * n2, vnf = uf * n1 + vf * n2;
        double rb, ub, vb, pb;
        if (fabs(vni) >= ci) {
            if (vni < 0.0) { rb = rf; ub = uf; vb = vf; pb = pf; }
            else { rb = ri; ub = ui; vb = vi; pb = pi; }
        } else {
            double Rp = vni + 2.0 * ci / GM1;
            double Rm = vnf - 2.0 * cf / GM1;
            double vnb = 0.5 * (Rp + Rm);
            double cb = 0.25 * GM1 * (Rp - Rm);
            double s, ut, vt;
            if (vnb < 0.0) {
                s = pf / pow(rf, GAMMA);
                ut = uf - vnf * n1;
                vt = vf - vnf * n2;
            } else {
                s = pi / pow(ri, GAMMA);
                ut = ui - vni * n1;
                vt = vi - vni * n2;
            }
            rb = pow(cb * cb / (GAMMA * s), 1.0 / GM1);
            pb = rb * cb * cb / GAMMA;
            ub = ut + vnb * n1;
            vb = vt + vnb * n2;
        }
        cons_from_prim(rb, ub, vb, pb, Ug);
        state_from_U(Ug, g);
        return;
    }
    default:
        *g = *in;
        return;
    }
}

void on_ghost_state(int bc, const double *Win5, double nx, double ny, const double *fs5,
                    double *Wg5) {
    FState in, g;
    for (int k = 0; k < 5; k++) in.W[k] = Win5[k];
    cons_from_prim(Win5[0], Win5[1], Win5[2], Win5[3], in.U);
    ghost_fstate(bc, &in, nx, ny, fs5, &g);
    for (int k = 0; k < 5; k++) Wg5[k] = g.W[k];
}

void on_ghost_cons(int bc, const double *U4, double nx, double ny, const double *fs5, double *Ug4) {
    FState in, g;
    state_from_U(U4, &in);
    ghost_fstate(bc, &in, nx, ny, fs5, &g);
    for (int k = 0; k < 4; k++) Ug4[k] = g.U[k];
}

void or_apply_bc_ref(const OMesh *m, int inlet_bc, double rho_in, double p_in, double *U_all,
                     double *prim_all, const double *U_exit_src) {
    double fs[4] = {rho_in, 0.0, 0.0, p_in};
    const int nc = m->nc;
    FState in, g;
    for (int i = 0; i < m->nin; i++) {
        int c = m->inl[3 * i], f = m->inl[3 * i + 1], gh = m->inl[3 * i + 2];
        state_from_U(&U_all[4 * c], &in);
        in.W[1] = prim_all[11 * c + 1];
        in.W[2] = prim_all[11 * c + 2];
        ghost_fstate(inlet_bc, &in, m->nrm[8 * c + 2 * f], m->nrm[8 * c + 2 * f + 1], fs, &g);
        for (int k = 0; k < 4; k++) U_all[4 * gh + k] = g.U[k];
        or_prim(g.U, &prim_all[11 * gh]);
    }
    for (int i = 0; i < m->nex; i++) {
        int c = m->exl[3 * i], gh = m->exl[3 * i + 2];
        const double *src = U_exit_src ? &U_exit_src[4 * c] : &U_all[4 * c];
        for (int k = 0; k < 4; k++) U_all[4 * gh + k] = src[k];
        or_prim(&U_all[4 * gh], &prim_all[11 * gh]);
    }
    for (int i = 0; i < m->nwall; i++) {
        int c = m->wal[3 * i], f = m->wal[3 * i + 1], gh = m->wal[3 * i + 2];
        for (int k = 0; k < 5; k++) in.W[k] = 0.0;
        in.W[0] = prim_all[11 * c + 0];
        in.W[1] = prim_all[11 * c + 1];
        in.W[2] = prim_all[11 * c + 2];
        in.W[3] = prim_all[11 * c + 4];
        ghost_fstate(OBC_WALL, &in, m->nrm[8 * c + 2 * f], m->nrm[8 * c + 2 * f + 1], fs, &g);
        for (int k = 0; k < 4; k++) U_all[4 * gh + k] = g.U[k];
        or_prim(g.U, &prim_all[11 * gh]);
    }
    (void)nc;
}

/* ------------------------------------------------------------------------------------------
 * Face fluxes
 * ---------------------------------------------------------------------------------------- */
/* Calculate_Face_Average_Flux (src/Average_Interface_Flux.cpp:20-82), wall flag never set */
static void avg_flux(const FState *L, const FState *R, double nx, double ny, double dl, double *A) {
    double Rho_L = L->W[0], u_L = L->W[1], v_L = L->W[2], P_L = L->W[3];
    double Rho_R = R->W[0], u_R = R->W[1], v_R = R->W[2], P_R = R->W[3];
    double Vdotn_L = (u_L * nx + v_L * ny);
    double Vdotn_R = (u_R * nx + v_R * ny);
    double Vmag_L = 0.5 * (u_L * u_L + v_L * v_L);
    double Vmag_R = 0.5 * (u_R * u_R + v_R * v_R);
    double Vdotn_L_dl = Vdotn_L * dl;
    double Vdotn_R_dl = Vdotn_R * dl;
    double FL[4], FR[4];
    FL[0] = Rho_L * Vdotn_L_dl;
    FL[1] = Rho_L * u_L * Vdotn_L_dl + P_L * nx * dl;
    FL[2] = Rho_L * v_L * Vdotn_L_dl + P_L * ny * dl;
    FL[3] = (GAMMA1 * P_L + Rho_L * Vmag_L) * Vdotn_L_dl;
    FR[0] = Rho_R * Vdotn_R_dl;
    FR[1] = Rho_R * u_R * Vdotn_R_dl + P_R * nx * dl;
    FR[2] = Rho_R * v_R * Vdotn_R_dl + P_R * ny * dl;
    FR[3] = (GAMMA1 * P_R + Rho_R * Vmag_R) * Vdotn_R_dl;
    for (int i = 0; i < 4; i++) A[i] = 0.5 * (FL[i] + FR[i]);
}

/* ROE dissipation (src/Roe_Scheme.cpp:6-256) with Harten's entropy fix delta = 0.1 a~ */
static void roe_diss(const FState *L, const FState *R, double nx, double ny, double dl, double *D) {
    D[0] = D[1] = D[2] = D[3] = 0.0;
    double normal_magnitude = sqrt(nx * nx + ny * ny);
    if (normal_magnitude < 1e-12) return;
    double Rho_L = L->W[0], u_L = L->W[1], v_L = L->W[2], P_L = L->W[3], C_L = L->W[4];
    if (Rho_L <= 0.0 || P_L <= 0.0 || C_L <= 0.0) return;
    double Vmag_L = 0.5 * (u_L * u_L + v_L * v_L);
    double H_L = ((C_L * C_L) / (GAMMA - 1.0)) + Vmag_L;
    double Rho_R = R->W[0], u_R = R->W[1], v_R = R->W[2], P_R = R->W[3], C_R = R->W[4];
    if (Rho_R <= 0.0 || P_R <= 0.0 || C_R <= 0.0) return;
    double Vmag_R = 0.5 * (u_R * u_R + v_R * v_R);
    double H_R = ((C_R * C_R) / (GAMMA - 1.0)) + Vmag_R;
    double sqrt_Rho_L = sqrt(Rho_L);
    double sqrt_Rho_R = sqrt(Rho_R);
    double Term1 = 1.0 / (sqrt_Rho_L + sqrt_Rho_R);
    double Roe_Rho = sqrt(Rho_L * Rho_R);
    double Roe_u = Term1 * (u_R * sqrt_Rho_R + u_L * sqrt_Rho_L);
    double Roe_v = Term1 * (v_R * sqrt_Rho_R + v_L * sqrt_Rho_L);
    double Roe_H = Term1 * (H_R * sqrt_Rho_R + H_L * sqrt_Rho_L);
    double Roe_Vmag = 0.5 * (Roe_u * Roe_u + Roe_v * Roe_v);
    double Roe_a = sqrt((GAMMA - 1.0) * (Roe_H - Roe_Vmag));
    if (Roe_a <= 0.0) {
        Roe_a = 0.5 * (C_L + C_R);
        if (Roe_a <= 0.0) return;
    }
    double Un = nx * Roe_u + ny * Roe_v;
    double Ut = -ny * Roe_u + nx * Roe_v;
    double du = (u_R - u_L), dv = (v_R - v_L), dP = (P_R - P_L), dRho = (Rho_R - Rho_L);
    double dUn = nx * du + ny * dv;
    double dUt = -ny * du + nx * dv;
    double R12 = Roe_u - nx * Roe_a, R13 = Roe_v - ny * Roe_a, R14 = Roe_H - Roe_a * Un;
    double R22 = Roe_u, R23 = Roe_v, R24 = Roe_Vmag;
    double R32 = -ny, R33 = nx, R34 = Ut;
    double R42 = Roe_u + nx * Roe_a, R43 = Roe_v + ny * Roe_a, R44 = Roe_H + Roe_a * Un;
    double Term2 = 1.0 / (Roe_a * Roe_a);
    double a1 = 0.5 * Term2 * (dP - Roe_Rho * Roe_a * dUn);
    double a2 = dRho - Term2 * dP;
    double a3 = Roe_Rho * dUt;
    double a4 = 0.5 * Term2 * (dP + Roe_Rho * Roe_a * dUn);
    double L1 = fabs(Un - Roe_a), L2 = fabs(Un), L3 = fabs(Un), L4 = fabs(Un + Roe_a);
    double ef = 0.1 * Roe_a;
    if (L1 < ef) L1 = 0.5 * (L1 * L1 / ef + ef);
    if (L4 < ef) L4 = 0.5 * (L4 * L4 / ef + ef);
    D[0] = 0.5 * (L1 * a1 * 1.0 + L2 * a2 * 1.0 + L3 * a3 * 0.0 + L4 * a4 * 1.0) * dl;
    D[1] = 0.5 * (L1 * a1 * R12 + L2 * a2 * R22 + L3 * a3 * R32 + L4 * a4 * R42) * dl;
    D[2] = 0.5 * (L1 * a1 * R13 + L2 * a2 * R23 + L3 * a3 * R33 + L4 * a4 * R43) * dl;
    D[3] = 0.5 * (L1 * a1 * R14 + L2 * a2 * R24 + L3 * a3 * R34 + L4 * a4 * R44) * dl;
}

static double max3(double a, double b, double c) {
    /* Maximum(a,b,c) (src/Utility_Files.cpp:5-13) */
    if ((a >= b) && (a >= c)) return a;
    if ((b >= a) && (b >= c)) return b;
    return c;
}

/* LLF dissipation (src/LLF.cpp:8-83) */
static void llf_diss(const FState *L, const FState *R, double nx, double ny, double dl, double *D) {
    double Vdotn_L = (L->W[1] * nx + L->W[2] * ny);
    double Vdotn_R = (R->W[1] * nx + R->W[2] * ny);
    double C_L = L->W[4], C_R = R->W[4];
    double Max1 = max3(fabs(Vdotn_L - C_L) * dl, fabs(Vdotn_L + C_L) * dl, fabs(Vdotn_L) * dl);
    double Max2 = max3(fabs(Vdotn_R - C_R) * dl, fabs(Vdotn_R + C_R) * dl, fabs(Vdotn_R) * dl);
    double mev = (Max1 >= Max2) ? Max1 : Max2;
    for (int k = 0; k < 4; k++) D[k] = 0.5 * mev * (R->U[k] - L->U[k]);
}

/* HLLC (CUDA_KERNELS/HLLC_LLF_Flux_Cuda_Kernels.cu:25-208), flux times face length */
static void hllc_flux(const FState *L, const FState *R, double nx, double ny, double dl, double *F) {
    double rho_L = L->W[0], u_L = L->W[1], v_L = L->W[2], P_L = L->W[3], a_L = L->W[4];
    double rho_R = R->W[0], u_R = R->W[1], v_R = R->W[2], P_R = R->W[3], a_R = R->W[4];
    double un_L = nx * u_L + ny * v_L;
    double un_R = nx * u_R + ny * v_R;
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    /* one reciprocal for the three sqrt(rho)-weighted averages */
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
        f[3] = (L->U[3] + P_L) * un_L;
    } else if (S_R <= 0.0) {
        f[0] = rho_R * un_R;
        f[1] = rho_R * u_R * un_R + P_R * nx;
        f[2] = rho_R * v_R * un_R + P_R * ny;
        f[3] = (R->U[3] + P_R) * un_R;
    } else {
        const FState *K = (S_star >= 0.0) ? L : R;
        double S_K = (S_star >= 0.0) ? S_L : S_R;
        double rho = K->W[0], u = K->W[1], v = K->W[2], P = K->W[3], E = K->U[3];
        double un = nx * u + ny * v;
        double FK[4], Us[4];
        FK[0] = rho * un;
        FK[1] = rho * u * un + P * nx;
        FK[2] = rho * v * un + P * ny;
        FK[3] = (E + P) * un;
        /* star state (Toro eq. 10.39) with one reciprocal of S_K - S*:
         * U*_3 = (E (S_K - un) + (S* - un)(rho (S_K - un) S* + P)) / (S_K - S*) */
        double inv_k = 1.0 / (S_K - S_star);
        double mK = rho * (S_K - un);
        double factor = mK * inv_k;
        Us[0] = factor;
        Us[1] = factor * (u + (S_star - un) * nx);
        Us[2] = factor * (v + (S_star - un) * ny);
        Us[3] = (E * (S_K - un) + (S_star - un) * (mK * S_star + P)) * inv_k;
        for (int i = 0; i < 4; i++) f[i] = FK[i] + S_K * (Us[i] - K->U[i]);
    }
    for (int i = 0; i < 4; i++) F[i] = f[i] * dl;
}

/* AUSM (src/Ausm_Flux.cpp:36-160), states from conserved variables */
static void ausm_flux(const double *UL, const double *UR, double nx, double ny, double face_area,
                      double *F) {
    double rho_L = UL[0], rhou_L = UL[1], rhov_L = UL[2], rhoE_L = UL[3];
    double u_L = rhou_L / rho_L, v_L = rhov_L / rho_L;
    double p_L = (GAMMA - 1.0) * (rhoE_L - 0.5 * rho_L * (u_L * u_L + v_L * v_L));
    double H_L = (rhoE_L + p_L) / rho_L;
    double rho_R = UR[0], rhou_R = UR[1], rhov_R = UR[2], rhoE_R = UR[3];
    double u_R = rhou_R / rho_R, v_R = rhov_R / rho_R;
    double p_R = (GAMMA - 1.0) * (rhoE_R - 0.5 * rho_R * (u_R * u_R + v_R * v_R));
    double H_R = (rhoE_R + p_R) / rho_R;
    double Vn_L = u_L * nx + v_L * ny;
    double Vn_R = u_R * nx + v_R * ny;
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    double inv_sqrt_sum = 1.0 / (sqrt_rho_L + sqrt_rho_R);
    double u_roe = (sqrt_rho_L * u_L + sqrt_rho_R * u_R) * inv_sqrt_sum;
    double v_roe = (sqrt_rho_L * v_L + sqrt_rho_R * v_R) * inv_sqrt_sum;
    double H_roe = (sqrt_rho_L * H_L + sqrt_rho_R * H_R) * inv_sqrt_sum;
    double a_roe = sqrt((GAMMA - 1.0) * (H_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe)));
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
    if (M_interface >= 0.0) {
        double mass_flux = M_interface * a_roe * rho_L;
        F[0] = mass_flux * face_area;
        F[1] = (mass_flux * u_L + p_interface * nx) * face_area;
        F[2] = (mass_flux * v_L + p_interface * ny) * face_area;
        F[3] = mass_flux * H_L * face_area;
    } else {
        double mass_flux = M_interface * a_roe * rho_R;
        F[0] = mass_flux * face_area;
        F[1] = (mass_flux * u_R + p_interface * nx) * face_area;
        F[2] = (mass_flux * v_R + p_interface * ny) * face_area;
        F[3] = mass_flux * H_R * face_area;
    }
}

static void face_flux(const ONatCfg *cfg, int boundary, const FState *L, const FState *R,
                      double nx, double ny, double dl, double *F) {
    double A[4], D[4];
    switch (cfg->flux) {
    case OFLUX_ROE:
        avg_flux(L, R, nx, ny, dl, A);
        if (boundary && !cfg->roe_bdiss) D[0] = D[1] = D[2] = D[3] = 0.0;
        else roe_diss(L, R, nx, ny, dl, D);
        for (int k = 0; k < 4; k++) F[k] = A[k] - D[k];
        return;
    case OFLUX_LLF:
        avg_flux(L, R, nx, ny, dl, A);
        llf_diss(L, R, nx, ny, dl, D);
        for (int k = 0; k < 4; k++) F[k] = A[k] - D[k];
        return;
    case OFLUX_HLLC:
        hllc_flux(L, R, nx, ny, dl, F);
        return;
    case OFLUX_AUSM:
    default:
        ausm_flux(L->U, R->U, nx, ny, dl, F);
        return;
    }
}

void on_face_flux(const ONatCfg *cfg, int boundary, const double *WL5, const double *WR5,
                  double nx, double ny, double dl, double *F4) {
    FState L, R;
    for (int k = 0; k < 5; k++) {
        L.W[k] = WL5[k];
        R.W[k] = WR5[k];
    }
    cons_from_prim(WL5[0], WL5[1], WL5[2], WL5[3], L.U);
    cons_from_prim(WR5[0], WR5[1], WR5[2], WR5[3], R.U);
    face_flux(cfg, boundary, &L, &R, nx, ny, dl, F4);
}

/* ------------------------------------------------------------------------------------------
 * Limited slopes (limiter functions of CUDA_KERNELS/Limiter_Cuda_Kernels.cu:89-157 applied to
 * r = dminus/dplus, returned as phi(r) * dplus without dividing)
 * ---------------------------------------------------------------------------------------- */
double or_limiter_slope(int limiter, double a, double b) {
    switch (limiter) {
    case OLIM_NONE:
        return 0.5 * (a + b);
    case OLIM_MINMOD:
        if (a * b <= 0.0) return 0.0;
        return (b > 0.0) ? fmin(a, b) : fmax(a, b);
    case OLIM_VANLEER:
        if (a * b <= 0.0) return 0.0;
        return 2.0 * a * b / (a + b);
    case OLIM_SUPERBEE: {
        if (a * b <= 0.0) return 0.0;
        double aa = fabs(a), bb = fabs(b);
        double s = fmax(fmin(2.0 * aa, bb), fmin(aa, 2.0 * bb));
        return (b > 0.0) ? s : -s;
    }
    case OLIM_VANALBADA: {
        /* van Albada phi(r) = (r^2 + r) / (r^2 + 1) (CUDA_KERNELS/Limiter_Cuda_Kernels.cu:146-154)
         * in the differentiable form with e = 1e-4 (no cut-off at r <= 0) */
        const double e = 1e-4;
        return (a * (b * b + e) + b * (a * a + e)) / (a * a + b * b + 2.0 * e);
    }
    default:
        return 0.0;
    }
}

/* ------------------------------------------------------------------------------------------
 * Gradients of (rho, u, v, p)
 * ---------------------------------------------------------------------------------------- */
/* The prescribed state (rho, u, v, p) and the exits' static pressure (p_exit, else p). */
static void fs5_of(const ONatCfg *cfg, double *fs5) {
    for (int k = 0; k < 4; k++) fs5[k] = cfg->fs[k];
    fs5[4] = (cfg->p_exit > 0.0) ? cfg->p_exit : cfg->fs[3];
}

static int bc_of(const ONatCfg *cfg, int kind) { return (kind >= 1 && kind <= 5) ? cfg->bc[kind] : OBC_WALL; }

/* W5 per physical cell; grad[8*nc]: (drho/dx, drho/dy, du/dx, du/dy, dv/dx, dv/dy, dp/dx, dp/dy) */
void on_gradients(const OMesh *m, const ONatCfg *cfg, const double *W5, double *grad) {
    const int nc = m->nc;
    double fs5[5];
    fs5_of(cfg, fs5);
    if (cfg->gradient == OGRAD_GG) {
        /* Calculate_Gradient_At_Cell_Center (src/Grid_Refine_Functions.cpp:221-246) for each of
         * (rho,u,v,p); ghost faces use the cell's own value */
#pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; c++) {
            for (int k = 0; k < 4; k++) {
                double gx = 0.0, gy = 0.0;
                double phi_c = W5[5 * c + k];
                for (int f = 0; f < m->nf[c]; f++) {
                    int j = m->nbr[4 * c + f];
                    double phi_n = (j >= 0 && j < nc) ? W5[5 * j + k] : phi_c;
                    double phi_face = 0.5 * (phi_c + phi_n);
                    gx += phi_face * m->nrm[8 * c + 2 * f] * m->len[4 * c + f];
                    gy += phi_face * m->nrm[8 * c + 2 * f + 1] * m->len[4 * c + f];
                }
                grad[8 * c + 2 * k] = gx * m->inv_area[c];
                grad[8 * c + 2 * k + 1] = gy * m->inv_area[c];
            }
        }
        return;
    }
    /* weighted least squares, weights 1/|d|^2, ghost neighbours at the reflected centre with the
     * boundary-condition ghost state */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; c++) {
        double a = 0.0, b = 0.0, d = 0.0;
        int n = m->nf[c];
        double dx[4], dy[4], w[4];
        for (int f = 0; f < n; f++) {
            int j = m->nbr[4 * c + f];
            dx[f] = m->ctr[3 * j] - m->ctr[3 * c];
            dy[f] = m->ctr[3 * j + 1] - m->ctr[3 * c + 1];
            w[f] = 1.0 / (dx[f] * dx[f] + dy[f] * dy[f]);
            a += w[f] * dx[f] * dx[f];
            b += w[f] * dx[f] * dy[f];
            d += w[f] * dy[f] * dy[f];
        }
        double det = a * d - b * b;
        double i11 = d / det, i12 = -b / det, i22 = a / det;
        for (int k = 0; k < 8; k++) grad[8 * c + k] = 0.0;
        FState in;
        for (int q = 0; q < 5; q++) in.W[q] = W5[5 * c + q];
        cons_from_prim(in.W[0], in.W[1], in.W[2], in.W[3], in.U);
        for (int f = 0; f < n; f++) {
            int j = m->nbr[4 * c + f];
            double cx = (i11 * dx[f] + i12 * dy[f]) * w[f];
            double cy = (i12 * dx[f] + i22 * dy[f]) * w[f];
            double Wn[4];
            if (j < nc) {
                for (int k = 0; k < 4; k++) Wn[k] = W5[5 * j + k];
            } else {
                FState g;
                ghost_fstate(bc_of(cfg, m->kind[4 * c + f]), &in, m->nrm[8 * c + 2 * f],
                             m->nrm[8 * c + 2 * f + 1], fs5, &g);
                for (int k = 0; k < 4; k++) Wn[k] = g.W[k];
            }
            for (int k = 0; k < 4; k++) {
                double dW = Wn[k] - W5[5 * c + k];
                grad[8 * c + 2 * k] += cx * dW;
                grad[8 * c + 2 * k + 1] += cy * dW;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Face-loop residual
 * ---------------------------------------------------------------------------------------- */
/* Limiter freezing: the limited slope equals psi * c with c = grad W . d the unlimited central
 * slope ((d- + d+) / 2) and psi in [0, 1] for minmod and Van Leer (<= 4/3 for superbee). Mode 1
 * records psi = slope / c (0 where c = 0); mode 2 uses slope = psi * c with the recorded psi,
 * which makes the iteration near a converged state differentiable (limiter switching at shocks
 * otherwise stalls steady convergence). */
static void reconstruct_side(const ONatCfg *cfg, const FState *cellS, const FState *otherS,
                             const double *g, double dx, double dy, double sgn, float *phi,
                             FState *out) {
    double Wf[4];
    for (int k = 0; k < 4; k++) {
        double dplus = sgn * (otherS->W[k] - cellS->W[k]);
        double c = g[2 * k] * dx + g[2 * k + 1] * dy;
        double slope;
        if (cfg->freeze == 2) {
            slope = (double)phi[k] * c;
        } else {
            double dminus = 2.0 * c - dplus;
            slope = or_limiter_slope(cfg->limiter, dminus, dplus);
            if (cfg->freeze == 1) phi[k] = (float)((c != 0.0) ? slope / c : 0.0);
        }
        Wf[k] = cellS->W[k] + sgn * 0.5 * slope;
    }
    if (!(Wf[0] > 0.0) || !(Wf[3] > 0.0)) {
        /* first order on this side: the cell's primitives, conserved values formed from them */
        for (int k = 0; k < 5; k++) out->W[k] = cellS->W[k];
        cons_from_prim(cellS->W[0], cellS->W[1], cellS->W[2], cellS->W[3], out->U);
        return;
    }
    out->W[0] = Wf[0];
    out->W[1] = Wf[1];
    out->W[2] = Wf[2];
    out->W[3] = Wf[3];
    out->W[4] = sqrt(GAMMA * Wf[3] * (1.0 / Wf[0]));
    /* conserved values of the face state, with p / (gamma - 1) as a product by the reciprocal */
    const double vmag = 0.5 * (Wf[1] * Wf[1] + Wf[2] * Wf[2]);
    out->U[0] = Wf[0];
    out->U[1] = Wf[0] * Wf[1];
    out->U[2] = Wf[0] * Wf[2];
    out->U[3] = Wf[3] * (1.0 / GM1) + Wf[0] * vmag;
}

/* ------------------------------------------------------------------------------------------
 * Laminar viscous terms
 * ---------------------------------------------------------------------------------------- */
/* Dynamic viscosity at temperature T: constant mu, or Sutherland's law C1 T^1.5 / (T + S) with the
 * constants of CUDA_KERNELS/Viscous_Flux_Cuda_Kernels.cu:6-7,12-14 and T^1.5 as T sqrt(T). */
static double viscosity(const ONatCfg *cfg, double T) {
    if (cfg->viscous == 2) return 1.45793265452e-06 * (T * sqrt(T)) / (T + 110.4);
    return cfg->mu;
}

/* Face gradient of one variable: mean of the two cell gradients, its component along the
 * centre-to-centre unit vector t replaced by (phi_R - phi_L) / |d|. */
static void face_gradient(double axL, double ayL, double axR, double ayR, double phiL, double phiR,
                          double inv_dist, double tx, double ty, double *gx, double *gy) {
    double ax = 0.5 * (axL + axR);
    double ay = 0.5 * (ayL + ayR);
    double corr = (phiR - phiL) * inv_dist - (ax * tx + ay * ty);
    *gx = ax + corr * tx;
    *gy = ay + corr * ty;
}

/* Viscous face flux times the face length (Viscous_Flux_on_Face,
 * src/Grid_Refine_Functions.cpp:418-479, dimensional mu in place of mu / Re):
 *   tau_xx = 2/3 mu (2 u_x - v_y), tau_xy = mu (u_y + v_x), tau_yy = 2/3 mu (2 v_y - u_x),
 *   Fv = (0, tau_xx nx + tau_xy ny, tau_xy nx + tau_yy ny,
 *         (tau_xx u + tau_xy v + k T_x) nx + (tau_xy u + tau_yy v + k T_y) ny) dl
 * with face velocity, temperature and (directionally corrected) gradients from the two cells;
 * T = p / (rho R), grad T = (grad p - (p / rho) grad rho) / (rho R), k = mu gamma R / ((gamma-1) Pr). */
static void viscous_flux(const ONatCfg *cfg, const double *WL, const double *WR, const double *gL,
                         const double *gR, double nx, double ny, double dl, double dx, double dy,
                         double *Fv) {
    double kfac = GAMMA * cfg->r_gas / ((GAMMA - 1.0) * cfg->prandtl);
    double rgL = WL[0] * cfg->r_gas, rgR = WR[0] * cfg->r_gas;
    double TL = WL[3] / rgL, TR = WR[3] / rgR;
    double qL = WL[3] / WL[0], qR = WR[3] / WR[0];
    double TxL = (gL[6] - qL * gL[0]) / rgL, TyL = (gL[7] - qL * gL[1]) / rgL;
    double TxR = (gR[6] - qR * gR[0]) / rgR, TyR = (gR[7] - qR * gR[1]) / rgR;
    double dist = sqrt(dx * dx + dy * dy);
    double inv_dist = 1.0 / dist;
    double tx = dx * inv_dist, ty = dy * inv_dist;
    double ux, uy, vx, vy, Tx, Ty;
    face_gradient(gL[2], gL[3], gR[2], gR[3], WL[1], WR[1], inv_dist, tx, ty, &ux, &uy);
    face_gradient(gL[4], gL[5], gR[4], gR[5], WL[2], WR[2], inv_dist, tx, ty, &vx, &vy);
    face_gradient(TxL, TyL, TxR, TyR, TL, TR, inv_dist, tx, ty, &Tx, &Ty);
    double mu = viscosity(cfg, 0.5 * (TL + TR));
    double k = mu * kfac;
    double t11 = (2.0 / 3.0) * mu * (2.0 * ux - vy);
    double t12 = mu * (uy + vx);
    double t22 = (2.0 / 3.0) * mu * (2.0 * vy - ux);
    double uf = 0.5 * (WL[1] + WR[1]), vf = 0.5 * (WL[2] + WR[2]);
    double Qx = k * Tx, Qy = k * Ty;
    Fv[0] = 0.0;
    Fv[1] = (t11 * nx + t12 * ny) * dl;
    Fv[2] = (t12 * nx + t22 * ny) * dl;
    Fv[3] = ((t11 * uf + t12 * vf + Qx) * nx * dl + (t12 * uf + t22 * vf + Qy) * ny * dl);
}

/* Viscous part of the time-step denominator, Viscous_Time_Step_1 (src/Time_Step.cpp:103-109)
 * with the cell's own density and viscosity: 4 (1/A) max(4/(3 rho), gamma/rho) (mu/Pr) sum dl^2. */
static double viscous_dt_term(const ONatCfg *cfg, double rho, double p, double sdl2, double area) {
    double mu = viscosity(cfg, p / (rho * cfg->r_gas));
    return 4.0 * ((fmax(4.0 / (3.0 * rho), GAMMA / rho) * mu * (1.0 / cfg->prandtl) * sdl2) / area);
}

/* ------------------------------------------------------------------------------------------
 * WENO5 (order 3)
 * ---------------------------------------------------------------------------------------- */
/* WENO_Reconstruction (src/WENO2D.cpp:9-79): fifth-order WENO-JS (epsilon 1e-6, p = 2) of one
 * variable on the stencil a..e centred on c; shift 0 gives the value at the edge between c and d,
 * shift 1 at the edge between b and c. pow(x, 2) is written as x * x. */
double or_weno5(double a, double b, double c, double d, double e, int shift) {
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
    double s2a = a - 2.0 * b + c, s2b = a - 4.0 * b + 3.0 * c;
    double s1a = b - 2.0 * c + d, s1b = b - d;
    double s0a = c - 2.0 * d + e, s0b = 3.0 * c - 4.0 * d + e;
    double b2 = (13.0 / 12.0) * (s2a * s2a) + (1.0 / 4.0) * (s2b * s2b);
    double b1 = (13.0 / 12.0) * (s1a * s1a) + (1.0 / 4.0) * (s1b * s1b);
    double b0 = (13.0 / 12.0) * (s0a * s0a) + (1.0 / 4.0) * (s0b * s0b);
    double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
    double a0 = d0 / (e0 * e0);
    double a1 = d1 / (e1 * e1);
    double a2 = d2 / (e2 * e2);
    double sum = a0 + a1 + a2;
    double U;
    if (sum < eps) {
        U = (v0 + v1 + v2) / 3.0;
    } else {
        double w0 = a0 / sum, w1 = a1 / sum, w2 = a2 / sum;
        U = w0 * v0 + w1 * v1 + w2 * v2;
    }
    return isfinite(U) ? U : c;
}

/* The physical cell across the face opposite slot f of quadrilateral c, or -1. */
static int cell_behind(const OMesh *m, int c, int f) {
    if (c < 0 || f < 0 || m->nf[c] != 4) return -1;
    int j = m->nbr[4 * c + (f + 2) % 4];
    return (j >= 0 && j < m->nc) ? j : -1;
}

static int slot_to(const OMesh *m, int c, int nb) {
    if (c < 0) return -1;
    for (int f = 0; f < m->nf[c]; f++)
        if (m->nbr[4 * c + f] == nb) return f;
    return -1;
}

/* Line stencil of the interior face between L and R through opposite faces of quadrilaterals:
 * st = {L-1, L-2, R+1, R+2}; returns 0 when any of them is missing. */
static int line_stencil(const OMesh *m, int L, int R, int *st) {
    int l1 = cell_behind(m, L, slot_to(m, L, R));
    int l2 = cell_behind(m, l1, slot_to(m, l1, L));
    int r1 = cell_behind(m, R, slot_to(m, R, L));
    int r2 = cell_behind(m, r1, slot_to(m, r1, R));
    if (l1 < 0 || l2 < 0 || r1 < 0 || r2 < 0) return 0;
    st[0] = l1;
    st[1] = l2;
    st[2] = r1;
    st[3] = r2;
    return 1;
}

/* Face state from reconstructed conserved values (src/WENO2D.cpp:482-489); a non-positive
 * density or pressure falls back to the cell state. */
static void weno_face_state(const double *U, const FState *cell, FState *out) {
    state_from_U(U, out);
    if (!(out->W[0] > 0.0) || !(out->W[3] > 0.0)) *out = *cell;
}

/* The face-loop residual; fo (may be NULL) marks cells whose faces all take first-order states
 * (the positivity fallback's recomputation). */
static void residual_fo(const OMesh *m, const ONatCfg *cfg, const double *U, const double *U_exit_src,
                        double *R, double *dt, double *grad, const unsigned char *fo);

void on_residual(const OMesh *m, const ONatCfg *cfg, const double *U, const double *U_exit_src,
                 double *R, double *dt, double *grad) {
    residual_fo(m, cfg, U, U_exit_src, R, dt, grad, NULL);
}

static void residual_fo(const OMesh *m, const ONatCfg *cfg, const double *U, const double *U_exit_src,
                        double *R, double *dt, double *grad, const unsigned char *fo) {
    double fs5[5];
    fs5_of(cfg, fs5);
    const int nc = m->nc;
    FState *S = (FState *)malloc(sizeof(FState) * nc);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; c++) state_from_U(&U[4 * c], &S[c]);
    double *W5 = NULL;
    if (cfg->order >= 2) {
        W5 = (double *)malloc(sizeof(double) * 5 * nc);
#pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; c++)
            for (int k = 0; k < 5; k++) W5[5 * c + k] = S[c].W[k];
        on_gradients(m, cfg, W5, grad);
    }
    double *Ff = (double *)malloc(sizeof(double) * 4 * m->nfaces);
#pragma omp parallel for schedule(static)
    for (int fi = 0; fi < m->nfaces; fi++) {
        int L = m->fL[fi], Rc = m->fR[fi], s = m->fslot[fi];
        double nx = m->nrm[8 * L + 2 * s], ny = m->nrm[8 * L + 2 * s + 1], dl = m->len[4 * L + s];
        int boundary = Rc >= nc;
        int bc = boundary ? bc_of(cfg, m->kind[4 * L + s]) : 0;
        FState SL = S[L], SR;
        FState srcL = S[L];
        if (boundary && bc == OBC_EXIT && U_exit_src) state_from_U(&U_exit_src[4 * L], &srcL);
        if (boundary) ghost_fstate(bc, &srcL, nx, ny, fs5, &SR);
        else SR = S[Rc];
        int st[4];
        if (cfg->order == 3 && !boundary && line_stencil(m, L, Rc, st)) {
            /* WENO5 on the conserved variables, cells L-2, L-1, L | R, R+1, R+2 */
            double UL[4], UR[4];
            for (int k = 0; k < 4; k++) {
                UL[k] = or_weno5(U[4 * st[1] + k], U[4 * st[0] + k], U[4 * L + k], U[4 * Rc + k],
                                 U[4 * st[2] + k], 0);
                UR[k] = or_weno5(U[4 * st[0] + k], U[4 * L + k], U[4 * Rc + k], U[4 * st[2] + k],
                                 U[4 * st[3] + k], 1);
            }
            FState fL, fR;
            weno_face_state(UL, &SL, &fL);
            weno_face_state(UR, &SR, &fR);
            face_flux(cfg, 0, &fL, &fR, nx, ny, dl, &Ff[4 * fi]);
            if (cfg->viscous) {
                double Fv[4];
                double dx = m->ctr[3 * Rc] - m->ctr[3 * L];
                double dy = m->ctr[3 * Rc + 1] - m->ctr[3 * L + 1];
                viscous_flux(cfg, SL.W, SR.W, &grad[8 * L], &grad[8 * Rc], nx, ny, dl, dx, dy, Fv);
                for (int k = 0; k < 4; k++) Ff[4 * fi + k] = Ff[4 * fi + k] - Fv[k];
            }
            continue;
        }
        const int first = fo && (fo[L] || (!boundary && fo[Rc]));
        if (cfg->order >= 2 && !first) {
            double dx = m->ctr[3 * Rc] - m->ctr[3 * L];
            double dy = m->ctr[3 * Rc + 1] - m->ctr[3 * L + 1];
            FState fL, fR;
            float *phiL = cfg->phi ? &cfg->phi[8 * (size_t)fi] : NULL;
            reconstruct_side(cfg, &SL, &SR, &grad[8 * L], dx, dy, 1.0, phiL, &fL);
            if (boundary) {
                FState src = fL;
                if (bc == OBC_EXIT && U_exit_src) src = srcL;
                ghost_fstate(bc, &src, nx, ny, fs5, &fR);
            } else {
                reconstruct_side(cfg, &SR, &SL, &grad[8 * Rc], dx, dy, -1.0, phiL ? phiL + 4 : NULL, &fR);
            }
            face_flux(cfg, boundary, &fL, &fR, nx, ny, dl, &Ff[4 * fi]);
            if (cfg->viscous) {
                /* cell-centre states (ghost on a boundary face, with the cell's own gradient) */
                double Fv[4];
                const double *gR = boundary ? &grad[8 * L] : &grad[8 * Rc];
                viscous_flux(cfg, SL.W, SR.W, &grad[8 * L], gR, nx, ny, dl, dx, dy, Fv);
                for (int k = 0; k < 4; k++) Ff[4 * fi + k] = Ff[4 * fi + k] - Fv[k];
            }
        } else {
            face_flux(cfg, boundary, &SL, &SR, nx, ny, dl, &Ff[4 * fi]);
        }
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; c++) {
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < m->nf[c]; f++) {
            int fi = m->cface[4 * c + f];
            double sg = (double)m->csign[4 * c + f];
            for (int k = 0; k < 4; k++) r[k] += sg * Ff[4 * fi + k];
        }
        for (int k = 0; k < 4; k++) R[4 * c + k] = r[k];
        /* Inviscid_Time_Step (src/Time_Step.cpp:293-329) */
        double u0 = S[c].W[1], v0 = S[c].W[2], a0 = S[c].W[4];
        double denom = 0.0, sdl2 = 0.0;
        for (int f = 0; f < m->nf[c]; f++) {
            double vn = u0 * m->nrm[8 * c + 2 * f] + v0 * m->nrm[8 * c + 2 * f + 1];
            denom += (fabs(vn) + a0) * m->len[4 * c + f];
            sdl2 += m->len[4 * c + f] * m->len[4 * c + f];
        }
        if (cfg->order >= 2 && cfg->viscous) denom += viscous_dt_term(cfg, S[c].W[0], S[c].W[3], sdl2, m->area[c]);
        dt[c] = (denom <= 0.0 || !isfinite(denom)) ? 0.0 : cfg->cfl * m->area[c] / denom;
    }
    free(Ff);
    free(W5);
    free(S);
}

/* ------------------------------------------------------------------------------------------
 * Time integration
 * ---------------------------------------------------------------------------------------- */
void or_error(int n, const double *dU, const double *U, double *err) {
    /* Estimate_Error (src/Error_Estimate_Update.cpp:10-28) */
    for (int k = 0; k < 4; k++) err[k] = 0.0;
    for (int c = 0; c < n; c++)
        for (int k = 0; k < 4; k++) {
            double d = dU[4 * c + k];
            double t = (d < 1e-9) ? fabs(d) : (fabs(d) / U[4 * c + k]);
            err[k] += t * t;
        }
    for (int k = 0; k < 4; k++) err[k] = sqrt(err[k]);
}

static double min_dt(int n, const double *dt) {
    double m = DBL_MAX;
    for (int i = 0; i < n; i++)
        if (dt[i] < m) m = dt[i];
    return m;
}

/* One stage of the native integrators, elementwise over n cells (4 values each):
 *   RK3 (Shu-Osher):  s0: Un - hR;  s1: 3/4 Un + 1/4 (Us - hR);  s2: 1/3 Un + 2/3 (Us - hR)
 *   RK4 (classical):  s0: acc = R, Un - h/2 R;  s1: acc += 2R, Un - h/2 R;
 *                     s2: acc += 2R, Un - h R;  s3: Un - h/6 (acc + R)
 * with h = dt / area per cell, R the residual of the stage state Us. */
void or_rk_stage(int rk, int stage, int n, const double *Un, const double *Us, double *acc,
                 const double *R, const double *h, double *out) {
    for (int c = 0; c < n; c++)
        for (int k = 0; k < 4; k++) {
            size_t i = 4 * (size_t)c + k;
            double hc = h[c];
            if (rk == 4) {
                switch (stage) {
                case 0: acc[i] = R[i]; out[i] = Un[i] - 0.5 * hc * R[i]; break;
                case 1: acc[i] += 2.0 * R[i]; out[i] = Un[i] - 0.5 * hc * R[i]; break;
                case 2: acc[i] += 2.0 * R[i]; out[i] = Un[i] - hc * R[i]; break;
                default: out[i] = Un[i] - (hc / 6.0) * (acc[i] + R[i]); break;
                }
            } else {
                switch (stage) {
                case 0: out[i] = Un[i] - hc * R[i]; break;
                case 1: out[i] = 0.75 * Un[i] + 0.25 * (Us[i] - hc * R[i]); break;
                default: out[i] = (1.0 / 3.0) * Un[i] + (2.0 / 3.0) * (Us[i] - hc * R[i]); break;
                }
            }
        }
}


static int g_fallback_cells = 0;

/* Threads the parallel loops of the residual use (OpenMP; 1 when built without it). The loops are
 * over independent faces and cells, so the results do not depend on the thread count. */
int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int on_last_fallback_cells(void) { return g_fallback_cells; }

/* Positivity fallback, detection (no counterpart in the reference; an a-posteriori "MOOD"
 * check): flag[c] = 1 when the stage state Unew of cell c has a non-positive (or non-finite)
 * density or pressure, or a density or pressure outside [lo - d, hi + d], where lo and hi are the
 * minimum and maximum of the stage input Uin over the cell and its physical face neighbours and
 * d = 1e-3 max(|hi|, hi - lo) (the range test, when `range` is set). Returns the number of flagged
 * cells. */
static int stage_fallback_flags(const OMesh *m, const double *Uin, const double *Unew,
                                unsigned char *flag, int range) {
    const int nc = m->nc;
    int nb = 0;
    for (int c = 0; c < nc; c++) {
        double r0, p0, rn, pn;
        rho_p_of(&Uin[4 * c], &r0, &p0);
        rho_p_of(&Unew[4 * c], &rn, &pn);
        /* the range's bounds from the stage inputs rounded to FP32 (the margin is 1e-3 relative) */
        r0 = (double)(float)r0;
        p0 = (double)(float)p0;
        double rlo = r0, rhi = r0, plo = p0, phi = p0;
        for (int f = 0; f < m->nf[c]; f++) {
            int j = m->nbr[4 * c + f];
            if (j < 0 || j >= nc) continue;
            double rj, pj;
            rho_p_of(&Uin[4 * j], &rj, &pj);
            rj = (double)(float)rj;
            pj = (double)(float)pj;
            rlo = fmin(rlo, rj);
            rhi = fmax(rhi, rj);
            plo = fmin(plo, pj);
            phi = fmax(phi, pj);
        }
        const double dr = fmax(1e-3 * fabs(rhi), 1e-3 * (rhi - rlo));
        const double dp = fmax(1e-3 * fabs(phi), 1e-3 * (phi - plo));
        const int bad = !(rn > 0.0) || !(pn > 0.0) ||
                        (range && (!(rn >= rlo - dr && rn <= rhi + dr) || !(pn >= plo - dp && pn <= phi + dp)));
        flag[c] = (unsigned char)bad;
        nb += bad;
    }
    g_fallback_cells += nb;
    return nb;
}

void on_step(const OMesh *m, const ONatCfg *cfg, int rk, int lts, double *U, double *err4,
             double *dt_out) {
    const int nc = m->nc;
    size_t n4 = 4 * (size_t)nc;
    double *Un = (double *)malloc(sizeof(double) * n4);
    double *Us = (double *)malloc(sizeof(double) * n4);
    double *R = (double *)malloc(sizeof(double) * n4);
    double *acc = (double *)malloc(sizeof(double) * n4);
    double *dt = (double *)malloc(sizeof(double) * nc);
    double *h = (double *)malloc(sizeof(double) * nc);
    double *grad = (double *)malloc(sizeof(double) * 8 * nc);
    memcpy(Un, U, sizeof(double) * n4);
    memcpy(Us, U, sizeof(double) * n4);
    const int nst = (rk == 4) ? 4 : 3;
    const int pos = cfg->positivity && cfg->order == 2;
    double *Uin = pos ? (double *)malloc(sizeof(double) * n4) : NULL;
    double *acc_in = pos ? (double *)malloc(sizeof(double) * n4) : NULL;
    unsigned char *flag = pos ? (unsigned char *)malloc(nc) : NULL;
    double dtg = 0.0;
    g_fallback_cells = 0;
    for (int s = 0; s < nst; s++) {
        on_residual(m, cfg, Us, NULL, R, dt, grad);
        if (s == 0) {
            dtg = min_dt(nc, dt);
            for (int c = 0; c < nc; c++) h[c] = (lts ? dt[c] : dtg) * m->inv_area[c];
        }
        double *out = (s == nst - 1) ? U : Us;
        if (pos) {
            memcpy(Uin, Us, sizeof(double) * n4);
            memcpy(acc_in, acc, sizeof(double) * n4);
        }
        or_rk_stage(rk == 4 ? 4 : 3, s, nc, Un, Us, acc, R, h, out);
        if (pos && stage_fallback_flags(m, Uin, out, flag, cfg->positivity == 1)) {
            /* recompute the stage with first-order states on every face of a flagged cell */
            residual_fo(m, cfg, Uin, NULL, R, dt, grad, flag);
            memcpy(acc, acc_in, sizeof(double) * n4);
            or_rk_stage(rk == 4 ? 4 : 3, s, nc, Un, Uin, acc, R, h, out);
        }
    }
    free(Uin);
    free(acc_in);
    free(flag);
    for (size_t i = 0; i < n4; i++) acc[i] = U[i] - Un[i];
    or_error(nc, acc, Un, err4);
    *dt_out = dtg;
    free(Un); free(Us); free(R); free(acc); free(dt); free(h); free(grad);
}

void on_refstep(const OMesh *m, const ONatCfg *cfg, int rk, int lts, double *U, double *err4,
                double *dt_out) {
    const int nc = m->nc;
    size_t n4 = 4 * (size_t)nc;
    double *Un = (double *)malloc(sizeof(double) * n4);
    double *U1 = (double *)malloc(sizeof(double) * n4);
    double *U2 = (double *)malloc(sizeof(double) * n4);
    double *R = (double *)malloc(sizeof(double) * n4);
    double *dU = (double *)malloc(sizeof(double) * n4);
    double *dt = (double *)malloc(sizeof(double) * nc);
    double *grad = (double *)malloc(sizeof(double) * 8 * nc);
    memcpy(Un, U, sizeof(double) * n4);
    if (rk == 1) {
        /* Explicit_Method (src/Numerical_Method.cpp:26-103): global dt, LTS ignored */
        on_residual(m, cfg, Un, Un, R, dt, grad);
        double dtg = min_dt(nc, dt);
        for (int c = 0; c < nc; c++)
            for (int k = 0; k < 4; k++) dU[4 * c + k] = -dtg * m->inv_area[c] * R[4 * c + k];
        *dt_out = dtg;
    } else {
        /* Runge_Kutta_Method (src/Numerical_Method.cpp:342-457) */
        on_residual(m, cfg, Un, Un, R, dt, grad);
        double dtg = min_dt(nc, dt);
        for (int c = 0; c < nc; c++) {
            double d = lts ? dt[c] : dtg;
            for (int k = 0; k < 4; k++) U1[4 * c + k] = Un[4 * c + k] - d * m->inv_area[c] * (R[4 * c + k] - 0.0);
        }
        on_residual(m, cfg, U1, Un, R, dt, grad);
        dtg = min_dt(nc, dt);
        for (int c = 0; c < nc; c++)
            for (int k = 0; k < 4; k++)
                U2[4 * c + k] = 0.25 * (U1[4 * c + k] + 3.0 * Un[4 * c + k] - dtg * m->inv_area[c] * (R[4 * c + k] - 0.0));
        on_residual(m, cfg, U2, Un, R, dt, grad);
        dtg = min_dt(nc, dt);
        for (int c = 0; c < nc; c++)
            for (int k = 0; k < 4; k++)
                dU[4 * c + k] = (2.0 / 3.0) * (U2[4 * c + k] - Un[4 * c + k] - dtg * m->inv_area[c] * (R[4 * c + k] - 0.0));
        *dt_out = dtg;
    }
    /* Estimate_Error then Update (src/Error_Estimate_Update.cpp:10-44) */
    or_error(nc, dU, Un, err4);
    for (size_t i = 0; i < n4; i++) U[i] = Un[i] + dU[i];
    free(Un); free(U1); free(U2); free(R); free(dU); free(dt); free(grad);
}

/* ------------------------------------------------------------------------------------------
 * Reference per-cell functions
 * ---------------------------------------------------------------------------------------- */
void or_gg_gradient(const OMesh *m, const double *prim, int var, double *grad2) {
    const int nc = m->nc;
    for (int c = 0; c < nc; c++) {
        double g0 = 0.0, g1 = 0.0;
        double phi_c = prim[11 * c + var];
        for (int f = 0; f < m->nf[c]; f++) {
            double nx_f = m->nrm[8 * c + 2 * f], ny_f = m->nrm[8 * c + 2 * f + 1];
            double dl_f = m->len[4 * c + f];
            int j = m->nbr[4 * c + f];
            double phi_n = phi_c;
            if (j >= 0 && j < nc) phi_n = prim[11 * j + var];
            double phi_face = 0.5 * (phi_c + phi_n);
            g0 += phi_face * nx_f * dl_f;
            g1 += phi_face * ny_f * dl_f;
        }
        grad2[2 * c] = g0 * m->inv_area[c];
        grad2[2 * c + 1] = g1 * m->inv_area[c];
    }
}

void or_ausm_cell(const OMesh *m, const double *U_all, double *R) {
    /* Ausm_Flux(cell) (src/Ausm_Flux.cpp:5-204): ghost faces skipped; the wall-pressure term
     * needs Cells_Face_Boundary_Type, which the reference never sets */
    const int nc = m->nc;
    for (int c = 0; c < nc; c++) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < m->nf[c]; f++) {
            int j = m->nbr[4 * c + f];
            if (j < 0 || j >= nc) continue;
            double F[4];
            ausm_flux(&U_all[4 * c], &U_all[4 * j], m->nrm[8 * c + 2 * f], m->nrm[8 * c + 2 * f + 1],
                      m->len[4 * c + f], F);
            for (int k = 0; k < 4; k++) acc[k] += F[k];
        }
        for (int k = 0; k < 4; k++) R[4 * c + k] = acc[k];
    }
}

static double sign_i(double a) { return (double)((a > 0.0) - (a < 0.0)); }

/* Second_Order_Limiter (src/Limiters.cpp:97-200) with the 2-argument MinMod (:58-63) */
static void second_order_limiter(const OMesh *m, const double *U_all, double zeta, double zeta1,
                                 int c, int f, double *d_U) {
    const int nc = m->nc;
    int N1 = 0, N2 = 0;
    double d1 = 0.0, d2 = 0.0;
    if (m->nf[c] != 4) {
        N1 = m->nbr[4 * c + f];
        d1 = m->dist[4 * c + f];
        if (d1 <= 0.0) d1 = 1.0;
        N2 = N1;
        d2 = d1;
    } else {
        switch (f) {
        case 0:
        case 1:
            N1 = m->nbr[4 * c + f];
            d1 = m->dist[4 * c + f];
            N2 = (N1 >= nc) ? N1 : m->nbr[4 * N1 + f];
            d2 = (N1 >= nc) ? d1 : m->dist[4 * N1 + f];
            break;
        case 2:
        case 3:
            N1 = m->nbr[4 * c + f];
            d1 = m->dist[4 * c + f];
            N2 = m->nbr[4 * c + f - 2];
            d2 = m->dist[4 * c + f - 2];
            break;
        }
    }
    for (int k = 0; k < 4; k++) {
        double Slope1 = zeta * (U_all[4 * c + k] - U_all[4 * N1 + k]) / d1;
        double Slope2 = zeta1 * (U_all[4 * N1 + k] - U_all[4 * N2 + k]) / d2;
        double a1 = fabs(Slope1), b1 = fabs(Slope2);
        double phi = 0.5 * (sign_i(Slope1) + sign_i(Slope2)) * (b1 < a1 ? b1 : a1);
        double dR = U_all[4 * N1 + k] + 0.5 * phi * d1;
        double dL = U_all[4 * c + k] - 0.5 * phi * d1;
        d_U[k] = dR - dL;
    }
}

/* Evaluate_Cell_Net_Flux_2O with ROE_2O (src/Net_Flux.cpp:97-128, src/Roe_Scheme.cpp:257-447) */
void or_roe2o_residual(const OMesh *m, const double *U_all, const double *prim, double zeta,
                       double zeta1, double cfl, double *Rout, double *dtout) {
    const int nc = m->nc;
    for (int c = 0; c < nc; c++) {
        double net[4] = {0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < m->nf[c]; f++) {
            int j = m->nbr[4 * c + f];
            double nx = m->nrm[8 * c + 2 * f], ny = m->nrm[8 * c + 2 * f + 1], dl = m->len[4 * c + f];
            FState L, R;
            L.W[0] = prim[11 * c]; L.W[1] = prim[11 * c + 1]; L.W[2] = prim[11 * c + 2];
            L.W[3] = prim[11 * c + 4]; L.W[4] = prim[11 * c + 5];
            R.W[0] = prim[11 * j]; R.W[1] = prim[11 * j + 1]; R.W[2] = prim[11 * j + 2];
            R.W[3] = prim[11 * j + 4]; R.W[4] = prim[11 * j + 5];
            double A[4];
            avg_flux(&L, &R, nx, ny, dl, A);
            double Rho_L = L.W[0], u_L = L.W[1], v_L = L.W[2], C_L = L.W[4];
            double Vmag_L = 0.5 * (u_L * u_L + v_L * v_L);
            double H_L = ((C_L * C_L) / (GAMMA - 1.0)) + Vmag_L;
            double Rho_R = R.W[0], u_R = R.W[1], v_R = R.W[2], C_R = R.W[4];
            double Vmag_R = 0.5 * (u_R * u_R + v_R * v_R);
            double H_R = ((C_R * C_R) / (GAMMA - 1.0)) + Vmag_R;
            double sqrt_Rho_R = sqrt(Rho_R);
            double sqrt_Rho_L = sqrt(Rho_L);
            double Term1 = 1.0 / (sqrt_Rho_L + sqrt_Rho_R);
            double Roe_Rho = sqrt(Rho_L * Rho_R);
            double Roe_u = Term1 * (u_R * sqrt_Rho_R + u_L * sqrt_Rho_L);
            double Roe_v = Term1 * (v_R * sqrt_Rho_R + v_L * sqrt_Rho_L);
            double Roe_H = Term1 * (H_R * sqrt_Rho_R + H_L * sqrt_Rho_L);
            double Roe_Vmag = 0.5 * (Roe_u * Roe_u + Roe_v * Roe_v);
            double Un = nx * Roe_u + ny * Roe_v;
            double Ut = -ny * Roe_u + nx * Roe_v;
            double Roe_a = sqrt((GAMMA - 1.0) * (Roe_H - Roe_Vmag));
            double d_U[4];
            second_order_limiter(m, U_all, zeta, zeta1, c, f, d_U);
            double du = d_U[1] / Roe_Rho - d_U[0] * Roe_u / (Roe_Rho * Roe_Rho);
            double dv = d_U[2] / Roe_Rho - d_U[0] * Roe_v / (Roe_Rho * Roe_Rho);
            double dE = d_U[3] - 0.5 * Roe_Rho * (2.0 * (Roe_u * du + Roe_v * dv) + du * du + dv * dv) - 0.5 * d_U[0] * Roe_Vmag;
            double dP = (GAMMA - 1.0) * dE;
            double dRho = d_U[0];
            double dUn = nx * du + ny * dv;
            double dUt = -ny * du + nx * dv;
            double L1 = fabs(Un - Roe_a), L2 = fabs(Un), L3 = fabs(Un), L4 = fabs(Un + Roe_a);
            double ef = 0.1 * Roe_a;
            if (L1 < ef) L1 = 0.5 * (L1 * L1 / ef + ef);
            if (L4 < ef) L4 = 0.5 * (L4 * L4 / ef + ef);
            double R12 = Roe_u - nx * Roe_a, R13 = Roe_v - ny * Roe_a, R14 = Roe_H - Roe_a * Un;
            double R22 = Roe_u, R23 = Roe_v, R24 = Roe_Vmag;
            double R32 = -ny, R33 = nx, R34 = Ut;
            double R42 = Roe_u + nx * Roe_a, R43 = Roe_v + ny * Roe_a, R44 = Roe_H + Roe_a * Un;
            double Term2 = 1.0 / (Roe_a * Roe_a);
            double a1 = 0.5 * Term2 * (dP - Roe_Rho * Roe_a * dUn);
            double a2 = dRho - Term2 * dP;
            double a3 = Roe_Rho * dUt;
            double a4 = 0.5 * Term2 * (dP + Roe_Rho * Roe_a * dUn);
            double D[4];
            D[0] = 0.5 * (L1 * a1 * 1.0 + L2 * a2 * 1.0 + L3 * a3 * 0.0 + L4 * a4 * 1.0) * dl;
            D[1] = 0.5 * (L1 * a1 * R12 + L2 * a2 * R22 + L3 * a3 * R32 + L4 * a4 * R42) * dl;
            D[2] = 0.5 * (L1 * a1 * R13 + L2 * a2 * R23 + L3 * a3 * R33 + L4 * a4 * R43) * dl;
            D[3] = 0.5 * (L1 * a1 * R14 + L2 * a2 * R24 + L3 * a3 * R34 + L4 * a4 * R44) * dl;
            for (int k = 0; k < 4; k++) net[k] += A[k] - D[k];
        }
        for (int k = 0; k < 4; k++) Rout[4 * c + k] = net[k];
        double u0 = prim[11 * c + 1], v0 = prim[11 * c + 2], a0 = prim[11 * c + 5];
        double denom = 0.0;
        for (int f = 0; f < m->nf[c]; f++) {
            double vn = u0 * m->nrm[8 * c + 2 * f] + v0 * m->nrm[8 * c + 2 * f + 1];
            denom += (fabs(vn) + a0) * m->len[4 * c + f];
        }
        dtout[c] = (denom <= 0.0 || !isfinite(denom)) ? 0.0 : cfl * m->area[c] / denom;
    }
}
