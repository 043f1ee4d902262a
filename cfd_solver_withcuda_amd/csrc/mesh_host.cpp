// Host-side mesh processing (see mesh_host.h). Compiled without floating-point contraction so
// that geometry is bit-identical to the reference's x86-64 build.
#include "mesh_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace cfd {

namespace {

constexpr double kBboxEps = 1e-6;  // EPSILON, include/definitions.h:70

// Sort_Points_AntiClockWise(V, idx) (src/Grid_Computations.cpp:1092-1142): atan2 about the
// vertex average; already-ascending input is kept; otherwise std::sort with the reference's
// comparator, which for <= 16 elements is a stable insertion sort.
void sort_ccw(int n, double *v, int *idx) {
    double x = 0.0, y = 0.0;
    for (int i = 0; i < n; i++) {
        x += v[3 * i];
        y += v[3 * i + 1];
    }
    const double cx = x / n, cy = y / n;
    bool sorted = true;
    for (int i = 1; i < n && sorted; i++) {
        double ap = std::atan2(v[3 * (i - 1) + 1] - cy, v[3 * (i - 1)] - cx);
        double ac = std::atan2(v[3 * i + 1] - cy, v[3 * i] - cx);
        if (ap > ac) sorted = false;
    }
    if (sorted) return;
    auto less = [&](const double *A, const double *B) {
        double aA = std::atan2(A[1] - cy, A[0] - cx);
        double aB = std::atan2(B[1] - cy, B[0] - cx);
        if (std::fabs(aA - aB) < 1e-9) return A[2] < B[2];
        return aA < aB;
    };
    for (int i = 1; i < n; i++) {
        double tv[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        int ti = idx[i];
        int j = i;
        while (j > 0 && less(tv, &v[3 * (j - 1)])) {
            std::memcpy(&v[3 * j], &v[3 * (j - 1)], 3 * sizeof(double));
            idx[j] = idx[j - 1];
            j--;
        }
        std::memcpy(&v[3 * j], tv, 3 * sizeof(double));
        idx[j] = ti;
    }
}

struct Edge {
    int a, b, cell, slot;
};

uint64_t hilbert_d(uint32_t x, uint32_t y, int order) {
    uint64_t d = 0;
    for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
        uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

}  // namespace

void build_ref_mesh(int npts, const double *pts, int nc, const int *cells4, RefMesh &m) {
    m = RefMesh();
    m.np = npts;
    m.nc = nc;
    m.nf.assign(nc, 0);
    m.nodes.assign(4 * (size_t)nc, -1);
    m.nbr.assign(4 * (size_t)nc, -1);
    m.kind.assign(4 * (size_t)nc, 0);
    m.vtx.assign(12 * (size_t)nc, 0.0);
    m.nrm.assign(8 * (size_t)nc, 0.0);
    m.len.assign(4 * (size_t)nc, 0.0);
    m.dist.assign(4 * (size_t)nc, 0.0);
    std::vector<double> ctr(3 * (size_t)nc), area(nc), inv(nc);
    for (int c = 0; c < nc; c++) {
        const int n = cells4[4 * c + 3] < 0 ? 3 : 4;
        m.nf[c] = n;
        double *v = &m.vtx[12 * (size_t)c];
        int *id = &m.nodes[4 * (size_t)c];
        for (int i = 0; i < n; i++) {
            id[i] = cells4[4 * c + i];
            if (id[i] < 0 || id[i] >= npts) throw std::runtime_error("cell node index out of range");
            v[3 * i] = pts[3 * id[i]];
            v[3 * i + 1] = pts[3 * id[i] + 1];
            v[3 * i + 2] = pts[3 * id[i] + 2];
        }
        sort_ccw(n, v, id);
        double x = 0.0, y = 0.0, z = 0.0;  // Compute_Centroid (:957-980)
        for (int i = 0; i < n; i++) {
            x += v[3 * i];
            y += v[3 * i + 1];
            z += v[3 * i + 2];
        }
        ctr[3 * c] = x / n;
        ctr[3 * c + 1] = y / n;
        ctr[3 * c + 2] = z / n;
        for (int f = 0; f < n; f++) {  // Construct_Face (:691-752)
            const double *a = &v[3 * f], *b = &v[3 * ((f + 1) % n)];
            double dy = b[1] - a[1];
            double dx = b[0] - a[0];
            double dL = std::sqrt(dx * dx + dy * dy);
            if (dL == 0.0) throw std::runtime_error("zero-length face");
            m.len[4 * c + f] = dL;
            m.nrm[8 * c + 2 * f] = dy / dL;
            m.nrm[8 * c + 2 * f + 1] = -dx / dL;
        }
        double twice = 0.0;  // shoelace (:541-557)
        for (int i = 0; i < n; i++) {
            int j = (i + 1) % n;
            twice += (v[3 * i] * v[3 * j + 1] - v[3 * j] * v[3 * i + 1]);
        }
        area[c] = 0.5 * std::fabs(twice);
        if (!(area[c] > 0.0) || !std::isfinite(area[c])) throw std::runtime_error("invalid cell area");
        inv[c] = 1.0 / area[c];
    }
    // mapFacesToCells + fillFaceOrderedNeighbours (src/Read_Gmsh_File.cpp:325-402)
    std::vector<Edge> E;
    E.reserve(4 * (size_t)nc);
    for (int c = 0; c < nc; c++) {
        const int n = m.nf[c];
        for (int f = 0; f < n; f++) {
            int a = m.nodes[4 * c + f], b = m.nodes[4 * c + (f + 1) % n];
            if (a > b) std::swap(a, b);
            E.push_back({a, b, c, f});
        }
    }
    std::sort(E.begin(), E.end(), [](const Edge &p, const Edge &q) {
        if (p.a != q.a) return p.a < q.a;
        if (p.b != q.b) return p.b < q.b;
        if (p.cell != q.cell) return p.cell < q.cell;
        return p.slot < q.slot;
    });
    for (size_t i = 0; i < E.size();) {
        size_t j = i;
        while (j < E.size() && E[j].a == E[i].a && E[j].b == E[i].b) j++;
        int distinct = 0, c0 = -1, c1 = -1;
        for (size_t q = i; q < j; q++)
            if (E[q].cell != c0 && E[q].cell != c1) {
                if (distinct == 0) c0 = E[q].cell;
                else if (distinct == 1) c1 = E[q].cell;
                distinct++;
            }
        if (distinct == 2)
            for (size_t q = i; q < j; q++) m.nbr[4 * (size_t)E[q].cell + E[q].slot] = (E[q].cell == c0) ? c1 : c0;
        i = j;
    }
    int ghost = nc;
    std::vector<char> hasb(nc, 0);
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m.nf[c]; f++)
            if (m.nbr[4 * c + f] < 0) {
                m.nbr[4 * c + f] = ghost++;
                hasb[c] = 1;
            }
    const int ng = ghost - nc;
    m.ng = ng;
    // BoundingBox (:441-472)
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    for (int i = 0; i < npts; i++) {
        double x = pts[3 * i], y = pts[3 * i + 1];
        if (i == 0) {
            xmin = xmax = x;
            ymin = ymax = y;
        } else {
            xmin = x < xmin ? x : xmin;
            xmax = x > xmax ? x : xmax;
            ymin = y < ymin ? y : ymin;
            ymax = y > ymax ? y : ymax;
        }
    }
    // Classify_Domain_Boundaries (:490-577)
    for (int c = 0; c < nc; c++) {
        const int n = m.nf[c];
        for (int e = 0; e < n; e++) {
            const double *p1 = &m.vtx[12 * c + 3 * e], *p2 = &m.vtx[12 * c + 3 * ((e + 1) % n)];
            int kd = 0;
            if (hasb[c]) {
                if (std::fabs(p1[0] - xmin) < kBboxEps && std::fabs(p2[0] - xmin) < kBboxEps) kd = 1;
                else if (std::fabs(p1[0] - xmax) < kBboxEps && std::fabs(p2[0] - xmax) < kBboxEps) kd = 2;
                else if (std::fabs(p1[1] - ymin) < kBboxEps && std::fabs(p2[1] - ymin) < kBboxEps) kd = 4;
                else if (std::fabs(p1[1] - ymax) < kBboxEps && std::fabs(p2[1] - ymax) < kBboxEps) kd = 3;
                else kd = 5;
            }
            m.kind[4 * c + e] = kd;
        }
    }
    // Create_Boundary_Cells_Lists (:579-621)
    m.gcell.assign(ng, 0);
    m.gface.assign(ng, 0);
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m.nf[c]; f++) {
            int g = m.nbr[4 * c + f];
            if (g < nc) continue;
            m.gcell[g - nc] = c;
            m.gface[g - nc] = f;
            int kd = m.kind[4 * c + f];
            std::vector<int> *L = kd == 1 ? &m.inl : kd == 2 ? &m.exl : (kd >= 3 ? &m.wal : nullptr);
            if (L) {
                L->push_back(c);
                L->push_back(f);
                L->push_back(g);
            }
        }
    // Construct_Ghost_Cells (:394-442, :589-629)
    const size_t nt = (size_t)nc + ng;
    m.ctr.assign(3 * nt, 0.0);
    m.area.assign(nt, 0.0);
    m.inv_area.assign(nt, 0.0);
    std::copy(ctr.begin(), ctr.end(), m.ctr.begin());
    std::copy(area.begin(), area.end(), m.area.begin());
    std::copy(inv.begin(), inv.end(), m.inv_area.begin());
    for (int gi = 0; gi < ng; gi++) {
        const int c = m.gcell[gi], f = m.gface[gi], g = nc + gi, nv = m.nf[c];
        const double *a = &m.vtx[12 * c + 3 * (f % nv)], *b = &m.vtx[12 * c + 3 * ((f + 1) % nv)];
        for (int q = 0; q < 3; q++) {
            double mid = 0.5 * (a[q] + b[q]);
            m.ctr[3 * (size_t)g + q] = m.ctr[3 * c + q] + 2.0 * (mid - m.ctr[3 * c + q]);
        }
        m.area[g] = m.area[c];
        m.inv_area[g] = m.inv_area[c];
    }
    // Calculate_Cell_Center_Distances (:802-827)
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m.nf[c]; f++) {
            const int j = m.nbr[4 * c + f];
            double D = 0.0;
            for (int q = 0; q < 3; q++) {
                double t = m.ctr[3 * (size_t)j + q] - m.ctr[3 * c + q];
                D += t * t;
            }
            m.dist[4 * c + f] = std::sqrt(D);
        }
}

void wlsq_coefficients(const RefMesh &m, std::vector<double> &CX, std::vector<double> &CY) {
    CX.assign(4 * (size_t)m.nc, 0.0);
    CY.assign(4 * (size_t)m.nc, 0.0);
    for (int c = 0; c < m.nc; c++) {
        double a = 0.0, b = 0.0, d = 0.0;
        const int n = m.nf[c];
        double dx[4], dy[4], w[4];
        for (int f = 0; f < n; f++) {
            const int j = m.nbr[4 * c + f];
            dx[f] = m.ctr[3 * (size_t)j] - m.ctr[3 * c];
            dy[f] = m.ctr[3 * (size_t)j + 1] - m.ctr[3 * c + 1];
            w[f] = 1.0 / (dx[f] * dx[f] + dy[f] * dy[f]);
            a += w[f] * dx[f] * dx[f];
            b += w[f] * dx[f] * dy[f];
            d += w[f] * dy[f] * dy[f];
        }
        const double det = a * d - b * b;
        const double i11 = d / det, i12 = -b / det, i22 = a / det;
        for (int f = 0; f < n; f++) {
            CX[4 * c + f] = (i11 * dx[f] + i12 * dy[f]) * w[f];
            CY[4 * c + f] = (i12 * dx[f] + i22 * dy[f]) * w[f];
        }
    }
}

void build_layout(const RefMesh &m, int tile, int renumber, Layout &L, int face_budget) {
    if (tile < 16 || tile > 4096) throw std::runtime_error("tile size out of range [16, 4096]");
    if (face_budget != 0 && face_budget < 4)
        throw std::runtime_error("face budget must be 0 (none) or at least 4 faces");
    L = Layout();
    const int nc = m.nc;
    L.tile = tile;
    L.face_budget = face_budget;
    L.perm.resize(nc);
    std::iota(L.perm.begin(), L.perm.end(), 0);
    if (renumber == 1 && nc > 0) {
        double xmin = m.ctr[0], xmax = m.ctr[0], ymin = m.ctr[1], ymax = m.ctr[1];
        for (int c = 0; c < nc; c++) {
            xmin = std::min(xmin, m.ctr[3 * c]);
            xmax = std::max(xmax, m.ctr[3 * c]);
            ymin = std::min(ymin, m.ctr[3 * c + 1]);
            ymax = std::max(ymax, m.ctr[3 * c + 1]);
        }
        const double span = std::max(xmax - xmin, ymax - ymin);
        const double s = span > 0 ? 65535.0 / span : 0.0;
        std::vector<uint64_t> key(nc);
        for (int c = 0; c < nc; c++) {
            uint32_t x = (uint32_t)std::lround((m.ctr[3 * c] - xmin) * s);
            uint32_t y = (uint32_t)std::lround((m.ctr[3 * c + 1] - ymin) * s);
            key[c] = hilbert_d(x, y, 16);
        }
        std::stable_sort(L.perm.begin(), L.perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    L.iperm.assign(nc, 0);
    for (int k = 0; k < nc; k++) L.iperm[L.perm[k]] = k;
    L.area.resize(nc);
    L.inv_area.resize(nc);
    for (int k = 0; k < nc; k++) {
        L.area[k] = m.area[L.perm[k]];
        L.inv_area[k] = m.inv_area[L.perm[k]];
    }
    // unique faces: canonical side = lower original index (or the owner of a boundary face)
    std::vector<int> face_of(4 * (size_t)nc, -1), canon_cell, canon_slot;
    for (int c = 0; c < nc; c++)
        for (int f = 0; f < m.nf[c]; f++) {
            const int j = m.nbr[4 * c + f];
            if (j < nc && j < c) continue;
            const int id = (int)canon_cell.size();
            canon_cell.push_back(c);
            canon_slot.push_back(f);
            face_of[4 * c + f] = id;
            if (j < nc)
                for (int s = 0; s < m.nf[j]; s++)
                    if (m.nbr[4 * j + s] == c) face_of[4 * (size_t)j + s] = id;
        }
    L.nfaces = (int)canon_cell.size();
    std::vector<double> CX, CY;
    wlsq_coefficients(m, CX, CY);
    // tile boundaries: contiguous cells, at most `tile` cells and `face_budget` faces each
    std::vector<int> stamp(L.nfaces, -1), local(L.nfaces, 0);
    L.tile_cstart.clear();
    {
        int k = 0, t = 0;
        while (k < nc) {
            L.tile_cstart.push_back(k);
            int ncell = 0, nface = 0;
            while (k < nc && ncell < tile) {
                const int c = L.perm[k];
                int fresh = 0;
                for (int f = 0; f < m.nf[c]; f++)
                    if (stamp[face_of[4 * (size_t)c + f]] != t) fresh++;
                if (face_budget && ncell > 0 && nface + fresh > face_budget) break;
                for (int f = 0; f < m.nf[c]; f++) stamp[face_of[4 * (size_t)c + f]] = t;
                nface += fresh;
                ncell++;
                k++;
            }
            t++;
        }
        L.tile_cstart.push_back(nc);
        std::fill(stamp.begin(), stamp.end(), -1);
    }
    L.ntiles = (int)L.tile_cstart.size() - 1;
    L.tile_fstart.assign(L.ntiles + 1, 0);
    L.slots.assign(4 * (size_t)nc, 0xFFFF);
    std::vector<int> ids, grp(L.nfaces, 0);
    for (int t = 0; t < L.ntiles; t++) {
        L.tile_fstart[t] = (int)L.fL.size();
        const int k0 = L.tile_cstart[t], k1 = L.tile_cstart[t + 1];
        // the tile's faces in first-touch order, then grouped: faces between two cells of the
        // tile, faces to a cell of another tile (halo), boundary faces. Lanes of one wave then
        // mostly take the same branches in the flux and gradient passes.
        ids.clear();
        for (int k = k0; k < k1; k++) {
            const int c = L.perm[k];
            for (int f = 0; f < m.nf[c]; f++) {
                const int id = face_of[4 * (size_t)c + f];
                if (stamp[id] == t) continue;
                stamp[id] = t;
                const int a = canon_cell[id], j = m.nbr[4 * (size_t)a + canon_slot[id]];
                if (j >= nc) {
                    grp[id] = 2;
                } else {
                    const int kl = L.iperm[a], kr = L.iperm[j];
                    grp[id] = (kl >= k0 && kl < k1 && kr >= k0 && kr < k1) ? 0 : 1;
                }
                ids.push_back(id);
            }
        }
        std::stable_sort(ids.begin(), ids.end(), [&](int p, int q) { return grp[p] < grp[q]; });
        const int nloc = (int)ids.size();
        for (int i = 0; i < nloc; i++) {
            const int id = ids[i];
            local[id] = i;
            const int a = canon_cell[id], s = canon_slot[id];
            const int j = m.nbr[4 * (size_t)a + s];
            L.fL.push_back(L.iperm[a]);
            L.fR.push_back(j < nc ? L.iperm[j] : -m.kind[4 * (size_t)a + s]);
            L.fnx.push_back(m.nrm[8 * (size_t)a + 2 * s]);
            L.fny.push_back(m.nrm[8 * (size_t)a + 2 * s + 1]);
            L.fdl.push_back(m.len[4 * (size_t)a + s]);
            L.fdx.push_back(m.ctr[3 * (size_t)j] - m.ctr[3 * (size_t)a]);
            L.fdy.push_back(m.ctr[3 * (size_t)j + 1] - m.ctr[3 * (size_t)a + 1]);
            L.fwL.push_back(CX[4 * (size_t)a + s]);
            L.fwL.push_back(CY[4 * (size_t)a + s]);
            double rx = 0.0, ry = 0.0;
            if (j < nc)
                for (int q = 0; q < m.nf[j]; q++)
                    if (m.nbr[4 * (size_t)j + q] == a) {
                        rx = CX[4 * (size_t)j + q];
                        ry = CY[4 * (size_t)j + q];
                    }
            L.fwR.push_back(rx);
            L.fwR.push_back(ry);
        }
        if (nloc > 32767) throw std::runtime_error("too many faces in one tile");
        for (int k = k0; k < k1; k++) {
            const int c = L.perm[k];
            for (int f = 0; f < m.nf[c]; f++) {
                const int id = face_of[4 * (size_t)c + f];
                const bool isR = canon_cell[id] != c;
                L.slots[4 * (size_t)k + f] = (uint16_t)((local[id] << 1) | (isR ? 1 : 0));
            }
        }
        L.max_tile_faces = std::max(L.max_tile_faces, nloc);
    }
    L.tile_fstart[L.ntiles] = (int)L.fL.size();
}

// The physical cell across the face opposite slot f of quadrilateral c, or -1.
static int cell_behind(const RefMesh &m, int c, int f) {
    if (c < 0 || f < 0 || m.nf[c] != 4) return -1;
    const int j = m.nbr[4 * (size_t)c + (f + 2) % 4];
    return (j >= 0 && j < m.nc) ? j : -1;
}

// Slot of cell c whose neighbour is nb, or -1.
static int slot_to(const RefMesh &m, int c, int nb) {
    if (c < 0) return -1;
    for (int f = 0; f < m.nf[c]; f++)
        if (m.nbr[4 * (size_t)c + f] == nb) return f;
    return -1;
}

void line_stencils(const RefMesh &m, const Layout &L, std::vector<int> &st) {
    const size_t n = L.fL.size();
    st.assign(4 * n, -1);
    for (size_t e = 0; e < n; e++) {
        if (L.fR[e] < 0) continue;
        const int a = L.perm[L.fL[e]], b = L.perm[L.fR[e]];
        const int l1 = cell_behind(m, a, slot_to(m, a, b));
        const int l2 = cell_behind(m, l1, slot_to(m, l1, a));
        const int r1 = cell_behind(m, b, slot_to(m, b, a));
        const int r2 = cell_behind(m, r1, slot_to(m, r1, b));
        if (l1 < 0 || l2 < 0 || r1 < 0 || r2 < 0) continue;
        st[4 * e] = L.iperm[l1];
        st[4 * e + 1] = L.iperm[l2];
        st[4 * e + 2] = L.iperm[r1];
        st[4 * e + 3] = L.iperm[r2];
    }
}

void pack_face_records(const Layout &L, int cell_off, std::vector<double> &geo, std::vector<double> &w,
                       std::vector<int> &lr) {
    const size_t n = L.fL.size();
    geo.resize(6 * n);
    w.resize(4 * n);
    lr.resize(2 * n);
    for (size_t e = 0; e < n; e++) {
        lr[2 * e] = L.fL[e] + cell_off;
        lr[2 * e + 1] = L.fR[e] >= 0 ? L.fR[e] + cell_off : L.fR[e];
        std::memcpy(&geo[6 * e], &lr[2 * e], 2 * sizeof(int));
        geo[6 * e + 1] = L.fnx[e];
        geo[6 * e + 2] = L.fny[e];
        geo[6 * e + 3] = L.fdl[e];
        geo[6 * e + 4] = L.fdx[e];
        geo[6 * e + 5] = L.fdy[e];
        w[4 * e] = L.fwL[2 * e];
        w[4 * e + 1] = L.fwL[2 * e + 1];
        w[4 * e + 2] = L.fwR[2 * e];
        w[4 * e + 3] = L.fwR[2 * e + 1];
    }
}

}  // namespace cfd
