// Host mesh processing under the address and undefined-behaviour sanitizers (built with g++, no
// GPU code): connectivity, tiled layouts in several shapes, the packed face records, the WENO line
// stencils, and the rejection of malformed input, on meshes given as files.
//
//   mesh_checks <mesh.bin> [<mesh.bin> ...]
// mesh.bin: int32 npts, int32 ncells, npts x 3 doubles, ncells x 4 int32 (-1 pads triangles).
// Exits 0 when every check holds; prints the failed check otherwise.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "../../cfd_solver_withcuda_amd/csrc/mesh_host.h"

using namespace cfd;

static int fails = 0;
#define CHECK(cond)                                                         \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            fails++;                                                        \
        }                                                                   \
    } while (0)

static bool read_mesh(const char *path, std::vector<double> &pts, std::vector<int> &cells) {
    std::ifstream in(path, std::ios::binary);
    int32_t np = 0, nc = 0;
    if (!in.read(reinterpret_cast<char *>(&np), 4) || !in.read(reinterpret_cast<char *>(&nc), 4)) return false;
    pts.resize(3 * (size_t)np);
    cells.resize(4 * (size_t)nc);
    in.read(reinterpret_cast<char *>(pts.data()), (std::streamsize)(pts.size() * sizeof(double)));
    in.read(reinterpret_cast<char *>(cells.data()), (std::streamsize)(cells.size() * sizeof(int)));
    return (bool)in;
}

static void check_layout(const RefMesh &m, int tile, int renumber, int budget) {
    Layout L;
    build_layout(m, tile, renumber, L, budget);
    const int nc = m.nc;
    CHECK((int)L.perm.size() == nc && (int)L.iperm.size() == nc);
    CHECK(L.tile_cstart.front() == 0 && L.tile_cstart.back() == nc);
    CHECK(L.tile_fstart.front() == 0 && L.tile_fstart.back() == (int)L.fL.size());
    for (int t = 0; t < L.ntiles; t++) {
        const int k0 = L.tile_cstart[t], k1 = L.tile_cstart[t + 1];
        const int e0 = L.tile_fstart[t], e1 = L.tile_fstart[t + 1];
        CHECK(k1 > k0 && k1 - k0 <= tile);
        if (budget) CHECK(e1 - e0 <= budget || k1 - k0 == 1);
        CHECK(e1 - e0 <= L.max_tile_faces);
        for (int k = k0; k < k1; k++) {
            const int c = L.perm[k];
            for (int f = 0; f < 4; f++) {
                const uint16_t s = L.slots[4 * (size_t)k + f];
                if (f >= m.nf[c]) {
                    CHECK(s == 0xFFFF);
                    continue;
                }
                const int e = e0 + (s >> 1);
                CHECK(e < e1);
                if (e >= e1) continue;
                const int me = (s & 1) ? L.fR[e] : L.fL[e];
                CHECK(me == k);
            }
        }
        // grouping: interior faces of the tile, then halo faces, then boundary faces
        int last = 0;
        for (int e = e0; e < e1; e++) {
            const bool bnd = L.fR[e] < 0;
            const bool inL = L.fL[e] >= k0 && L.fL[e] < k1;
            const bool inR = !bnd && L.fR[e] >= k0 && L.fR[e] < k1;
            const int grp = bnd ? 2 : (inL && inR) ? 0 : 1;
            CHECK(grp >= last);
            last = grp;
        }
    }
    // WENO line stencils: all four cells or none; each a neighbour of the previous one on the line
    std::vector<int> st;
    line_stencils(m, L, st);
    CHECK(st.size() == 4 * L.fL.size());
    auto adjacent = [&](int a, int b) {  // renumbered cells
        const int ca = L.perm[a], cb = L.perm[b];
        for (int f = 0; f < m.nf[ca]; f++)
            if (m.nbr[4 * (size_t)ca + f] == cb) return true;
        return false;
    };
    size_t with_line = 0;
    for (size_t e = 0; e < L.fL.size(); e++) {
        const int *q = &st[4 * e];
        const bool none = q[0] < 0 && q[1] < 0 && q[2] < 0 && q[3] < 0;
        bool all = true;
        for (int i = 0; i < 4; i++) all = all && q[i] >= 0 && q[i] < nc;
        CHECK(none || all);
        if (!all) continue;
        with_line++;
        CHECK(L.fR[e] >= 0);
        CHECK(adjacent(L.fL[e], q[0]) && adjacent(q[0], q[1]));
        CHECK(adjacent(L.fR[e], q[2]) && adjacent(q[2], q[3]));
    }
    CHECK(with_line > 0);
    std::vector<double> geo, w;
    std::vector<int> lr;
    pack_face_records(L, 7, geo, w, lr);
    const size_t n = L.fL.size();
    CHECK(geo.size() == 6 * n && w.size() == 4 * n && lr.size() == 2 * n);
    for (size_t e = 0; e < n; e++) {
        int32_t rec[2];
        std::memcpy(rec, &geo[6 * e], sizeof(rec));
        CHECK(rec[0] == L.fL[e] + 7 && lr[2 * e] == rec[0]);
        CHECK(rec[1] == (L.fR[e] >= 0 ? L.fR[e] + 7 : L.fR[e]) && lr[2 * e + 1] == rec[1]);
        CHECK(geo[6 * e + 1] == L.fnx[e] && geo[6 * e + 3] == L.fdl[e] && geo[6 * e + 5] == L.fdy[e]);
    }
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("usage: mesh_checks <mesh.bin> [...]\n");
        return 2;
    }
    for (int a = 1; a < argc; a++) {
        std::vector<double> pts;
        std::vector<int> cells;
        if (!read_mesh(argv[a], pts, cells)) {
            std::printf("cannot read %s\n", argv[a]);
            return 2;
        }
        const int np = (int)(pts.size() / 3), nc = (int)(cells.size() / 4);
        RefMesh m;
        build_ref_mesh(np, pts.data(), nc, cells.data(), m);
        CHECK(m.nc == nc && (int)m.gcell.size() == m.ng);
        const int shapes[5][3] = {{128, 1, 256}, {64, 1, 128}, {256, 1, 0}, {32, 0, 0}, {128, 1, 192}};
        for (const auto &s : shapes) check_layout(m, s[0], s[1], s[2]);
        // malformed input is rejected with an exception, not by reading out of bounds
        std::vector<int> bad = cells;
        bad[4 * (size_t)(nc / 2) + 1] = np + 5;  // node out of range
        bool threw = false;
        try {
            RefMesh mb;
            build_ref_mesh(np, pts.data(), nc, bad.data(), mb);
        } catch (const std::exception &) {
            threw = true;
        }
        CHECK(threw);
        std::printf("%s: %d cells, %d ghosts checked\n", argv[a], m.nc, m.ng);
    }
    return fails ? 1 : 0;
}
