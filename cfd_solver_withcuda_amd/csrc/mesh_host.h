// Host-side mesh processing: the reference's connectivity and geometry, and the tiled,
// renumbered layout the kernels use.
#pragma once
#include <cstdint>
#include <vector>

namespace cfd {

// Connectivity and geometry in the reference's cell order and layout (4 face slots per cell),
// as Read_VTK_Grid + Form_Cells produce them (src/Read_Gmsh_File.cpp:11-621,
// src/Grid_Computations.cpp:319-827,957-1142).
struct RefMesh {
    int np = 0, nc = 0, ng = 0;
    std::vector<int> nf, nodes, nbr, kind;        // [nc], [4nc] x3
    std::vector<double> vtx;                      // [12nc] CCW-sorted vertices
    std::vector<double> nrm, len, dist;           // [8nc], [4nc], [4nc]
    std::vector<double> ctr, area, inv_area;      // [3(nc+ng)], [nc+ng], [nc+ng]
    std::vector<int> inl, exl, wal;               // (cell, face, ghost) triplets
    std::vector<int> gcell, gface;                // per ghost: owner cell and face
};

void build_ref_mesh(int npts, const double *pts3, int ncells, const int *cells4, RefMesh &m);

// Least-squares coefficient vectors c_{cell,slot} = M^-1 w d (weights 1/|d|^2, ghost
// neighbours at the reflected centre): grad W_c = sum_slot c (W_n - W_c).
void wlsq_coefficients(const RefMesh &m, std::vector<double> &cx, std::vector<double> &cy);

// Tiled layout on the renumbered cells. A tile is a contiguous range of at most `tile` cells
// whose faces number at most `face_budget` (0 = no face limit).
struct Layout {
    int tile = 256;
    int face_budget = 0;
    int ntiles = 0;
    int max_tile_faces = 0;
    int nfaces = 0;                     // unique faces
    std::vector<int> perm, iperm;       // new -> old, old -> new
    std::vector<int> tile_cstart;       // [ntiles+1]
    std::vector<int> tile_fstart;       // [ntiles+1]
    std::vector<int> fL, fR;            // per face entry (renumbered; fR = -kind at boundary)
    std::vector<double> fnx, fny, fdl, fdx, fdy;
    std::vector<double> fwL, fwR;       // [2*nfe]
    std::vector<uint16_t> slots;        // [4*nc]
    std::vector<double> area, inv_area; // renumbered
};

// renumber: 0 = keep the input order, 1 = Hilbert order of the cell centres.
void build_layout(const RefMesh &m, int tile, int renumber, Layout &L, int face_budget = 0);

// The device records of the layout's face entries, with cell ids shifted by `cell_off` (boundary
// markers kept): geo gets 6 doubles per entry ({L, R} as two int32 in the first, then nx, ny, dl,
// dx, dy), w gets 4 (the least-squares vectors of L, then of R), lr gets {L, R}.
void pack_face_records(const Layout &L, int cell_off, std::vector<double> &geo, std::vector<double> &w,
                       std::vector<int> &lr);

// WENO line stencils of the layout's face entries: st gets 4 renumbered cells per entry,
// {L-1, L-2, R+1, R+2}, where L-1 is the neighbour of L across the face opposite the entry's face
// (L a quadrilateral), L-2 the neighbour of L-1 across the face opposite the one it shares with L,
// and likewise on the R side; all four are -1 when the entry is a boundary face or any of them is
// missing (a triangle on the line, or a ghost).
void line_stencils(const RefMesh &m, const Layout &L, std::vector<int> &st);

}  // namespace cfd
