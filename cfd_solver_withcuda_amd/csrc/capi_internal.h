// Internal state shared by the C-ABI translation units of libcfdhip.so (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/cfd_hip.h"
#include "cfd_types.h"
#include "mesh_host.h"

using cfd::DevMesh;
using cfd::Layout;
using cfd::RefMesh;

struct cfd_mesh_s {
    RefMesh ref;
    Layout lay;
    bool on_device = false;
    int *d_tile_cstart = nullptr, *d_tile_fstart = nullptr;
    double *d_fgeo = nullptr, *d_fw = nullptr, *d_fnx = nullptr, *d_fny = nullptr, *d_fdl = nullptr;
    int *d_fLR = nullptr;
    uint16_t *d_slots = nullptr;
    double *d_area = nullptr, *d_inv_area = nullptr;
    int *d_gcell = nullptr, *d_gkind = nullptr;  // ghost k: renumbered owner cell, face kind
    double *d_gnrm = nullptr;                    // ghost k: unit normal of the owner's face
    // the reference's per-cell arrays in renumbered order (built on first use by the reference's
    // second order)
    int *d_rc_nf = nullptr, *d_rc_nbr = nullptr;
    double *d_rc_dist = nullptr, *d_rc_nrm = nullptr, *d_rc_len = nullptr, *d_rc_area = nullptr;
    int *d_fst = nullptr;  // WENO line stencils of the face entries, 4 per entry (built on first use)
    // WENO: the tiles owning a cell of a face without a line stencil (the only gradients the
    // MUSCL fallback faces read), as a tile list for the gradient pass of the time steps
    int *d_gtiles = nullptr, *d_gcount = nullptr;
    int ngtiles = 0;
    mutable std::mutex lazy;  // guards the arrays built on first use (solvers may share the mesh)
    DevMesh dev{};
    ~cfd_mesh_s() {
        void *ps[] = {d_tile_cstart, d_tile_fstart, d_fgeo, d_fw, d_fnx, d_fny, d_fdl, d_fLR, d_slots, d_area, d_inv_area, d_gcell, d_gkind, d_gnrm,
                      d_rc_nf, d_rc_nbr, d_rc_dist, d_rc_nrm, d_rc_len, d_rc_area, d_fst, d_gtiles, d_gcount};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};


namespace cfd {
// One error message per thread (cfd_last_error).
int capi_fail(int code, const std::string &msg);
int capi_hip_fail(hipError_t e, const char *where);
// Validation of a configuration against a mesh, and the kernels' view of it.
int capi_check_config(const cfd_mesh_s *m, const cfd_config_t *c);
void capi_fill_devcfg(DevCfg &d, const cfd_config_t &c);
}  // namespace cfd
