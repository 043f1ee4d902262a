// C-ABI of the native path (include/cfd_hip.h): mesh handles, solver state, residual, steps.
#include "../../include/cfd_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "fv_kernels.h"
#include "mesh_host.h"

using namespace cfd;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *where) {
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    return (int)e;
}

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

template <typename T>
hipError_t upload(T **d, const std::vector<T> &h) {
    *d = nullptr;
    if (h.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void **)d, h.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

struct cfd_solver_s;
static int freeze_init(cfd_solver_s *s, int mode);

namespace cfd {
// Error reporting shared with the other C-ABI translation units (one message per thread).
int capi_fail(int code, const std::string &msg) { return fail(code, msg); }
int capi_hip_fail(hipError_t e, const char *where) { return hip_fail(e, where); }

// Viscous settings shared by single and batched solvers: validation and the device copy.
int capi_check_viscous(const cfd_config_t *c, const char *who) {
    if (c->viscous < 0 || c->viscous > 2) return fail(CFD_E_ARG, std::string(who) + ": viscous must be 0, 1 or 2");
    if (!c->viscous) return 0;
    if (c->order < 2) return fail(CFD_E_ARG, std::string(who) + ": viscous terms need order 2 or 3 (cell gradients)");
    if (c->ref_mode) return fail(CFD_E_ARG, std::string(who) + ": viscous terms exist in the native iteration only");
    if (!(c->prandtl > 0.0) || !std::isfinite(c->prandtl)) return fail(CFD_E_ARG, std::string(who) + ": prandtl must be positive");
    if (!(c->r_gas > 0.0) || !std::isfinite(c->r_gas)) return fail(CFD_E_ARG, std::string(who) + ": r_gas must be positive");
    if (c->viscous == 1 && (!(c->mu >= 0.0) || !std::isfinite(c->mu)))
        return fail(CFD_E_ARG, std::string(who) + ": mu must be finite and >= 0");
    return 0;
}

void capi_fill_viscous(DevCfg &d, const cfd_config_t &c) {
    d.viscous = c.viscous;
    d.mu = c.viscous ? c.mu : 0.0;
    d.r_gas = c.viscous ? c.r_gas : 1.0;
    d.kfac = c.viscous ? 1.4 * c.r_gas / ((1.4 - 1.0) * c.prandtl) : 0.0;
    d.inv_pr = c.viscous ? 1.0 / c.prandtl : 0.0;
}
}  // namespace cfd

struct cfd_solver_s {
    cfd_mesh_s *m = nullptr;
    // the solver's own copy of the mesh's device view, taken under the mesh's lock when the
    // solver is configured (other solvers sharing the mesh may add lazily built arrays)
    DevMesh dev{};
    int ngtiles = 0;
    cfd_config_t cfg{};
    DevCfg dcfg{};
    double4 *Un = nullptr, *Ua = nullptr, *Ub = nullptr, *acc = nullptr, *R = nullptr;
    // positivity fallback: the last stage's output (swapped with Un after the step), the second
    // RK4 accumulator, per-cell flags and per-tile marks (allocated on first use)
    double4 *Uc = nullptr, *acc2 = nullptr;
    uint8_t *pflag = nullptr;
    int *dirty = nullptr;
    unsigned pos_seq = 0;  // detecting launches so far: picks the marked-tile count of the next one
    double *dt = nullptr, *grad = nullptr, *hbuf = nullptr, *err_partial = nullptr, *dt_partial = nullptr;
    double *scal = nullptr;  // [0],[1] global dt slots, [2] stage minimum, [3..6] error norm
    double4 *Ug = nullptr;   // ghost states (reference ghost order), allocated on first use
    float *psi = nullptr;    // [8 * face entries] frozen limiter factors, allocated on first use
    double4 *Uall = nullptr; // [cells + ghosts] state for the reference's second order
    int *fz = nullptr;       // [0] freezing state, [1] steps since the best Error[0]
    double *fz_best = nullptr;
    bool fz_explicit = false;  // cfd_set_freeze chose a non-zero mode
    int cur = 0;
    bool dt_valid = false;
    bool has_state = false;
    ~cfd_solver_s() {
        void *ps[] = {Un, Ua, Ub, acc, R, dt, grad, hbuf, err_partial, dt_partial, scal, Ug, psi, fz, fz_best, Uall,
                      Uc, acc2, pflag, dirty};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};

extern "C" {

const char *cfd_last_error(void) { return g_err.c_str(); }
int cfd_version(void) { return 1; }

int cfd_device_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(CFD_E_ARG, "cfd_device_alloc: null argument");
    *ptr = nullptr;
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}
int cfd_device_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return 0;
}
int cfd_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int cfd_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int cfd_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

int cfd_mesh_create_host(int npts, const double *pts3, int ncells, const int *cells4, int tile,
                         int face_budget, int renumber, cfd_mesh_t **out) {
    if (!out || !pts3 || !cells4 || npts <= 0 || ncells <= 0) return fail(CFD_E_ARG, "cfd_mesh_create: bad arguments");
    *out = nullptr;
    cfd_mesh_s *m = new cfd_mesh_s();
    try {
        build_ref_mesh(npts, pts3, ncells, cells4, m->ref);
        build_layout(m->ref, tile, renumber, m->lay, face_budget);
    } catch (const std::exception &e) {
        delete m;
        return fail(CFD_E_MESH, std::string("cfd_mesh_create: ") + e.what());
    }
    *out = m;
    return 0;
}

int cfd_mesh_create(int npts, const double *pts3, int ncells, const int *cells4, int tile,
                    int face_budget, int renumber, cfd_mesh_t **out) {
    int rc = cfd_mesh_create_host(npts, pts3, ncells, cells4, tile, face_budget, renumber, out);
    if (rc) return rc;
    cfd_mesh_s *m = *out;
    *out = nullptr;
    const Layout &L = m->lay;
    hipError_t e = hipSuccess;
    std::vector<double> geo, w;
    std::vector<int> lr;
    pack_face_records(L, 0, geo, w, lr);
    if ((e = upload(&m->d_tile_cstart, L.tile_cstart)) != hipSuccess ||
        (e = upload(&m->d_tile_fstart, L.tile_fstart)) != hipSuccess ||
        (e = upload(&m->d_fgeo, geo)) != hipSuccess || (e = upload(&m->d_fw, w)) != hipSuccess ||
        (e = upload(&m->d_fLR, lr)) != hipSuccess || (e = upload(&m->d_fnx, L.fnx)) != hipSuccess ||
        (e = upload(&m->d_fny, L.fny)) != hipSuccess || (e = upload(&m->d_fdl, L.fdl)) != hipSuccess ||
        (e = upload(&m->d_slots, L.slots)) != hipSuccess ||
        (e = upload(&m->d_area, L.area)) != hipSuccess || (e = upload(&m->d_inv_area, L.inv_area)) != hipSuccess) {
        delete m;
        return hip_fail(e, "cfd_mesh_create upload");
    }
    {
        const RefMesh &R = m->ref;
        std::vector<int> gc(R.ng), gk(R.ng);
        std::vector<double> gn(2 * (size_t)R.ng);
        for (int k = 0; k < R.ng; k++) {
            const int c = R.gcell[k], f = R.gface[k];
            gc[k] = L.iperm[c];
            gk[k] = R.kind[4 * c + f];
            gn[2 * k] = R.nrm[8 * c + 2 * f];
            gn[2 * k + 1] = R.nrm[8 * c + 2 * f + 1];
        }
        if ((e = upload(&m->d_gcell, gc)) != hipSuccess || (e = upload(&m->d_gkind, gk)) != hipSuccess ||
            (e = upload(&m->d_gnrm, gn)) != hipSuccess) {
            delete m;
            return hip_fail(e, "cfd_mesh_create upload");
        }
    }
    DevMesh &d = m->dev;
    d.nc = m->ref.nc;
    d.ntiles = L.ntiles;
    d.tile = L.tile;
    d.max_tile_faces = L.max_tile_faces;
    d.nfe = (int)L.fL.size();
    d.tile_cstart = m->d_tile_cstart;
    d.tile_fstart = m->d_tile_fstart;
    d.fgeo = reinterpret_cast<const double2 *>(m->d_fgeo);
    d.fw = reinterpret_cast<const double2 *>(m->d_fw);
    d.fLR = reinterpret_cast<const int2 *>(m->d_fLR);
    d.fnx = m->d_fnx;
    d.fny = m->d_fny;
    d.fdl = m->d_fdl;
    d.slots = m->d_slots;
    d.area = m->d_area;
    d.inv_area = m->d_inv_area;
    m->on_device = true;
    *out = m;
    return 0;
}

void cfd_mesh_destroy(cfd_mesh_t *m) { delete m; }

int cfd_mesh_counts(const cfd_mesh_t *m, int *o) {
    if (!m || !o) return fail(CFD_E_ARG, "cfd_mesh_counts: null argument");
    o[0] = m->ref.nc;
    o[1] = m->ref.ng;
    o[2] = (int)m->ref.inl.size() / 3;
    o[3] = (int)m->ref.exl.size() / 3;
    o[4] = (int)m->ref.wal.size() / 3;
    o[5] = m->lay.nfaces;
    o[6] = m->lay.ntiles;
    o[7] = (int)m->lay.fL.size();
    o[8] = m->lay.max_tile_faces;
    return 0;
}

int cfd_mesh_export(const cfd_mesh_t *m, int *nf, int *nodes4, int *nbr4, double *nrm8,
                    double *len4, double *ctr3_all, double *area_all, double *inv_area_all,
                    double *dist4, int *kind4, int *inlet3, int *exit3, int *wall3) {
    if (!m) return fail(CFD_E_ARG, "cfd_mesh_export: null mesh");
    const RefMesh &r = m->ref;
    const size_t nc = r.nc, nt = (size_t)r.nc + r.ng;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(nf, r.nf.data(), nc * sizeof(int));
    for (size_t c = 0; c < nc; c++)
        for (int f = 0; f < r.nf[c]; f++) {
            if (nodes4) nodes4[4 * c + f] = r.nodes[4 * c + f];
            if (nbr4) nbr4[4 * c + f] = r.nbr[4 * c + f];
            if (nrm8) {
                nrm8[8 * c + 2 * f] = r.nrm[8 * c + 2 * f];
                nrm8[8 * c + 2 * f + 1] = r.nrm[8 * c + 2 * f + 1];
            }
            if (len4) len4[4 * c + f] = r.len[4 * c + f];
            if (dist4) dist4[4 * c + f] = r.dist[4 * c + f];
            if (kind4) kind4[4 * c + f] = r.kind[4 * c + f];
        }
    cp(ctr3_all, r.ctr.data(), 3 * nt * sizeof(double));
    cp(area_all, r.area.data(), nt * sizeof(double));
    cp(inv_area_all, r.inv_area.data(), nt * sizeof(double));
    cp(inlet3, r.inl.data(), r.inl.size() * sizeof(int));
    cp(exit3, r.exl.data(), r.exl.size() * sizeof(int));
    cp(wall3, r.wal.data(), r.wal.size() * sizeof(int));
    return 0;
}

int cfd_mesh_permutation(const cfd_mesh_t *m, int *new_to_old) {
    if (!m || !new_to_old) return fail(CFD_E_ARG, "cfd_mesh_permutation: null argument");
    std::memcpy(new_to_old, m->lay.perm.data(), m->lay.perm.size() * sizeof(int));
    return 0;
}

int cfd_mesh_layout_export(const cfd_mesh_t *m, int *tile_cstart, int *tile_fstart, int *fL,
                           int *fR, unsigned short *slots4) {
    if (!m) return fail(CFD_E_ARG, "cfd_mesh_layout_export: null mesh");
    const Layout &L = m->lay;
    if (tile_cstart) std::memcpy(tile_cstart, L.tile_cstart.data(), L.tile_cstart.size() * sizeof(int));
    if (tile_fstart) std::memcpy(tile_fstart, L.tile_fstart.data(), L.tile_fstart.size() * sizeof(int));
    if (fL) std::memcpy(fL, L.fL.data(), L.fL.size() * sizeof(int));
    if (fR) std::memcpy(fR, L.fR.data(), L.fR.size() * sizeof(int));
    if (slots4) std::memcpy(slots4, L.slots.data(), L.slots.size() * sizeof(uint16_t));
    return 0;
}


static int check_config(const cfd_mesh_s *m, const cfd_config_t *c) {
    if (c->flux < 0 || c->flux > 3) return fail(CFD_E_ARG, "config: flux must be 0..3");
    if (c->order < 1 || c->order > 3) return fail(CFD_E_ARG, "config: order must be 1, 2 or 3 (WENO5)");
    if (c->order == 3 && (c->freeze_tol > 0.0 || c->freeze_stall > 0))
        return fail(CFD_E_ARG, "config: limiter freezing needs order 2");
    if (c->limiter < 0 || c->limiter > 4) return fail(CFD_E_ARG, "config: limiter must be 0..4");
    if (c->gradient < 0 || c->gradient > 1) return fail(CFD_E_ARG, "config: gradient must be 0 or 1");
    if (!(c->cfl > 0.0)) return fail(CFD_E_ARG, "config: cfl must be positive");
    if (!(c->freeze_tol >= 0.0) || c->freeze_stall < 0) return fail(CFD_E_ARG, "config: freeze_tol and freeze_stall must be >= 0");
    if (c->rk != 1 && c->rk != 3 && c->rk != 4) return fail(CFD_E_ARG, "config: rk must be 1, 3 or 4");
    if (c->rk == 1 && !c->ref_mode) return fail(CFD_E_ARG, "config: forward Euler (rk 1) exists in reference mode only");
    if (c->ref_mode && c->rk == 4) return fail(CFD_E_ARG, "config: the reference iteration has no RK4");
    if (c->block != 64 && c->block != 128 && c->block != 256) return fail(CFD_E_ARG, "config: block must be 64, 128 or 256");
    if (c->xcd_remap < 0 || c->xcd_remap > 4096) return fail(CFD_E_ARG, "config: xcd_remap must be in 0..4096");
    for (int k = 1; k <= 5; k++)
        if (c->bc[k] < 1 || c->bc[k] > 10) return fail(CFD_E_ARG, "config: bc types must be 1..10");
    if (!(c->p_exit >= 0.0) || !std::isfinite(c->p_exit)) return fail(CFD_E_ARG, "config: p_exit must be finite and >= 0");
    if (int rc = capi_check_viscous(c, "config")) return rc;
    if (c->positivity < 0 || c->positivity > 2) return fail(CFD_E_ARG, "config: positivity must be 0, 1 or 2");
    if (c->positivity && (c->order != 2 || c->ref_mode || c->viscous))
        return fail(CFD_E_ARG, "config: the positivity fallback needs order 2, the native iteration and no viscous terms");
    if (m->dev.tile > c->block) return fail(CFD_E_ARG, "config: the mesh tile (cells per tile) must not exceed the block size");
    const size_t sm = flux_smem_bytes(m->dev, c->order, c->grad_lds);
    if (sm > 65536 || grad_smem_bytes(m->dev, c->gradient) > 65536)
        return fail(CFD_E_ARG, "config: tile too large for 64 KiB of LDS per workgroup; use a smaller tile");
    return 0;
}

// WENO line stencils of the mesh's face entries on the device (first order-3 configuration).
static int ensure_line_stencils(cfd_mesh_s *m) {
    std::lock_guard<std::mutex> lock(m->lazy);
    if (m->d_fst) return 0;  // the last array uploaded
    const Layout &L = m->lay;
    std::vector<int> st;
    line_stencils(m->ref, L, st);
    std::vector<int> tile_of(L.perm.size());
    for (int t = 0; t < L.ntiles; t++)
        for (int k = L.tile_cstart[t]; k < L.tile_cstart[t + 1]; k++) tile_of[k] = t;
    std::vector<char> need(L.ntiles, 0);
    for (size_t e = 0; e < L.fL.size(); e++)
        if (st[4 * e] < 0) {
            need[tile_of[L.fL[e]]] = 1;
            if (L.fR[e] >= 0) need[tile_of[L.fR[e]]] = 1;
        }
    std::vector<int> gt;
    for (int t = 0; t < L.ntiles; t++)
        if (need[t]) gt.push_back(t);
    const int count = (int)gt.size();
    if (!m->d_gtiles) HIP_TRY(upload(&m->d_gtiles, gt.empty() ? std::vector<int>{0} : gt));
    if (!m->d_gcount) HIP_TRY(upload(&m->d_gcount, std::vector<int>{count}));
    m->ngtiles = count;
    HIP_TRY(upload(&m->d_fst, st));
    m->dev.fst = reinterpret_cast<const int4 *>(m->d_fst);
    return 0;
}

// Gradient pass of a time step: at order 3 only over the tiles whose gradients a MUSCL face
// reads (none on an all-quadrilateral line mesh away from its boundary).
static int step_grad(cfd_solver_s *s, const double4 *in, hipStream_t st) {
    const cfd_config_t &c = s->cfg;
    if (c.order == 2) {
        HIP_TRY(launch_grad(s->dev, s->dcfg, c.gradient, in, s->grad, c.block, st));
    } else if (c.order == 3 && c.viscous) {  // every face reads gradients
        HIP_TRY(launch_grad(s->dev, s->dcfg, c.gradient, in, s->grad, c.block, st));
    } else if (c.order == 3 && s->ngtiles > 0) {
        DevMesh dm = s->dev;
        dm.tile_list = s->m->d_gtiles;
        dm.tile_count = s->m->d_gcount;
        dm.ntiles = s->ngtiles;
        HIP_TRY(launch_grad(dm, s->dcfg, c.gradient, in, s->grad, c.block, st));
    }
    return 0;
}

int cfd_solver_configure(cfd_solver_t *s, const cfd_config_t *cfg) {
    if (!s || !cfg) return fail(CFD_E_ARG, "cfd_solver_configure: null argument");
    int rc = check_config(s->m, cfg);
    if (rc) return rc;
    if (cfg->order == 3 && (rc = ensure_line_stencils(s->m))) return rc;
    {
        std::lock_guard<std::mutex> lock(s->m->lazy);
        s->dev = s->m->dev;
        s->ngtiles = s->m->ngtiles;
    }
    // recorded limiter factors belong to the scheme that recorded them: a change of scheme or of
    // the freezing settings returns the march to the evaluated limiter
    const cfd_config_t &o = s->cfg;
    const bool scheme_changed = o.flux != cfg->flux || o.order != cfg->order || o.limiter != cfg->limiter ||
                                o.gradient != cfg->gradient || o.ref_mode != cfg->ref_mode ||
                                o.viscous != cfg->viscous || o.freeze_tol != cfg->freeze_tol ||
                                o.freeze_stall != cfg->freeze_stall;
    if (s->fz && scheme_changed) {
        s->fz_explicit = false;
        if ((rc = freeze_init(s, 0))) return rc;
    }
    s->cfg = *cfg;
    capi_fill_devcfg(s->dcfg, *cfg);
    s->dt_valid = false;
    return 0;
}

int cfd_solver_create(cfd_mesh_t *m, const cfd_config_t *cfg, cfd_solver_t **out) {
    if (!m || !cfg || !out) return fail(CFD_E_ARG, "cfd_solver_create: null argument");
    *out = nullptr;
    if (!m->on_device) return fail(CFD_E_MESH, "cfd_solver_create: mesh was built with cfd_mesh_create_host");
    cfd_solver_s *s = new cfd_solver_s();
    s->m = m;
    int rc = cfd_solver_configure(s, cfg);
    if (rc) {
        delete s;
        return rc;
    }
    const size_t nc = m->ref.nc;
    const size_t nparts = std::max<size_t>(m->lay.ntiles, (nc + 63) / 64);
    hipError_t e;
    if ((e = hipMalloc((void **)&s->Un, nc * sizeof(double4))) != hipSuccess ||
        (e = hipMalloc((void **)&s->Ua, nc * sizeof(double4))) != hipSuccess ||
        (e = hipMalloc((void **)&s->Ub, nc * sizeof(double4))) != hipSuccess ||
        (e = hipMalloc((void **)&s->acc, nc * sizeof(double4))) != hipSuccess ||
        (e = hipMalloc((void **)&s->R, nc * sizeof(double4))) != hipSuccess ||
        (e = hipMalloc((void **)&s->dt, nc * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&s->grad, 8 * nc * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&s->hbuf, nc * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&s->err_partial, 4 * nparts * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&s->dt_partial, nparts * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&s->scal, 8 * sizeof(double))) != hipSuccess) {
        delete s;
        return hip_fail(e, "cfd_solver_create hipMalloc");
    }
    *out = s;
    return 0;
}

void cfd_solver_destroy(cfd_solver_t *s) { delete s; }

int cfd_set_state(cfd_solver_t *s, const double *U) {
    if (!s || !U) return fail(CFD_E_ARG, "cfd_set_state: null argument");
    const int nc = s->m->ref.nc;
    std::vector<double> h(4 * (size_t)nc);
    for (int k = 0; k < nc; k++) std::memcpy(&h[4 * (size_t)k], &U[4 * (size_t)s->m->lay.perm[k]], 4 * sizeof(double));
    // steps launched on any stream (also a non-blocking one) may still read the state
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(s->Un, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    s->has_state = true;
    s->dt_valid = false;
    if (s->fz) return freeze_init(s, 0);  // a new state starts with the limiter evaluated
    return 0;
}

int cfd_get_state(cfd_solver_t *s, double *U) {
    if (!s || !U) return fail(CFD_E_ARG, "cfd_get_state: null argument");
    const int nc = s->m->ref.nc;
    std::vector<double> h(4 * (size_t)nc);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(h.data(), s->Un, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < nc; k++) std::memcpy(&U[4 * (size_t)s->m->lay.perm[k]], &h[4 * (size_t)k], 4 * sizeof(double));
    return 0;
}

int cfd_state_device(cfd_solver_t *s, void **d_U) {
    if (!s || !d_U) return fail(CFD_E_ARG, "cfd_state_device: null argument");
    *d_U = s->Un;
    return 0;
}

static StageArgs resid_args(cfd_solver_s *s, bool want_min, const double4 *uexit) {
    StageArgs st{};
    st.R_out = reinterpret_cast<double *>(s->R);
    st.dt_out = s->dt;
    st.dt_partial = want_min ? s->dt_partial : nullptr;
    st.Uexit = uexit;
    return st;
}

static int ensure_ref_cells(cfd_mesh_s *m) {
    std::lock_guard<std::mutex> lock(m->lazy);
    if (m->d_rc_area) return 0;  // the last array uploaded
    const RefMesh &R = m->ref;
    const Layout &L = m->lay;
    const int nc = R.nc;
    std::vector<int> nf(nc), nbr(4 * (size_t)nc, -1);
    std::vector<double> dist(4 * (size_t)nc), nrm(8 * (size_t)nc), len(4 * (size_t)nc), area(nc);
    for (int k = 0; k < nc; k++) {
        const int c = L.perm[k];
        nf[k] = R.nf[c];
        for (int f = 0; f < 4; f++) {
            const int nb = R.nbr[4 * c + f];
            nbr[4 * k + f] = (nb < 0) ? -1 : (nb < nc ? L.iperm[nb] : nb);
            dist[4 * k + f] = R.dist[4 * c + f];
            len[4 * k + f] = R.len[4 * c + f];
            nrm[8 * k + 2 * f] = R.nrm[8 * c + 2 * f];
            nrm[8 * k + 2 * f + 1] = R.nrm[8 * c + 2 * f + 1];
        }
        area[k] = R.area[c];
    }
    if (!m->d_rc_nf) HIP_TRY(upload(&m->d_rc_nf, nf));
    if (!m->d_rc_nbr) HIP_TRY(upload(&m->d_rc_nbr, nbr));
    if (!m->d_rc_dist) HIP_TRY(upload(&m->d_rc_dist, dist));
    if (!m->d_rc_nrm) HIP_TRY(upload(&m->d_rc_nrm, nrm));
    if (!m->d_rc_len) HIP_TRY(upload(&m->d_rc_len, len));
    if (!m->d_rc_area) HIP_TRY(upload(&m->d_rc_area, area));
    return 0;
}

// The reference's Evaluate_Cell_Net_Flux_2O (Roe on limited differences): physical state and
// ghosts (exit ghosts from uexit when set) gathered into Uall, then the cell-by-cell kernel.
static int ref_roe2o_pass(cfd_solver_s *s, const double4 *in, double *dtmin, const double4 *uexit, hipStream_t st) {
    cfd_mesh_s *m = s->m;
    const int nc = m->ref.nc, ng = m->ref.ng;
    int rc = ensure_ref_cells(m);
    if (rc) return rc;
    if (!s->Uall) HIP_TRY(hipMalloc((void **)&s->Uall, (size_t)(nc + ng) * sizeof(double4)));
    HIP_TRY(hipMemcpyAsync(s->Uall, in, (size_t)nc * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(launch_ghosts(ng, m->d_gcell, reinterpret_cast<const double2 *>(m->d_gnrm), m->d_gkind, s->dcfg, in,
                          uexit, s->Uall + nc, st));
    const RefCells cells{m->d_rc_nf, m->d_rc_nbr, m->d_rc_dist, m->d_rc_nrm, m->d_rc_len, m->d_rc_area};
    HIP_TRY(launch_roe2o(nc, cells, s->Uall, s->cfg.zeta, s->cfg.zeta1, s->cfg.cfl, s->R, s->dt,
                         dtmin ? s->dt_partial : nullptr, st));
    if (dtmin) HIP_TRY(launch_min_reduce(s->dt_partial, (nc + 255) / 256, dtmin, st));
    return 0;
}

// Order of the pass that forms the first global time step: the inviscid time step does not
// depend on the order, so a first-order pass suffices; the viscous term exists at order 2 only.
static int dt_pass_order(const cfd_config_t &c) { return c.viscous ? 2 : 1; }

// residual of state `in` into s->R, s->dt (+ minimum dt into *dtmin when set)
static int residual_pass(cfd_solver_s *s, const double4 *in, double *dtmin, const double4 *uexit,
                         hipStream_t st, int order_override = 0) {
    const cfd_config_t &c = s->cfg;
    const int order = order_override ? order_override : c.order;
    if (c.ref_mode && order == 2 && c.flux == FLUX_ROE) return ref_roe2o_pass(s, in, dtmin, uexit, st);
    if (order >= 2) HIP_TRY(launch_grad(s->dev, s->dcfg, c.gradient, in, s->grad, c.block, st));
    StageArgs a = resid_args(s, dtmin != nullptr, uexit);
    HIP_TRY(launch_flux(s->dev, s->dcfg, c.flux, order, c.limiter, MODE_RESID, in, s->grad, a, c.block, st));
    if (dtmin) HIP_TRY(launch_min_reduce(s->dt_partial, s->dev.ntiles, dtmin, st));
    return 0;
}

int cfd_residual(cfd_solver_t *s, double *R, double *dt, double *grad, void *stream) {
    if (!s) return fail(CFD_E_ARG, "cfd_residual: null solver");
    if (!s->has_state) return fail(CFD_E_STATE, "cfd_residual: state not set");
    hipStream_t st = (hipStream_t)stream;
    int rc = residual_pass(s, s->Un, nullptr, nullptr, st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    const int nc = s->m->ref.nc;
    const std::vector<int> &perm = s->m->lay.perm;
    if (R) {
        std::vector<double> h(4 * (size_t)nc);
        HIP_TRY(hipMemcpy(h.data(), s->R, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < nc; k++) std::memcpy(&R[4 * (size_t)perm[k]], &h[4 * (size_t)k], 4 * sizeof(double));
    }
    if (dt) {
        std::vector<double> h(nc);
        HIP_TRY(hipMemcpy(h.data(), s->dt, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < nc; k++) dt[perm[k]] = h[k];
    }
    if (grad && s->cfg.order >= 2) {
        std::vector<double> h(8 * (size_t)nc);
        HIP_TRY(hipMemcpy(h.data(), s->grad, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < nc; k++) std::memcpy(&grad[8 * (size_t)perm[k]], &h[8 * (size_t)k], 8 * sizeof(double));
    }
    return 0;
}

int cfd_apply_bc(cfd_solver_t *s, double *U_ghost, void *stream) {
    if (!s) return fail(CFD_E_ARG, "cfd_apply_bc: null solver");
    if (!s->has_state) return fail(CFD_E_STATE, "cfd_apply_bc: state not set");
    const int ng = s->m->ref.ng;
    if (ng == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (!s->Ug) HIP_TRY(hipMalloc((void **)&s->Ug, (size_t)ng * sizeof(double4)));
    const cfd_mesh_s *m = s->m;
    HIP_TRY(launch_ghosts(ng, m->d_gcell, reinterpret_cast<const double2 *>(m->d_gnrm), m->d_gkind,
                          s->dcfg, s->Un, nullptr, s->Ug, st));
    if (U_ghost) {
        HIP_TRY(hipMemcpyAsync(U_ghost, s->Ug, (size_t)ng * sizeof(double4), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return 0;
}

static bool auto_freeze(const cfd_solver_s *s) {
    return s->cfg.order == 2 && !s->cfg.ref_mode && (s->cfg.freeze_tol > 0.0 || s->cfg.freeze_stall > 0);
}

// Allocates the limiter-freezing buffers and sets the state to `mode` (best Error[0] reset).
static int freeze_init(cfd_solver_s *s, int mode) {
    if (!s->psi) {
        HIP_TRY(hipMalloc((void **)&s->psi, 8 * (size_t)s->dev.nfe * sizeof(float)));
        HIP_TRY(hipMalloc((void **)&s->fz, 2 * sizeof(int)));
        HIP_TRY(hipMalloc((void **)&s->fz_best, sizeof(double)));
    }
    const int h[2] = {mode, 0};
    const double b = 1.7976931348623157e308;
    // steps in flight on any stream may still read the freezing state
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(s->fz, h, sizeof(h), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->fz_best, &b, sizeof(b), hipMemcpyHostToDevice));
    return 0;
}

int cfd_set_freeze(cfd_solver_t *s, int mode) {
    if (!s || mode < 0 || mode > 2) return fail(CFD_E_ARG, "cfd_set_freeze: mode must be 0, 1 or 2");
    if (s->cfg.order != 2 || s->cfg.ref_mode) return fail(CFD_E_ARG, "cfd_set_freeze: limiter freezing needs order 2 in native mode");
    s->fz_explicit = (mode != 0);
    return freeze_init(s, mode);
}

// Buffers of the positivity fallback (first use); the tile marks start cleared and every
// recomputing pass clears the marks it consumed.
static int ensure_positivity_buffers(cfd_solver_s *s) {
    if (s->dirty) return 0;
    const size_t nc = s->m->ref.nc;
    HIP_TRY(hipMalloc((void **)&s->Uc, nc * sizeof(double4)));
    HIP_TRY(hipMalloc((void **)&s->acc2, nc * sizeof(double4)));
    HIP_TRY(hipMalloc((void **)&s->pflag, nc));
    // marks [ntiles], the list of marked tiles [ntiles] and its counts (two alternate here)
    const size_t nd = 2 * (size_t)s->dev.ntiles + 4;
    HIP_TRY(hipMalloc((void **)&s->dirty, nd * sizeof(int)));
    HIP_TRY(hipMemset(s->dirty, 0, nd * sizeof(int)));
    return 0;
}

// The fallback's check flags: range test (bit 0), the marked-tile count of this detecting launch
// (bits 1-2) and of the next one (bits 3-4), alternating between two.
static int pos_flags(const cfd_solver_s *s) {
    return (s->cfg.positivity == 1 ? 1 : 0) | (int)((s->pos_seq & 1u) << 1) | (int)(((s->pos_seq + 1u) & 1u) << 3);
}

// One native step (fused stages). want_err: emit the error norm into scal[3..6].
static int native_step(cfd_solver_s *s, bool want_err, hipStream_t st) {
    const cfd_config_t &c = s->cfg;
    if (auto_freeze(s)) {
        want_err = true;
        if (!s->fz) {
            int rc = freeze_init(s, 0);
            if (rc) return rc;
        }
    }
    const DevMesh &dm = s->dev;
    if (!c.lts && !s->dt_valid) {
        int rc = residual_pass(s, s->Un, s->scal + s->cur, nullptr, st, dt_pass_order(c));
        if (rc) return rc;
        s->dt_valid = true;
    }
    const int nst = (c.rk == 4) ? 4 : 3;
    const bool pos = c.positivity != 0;
    if (pos) {
        if (int rc = ensure_positivity_buffers(s)) return rc;
    }
    // With the positivity fallback no stage writes an array it reads (its recomputation reads the
    // stage inputs again): the last stage writes Uc, swapped with Un after the step, and the RK4
    // accumulator alternates between acc and acc2.
    double4 *last = pos ? s->Uc : s->Un;
    const double4 *ins3[3] = {s->Un, s->Ua, s->Ub};
    double4 *outs3[3] = {s->Ua, s->Ub, last};
    const double4 *ins4[4] = {s->Un, s->Ua, s->Ub, s->Ua};
    double4 *outs4[4] = {s->Ua, s->Ub, s->Ua, last};
    double4 *acc_w[4] = {s->acc, pos ? s->acc2 : s->acc, s->acc, s->acc};
    const double4 *acc_r[4] = {s->acc, s->acc, pos ? s->acc2 : s->acc, s->acc};
    for (int k = 0; k < nst; k++) {
        const double4 *in = (nst == 4) ? ins4[k] : ins3[k];
        double4 *out = (nst == 4) ? outs4[k] : outs3[k];
        if (int rc = step_grad(s, in, st)) return rc;
        StageArgs a{};
        a.rk = nst;
        a.stage = k;
        a.lts = c.lts;
        a.last = (k == nst - 1);
        a.Un = s->Un;
        a.Uout = out;
        a.acc_in = acc_r[k];
        a.acc_out = acc_w[k];
        a.hbuf = s->hbuf;
        a.dt_global = s->scal + s->cur;
        a.err_partial = (a.last && want_err) ? s->err_partial : nullptr;
        a.dt_partial = (a.last && !c.lts) ? s->dt_partial : nullptr;
        // recorded factors are used only while freezing is automatic or was set explicitly
        a.freeze = (auto_freeze(s) || s->fz_explicit) ? s->fz : nullptr;
        a.psi = s->psi;
        if (pos) {
            a.pflag = s->pflag;
            a.dirty = s->dirty;
            a.pos_range = pos_flags(s);
            HIP_TRY(launch_flux_pos(dm, s->dcfg, c.flux, c.limiter, 1, in, s->grad, a, c.block, st));
            HIP_TRY(launch_flux_pos(dm, s->dcfg, c.flux, c.limiter, 2, in, s->grad, a, c.block, st));
            s->pos_seq++;
        } else {
            HIP_TRY(launch_flux(dm, s->dcfg, c.flux, c.order, c.limiter, MODE_STAGE, in, s->grad, a, c.block, st));
        }
    }
    if (pos) std::swap(s->Un, s->Uc);
    if (!c.lts) HIP_TRY(launch_min_reduce(s->dt_partial, dm.ntiles, s->scal + (1 - s->cur), st));
    if (want_err) HIP_TRY(launch_err_finalize(s->err_partial, dm.ntiles, s->scal + 3, st));
    if (want_err && auto_freeze(s))
        HIP_TRY(launch_freeze_update(s->scal + 3, s->fz, s->fz_best, s->fz + 1, c.freeze_tol, c.freeze_stall, st));
    return 0;
}

// One iteration of the reference's Inviscid_Solver loop body.
static int ref_step(cfd_solver_s *s, hipStream_t st) {
    const cfd_config_t &c = s->cfg;
    const int nc = s->m->ref.nc;
    const int nb = (nc + 255) / 256;
    double *dmin = s->scal + 2;
    const double4 *R = s->R;
    int rc;
    if (c.rk == 1) {
        if ((rc = residual_pass(s, s->Un, dmin, s->Un, st))) return rc;
        HIP_TRY(launch_ref_update(nc, 3, 0, s->m->d_inv_area, s->Un, s->Un, R, s->dt, dmin, s->Un, s->err_partial, st));
    } else {
        if ((rc = residual_pass(s, s->Un, dmin, s->Un, st))) return rc;
        HIP_TRY(launch_ref_update(nc, 0, c.lts, s->m->d_inv_area, s->Un, s->Un, R, s->dt, dmin, s->Ua, nullptr, st));
        if ((rc = residual_pass(s, s->Ua, dmin, s->Un, st))) return rc;
        HIP_TRY(launch_ref_update(nc, 1, 0, s->m->d_inv_area, s->Un, s->Ua, R, s->dt, dmin, s->Ub, nullptr, st));
        if ((rc = residual_pass(s, s->Ub, dmin, s->Un, st))) return rc;
        HIP_TRY(launch_ref_update(nc, 2, 0, s->m->d_inv_area, s->Un, s->Ub, R, s->dt, dmin, s->Un, s->err_partial, st));
    }
    HIP_TRY(launch_err_finalize(s->err_partial, nb, s->scal + 3, st));
    return 0;
}

int cfd_step(cfd_solver_t *s, int nsteps, double *err_hist, double *dt_hist, void *stream) {
    if (!s) return fail(CFD_E_ARG, "cfd_step: null solver");
    if (!s->has_state) return fail(CFD_E_STATE, "cfd_step: state not set");
    if (nsteps < 0) return fail(CFD_E_ARG, "cfd_step: negative step count");
    hipStream_t st = (hipStream_t)stream;
    const bool sync = err_hist || dt_hist;
    for (int n = 0; n < nsteps; n++) {
        const int cur_before = s->cur;
        int rc = s->cfg.ref_mode ? ref_step(s, st) : native_step(s, err_hist != nullptr, st);
        if (rc) return rc;
        if (sync) {
            double h[8];
            HIP_TRY(hipMemcpyAsync(h, s->scal, sizeof(h), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (err_hist)
                for (int q = 0; q < 4; q++) err_hist[4 * n + q] = h[3 + q];
            if (dt_hist) dt_hist[n] = s->cfg.ref_mode ? h[2] : (s->cfg.lts ? NAN : h[cur_before]);
        }
        if (!s->cfg.ref_mode && !s->cfg.lts) s->cur = 1 - s->cur;
    }
    return 0;
}

int cfd_time_pass(cfd_solver_t *s, int which, int reps, float *ms, void *stream) {
    if (!s || reps <= 0 || !ms) return fail(CFD_E_ARG, "cfd_time_pass: bad arguments");
    if (!s->has_state) return fail(CFD_E_STATE, "cfd_time_pass: state not set");
    if (which < 0 || which > 3) return fail(CFD_E_ARG, "cfd_time_pass: which must be 0, 1, 2 or 3");
    if (which == 0 && s->cfg.order == 1) return fail(CFD_E_ARG, "cfd_time_pass: no gradient pass at order 1");
    if (s->cfg.ref_mode) return fail(CFD_E_ARG, "cfd_time_pass: native mode only");
    hipStream_t st = (hipStream_t)stream;
    const cfd_config_t &c = s->cfg;
    const DevMesh &dm = s->dev;
    if (!c.lts && !s->dt_valid) {
        int rc = residual_pass(s, s->Un, s->scal + s->cur, nullptr, st, dt_pass_order(c));
        if (rc) return rc;
        s->dt_valid = true;
    }
    if (c.order >= 2) HIP_TRY(launch_grad(dm, s->dcfg, c.gradient, s->Un, s->grad, c.block, st));
    StageArgs a{};
    a.rk = (c.rk == 4) ? 4 : 3;
    a.stage = 0;
    a.lts = c.lts;
    a.last = 0;
    a.Un = s->Un;
    a.Uout = s->Ua;
    a.acc_in = s->acc;
    a.acc_out = s->acc;
    a.hbuf = s->hbuf;
    a.dt_global = s->scal + s->cur;
    // with the positivity fallback the stage pass is its detecting pass (the recomputing pass
    // runs once after the timing, to clear the tile marks)
    const bool pos = c.positivity != 0 && c.order == 2;
    if (pos) {
        if (int rc = ensure_positivity_buffers(s)) return rc;
        a.pflag = s->pflag;
        a.dirty = s->dirty;
        a.pos_range = pos_flags(s);
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) {
        if (which == 0) HIP_TRY(launch_grad(dm, s->dcfg, c.gradient, s->Un, s->grad, c.block, st));
        else if (which == 1 && pos) HIP_TRY(launch_flux_pos(dm, s->dcfg, c.flux, c.limiter, 1, s->Un, s->grad, a, c.block, st));
        else if (which == 1) HIP_TRY(launch_flux(dm, s->dcfg, c.flux, c.order, c.limiter, MODE_STAGE, s->Un, s->grad, a, c.block, st));
        else if (which == 3) {
            const StageArgs r = resid_args(s, false, nullptr);
            HIP_TRY(launch_flux(dm, s->dcfg, c.flux, c.order, c.limiter, MODE_RESID, s->Un, s->grad, r, c.block, st));
        }
        else {
            int rc = native_step(s, false, st);
            if (rc) return rc;
            if (!c.lts) s->cur = 1 - s->cur;
        }
    }
    HIP_TRY(hipEventRecord(e1, st));
    if (which == 1 && pos) {
        HIP_TRY(launch_flux_pos(dm, s->dcfg, c.flux, c.limiter, 2, s->Un, s->grad, a, c.block, st));
        s->pos_seq++;
    }
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int cfd_run_steady(cfd_solver_t *s, double tol, int max_steps, int check_every, int *steps,
                   double *err4, void *stream) {
    if (!s || check_every <= 0 || max_steps < 0) return fail(CFD_E_ARG, "cfd_run_steady: bad arguments");
    if (!s->has_state) return fail(CFD_E_STATE, "cfd_run_steady: state not set");
    hipStream_t st = (hipStream_t)stream;
    int done = 0;
    double e[4] = {NAN, NAN, NAN, NAN};
    while (done < max_steps) {
        const int chunk = std::min(check_every, max_steps - done);
        for (int n = 0; n < chunk; n++) {
            const bool want = (n == chunk - 1);
            int rc = s->cfg.ref_mode ? ref_step(s, st) : native_step(s, want, st);
            if (rc) return rc;
            if (!s->cfg.ref_mode && !s->cfg.lts) s->cur = 1 - s->cur;
        }
        done += chunk;
        double h[8];
        HIP_TRY(hipMemcpyAsync(h, s->scal, sizeof(h), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int q = 0; q < 4; q++) e[q] = h[3 + q];
        if (!std::isfinite(e[0])) break;
        if (e[0] <= tol) break;
    }
    if (steps) *steps = done;
    if (err4)
        for (int q = 0; q < 4; q++) err4[q] = e[q];
    if (!(e[0] <= tol)) return fail(CFD_E_NOTCONV, "cfd_run_steady: not converged");
    return 0;
}

}  // extern "C"

namespace cfd {
int capi_check_config(const cfd_mesh_s *m, const cfd_config_t *c) { return check_config(m, c); }

void capi_fill_devcfg(DevCfg &d, const cfd_config_t &c) {
    for (int k = 0; k < 6; k++) d.bc[k] = c.bc[k];
    for (int k = 0; k < 4; k++) d.fs[k] = c.fs[k];
    d.cfl = c.cfl;
    d.roe_bdiss = c.roe_bdiss;
    d.xcd_remap = c.xcd_remap;
    d.grad_lds = c.grad_lds;
    d.p_exit = c.p_exit;
    capi_fill_viscous(d, c);
}
}  // namespace cfd
