// Batched steady solves (include/cfd_hip.h, cfd_batch_*): many independent meshes of uneven
// size concatenated into one tiled device mesh, marched together with local time stepping.
//
// Every stage kernel runs once over the tiles of all instances still marching (a compacted
// tile list rebuilt on the device after each step), so large and small instances share the
// machine without per-instance launches. One whole step -- gradient and flux/stage passes of
// each RK stage, the per-instance error norm and convergence test, and the tile-list rebuild --
// is captured once in a hipGraph and replayed; the host only reads the instance status every
// `check_every` steps.
#include "../../include/cfd_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fv_kernels.h"
#include "mesh_host.h"

using namespace cfd;

namespace cfd {
int capi_fail(int code, const std::string &msg);
int capi_hip_fail(hipError_t e, const char *where);
int capi_check_viscous(const cfd_config_t *c, const char *who);
void capi_fill_viscous(DevCfg &d, const cfd_config_t &c);
}  // namespace cfd

namespace {

#define BHIP(expr)                                               \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return cfd::capi_hip_fail(_e, #expr); \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) return hipSuccess;
    return hipMalloc((void **)p, n * sizeof(T));
}

template <typename T>
hipError_t h2d(T *d, const T *h, size_t n) {
    if (n == 0) return hipSuccess;
    return hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice);
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct cfd_batch_s {
    int ninst = 0;
    cfd_config_t cfg{};
    DevCfg dcfg{};
    int tile = 0, face_budget = 0, block = 0;
    int nc = 0, ntiles = 0, nfe = 0, max_tile_faces = 0;
    std::vector<int> cell_off;   // [ninst+1]
    std::vector<int> inst_tile;  // [ninst+1]
    std::vector<int> iperm;      // caller cell (global) -> renumbered cell (global)
    // device mesh
    int *d_tile_cstart = nullptr, *d_tile_fstart = nullptr;
    double *d_fgeo = nullptr, *d_fw = nullptr, *d_area = nullptr, *d_inv_area = nullptr;
    double *d_fnx = nullptr, *d_fny = nullptr, *d_fdl = nullptr;
    int *d_fLR = nullptr;
    int *d_fst = nullptr;  // WENO line stencils, 4 per face entry (order 3)
    uint16_t *d_slots = nullptr;
    int *d_tile_inst = nullptr, *d_inst_tile = nullptr, *d_tile_list = nullptr, *d_tile_count = nullptr;
    int *d_inst_off = nullptr, *d_status = nullptr, *d_steps = nullptr;
    int *d_fz = nullptr, *d_fz_since = nullptr;  // per instance limiter-freezing state
    double *d_fz_best = nullptr;
    float *psi = nullptr;  // [8 * face entries] frozen limiter factors
    double4 *d_inst_fs = nullptr;
    double *d_err4 = nullptr;
    // solver state
    double4 *Un = nullptr, *Ua = nullptr, *Ub = nullptr, *acc = nullptr;
    // positivity fallback: the last stage's output (copied into Un over the marching tiles), the
    // second RK4 accumulator, per-cell flags, per-tile marks with the list of marked tiles
    double4 *Uc = nullptr, *acc2 = nullptr;
    uint8_t *pflag = nullptr;
    int *dirty = nullptr;
    double *grad = nullptr, *hbuf = nullptr, *err_partial = nullptr;
    DevMesh dev{};
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_tol = -1.0;
    int graph_tiles = 0;         // tiles the captured step's launch grids cover
    std::vector<int> fz_start;   // per instance freeze mode each solve starts with (empty: 0)
    std::vector<double> hist;    // Error[0] per instance at each check of the last solve
    int hist_checks = 0;
    cfd_progress_fn progress = nullptr;  // called every progress_every checks of a solve
    void *progress_user = nullptr;
    int progress_every = 0;
    double build_s = 0.0, tune_s = 0.0;
    int tune_src = 0;  // launch shape: 0 given, 1 timed by the autotuner, 2 from the autotune cache
    ~cfd_batch_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
        void *ps[] = {d_tile_cstart, d_tile_fstart, d_fgeo, d_fw, d_fnx, d_fny, d_fdl, d_fLR, d_fst, d_area, d_inv_area, d_slots, d_tile_inst, d_inst_tile,
                      d_tile_list, d_tile_count, d_inst_off, d_status, d_steps, d_inst_fs, d_err4,
                      Un, Ua, Ub, acc, grad, hbuf, err_partial, d_fz, d_fz_since, d_fz_best, psi,
                      Uc, acc2, pflag, dirty};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};

namespace {

struct InstanceInput {
    int npts, ncells;
    const double *pts;
    const int *cells;
};

int check_batch_config(const cfd_config_t *c) {
    if (c->flux < 0 || c->flux > 3) return capi_fail(CFD_E_ARG, "batch config: flux must be 0..3");
    if (c->order < 1 || c->order > 3) return capi_fail(CFD_E_ARG, "batch config: order must be 1, 2 or 3 (WENO5)");
    if (c->order == 3 && (c->freeze_tol > 0.0 || c->freeze_stall > 0))
        return capi_fail(CFD_E_ARG, "batch config: limiter freezing needs order 2");
    if (c->limiter < 0 || c->limiter > 4) return capi_fail(CFD_E_ARG, "batch config: limiter must be 0..4");
    if (c->gradient < 0 || c->gradient > 1) return capi_fail(CFD_E_ARG, "batch config: gradient must be 0 or 1");
    if (c->xcd_remap < 0 || c->xcd_remap > 4096) return capi_fail(CFD_E_ARG, "batch config: xcd_remap must be in 0..4096");
    if (!(c->cfl > 0.0)) return capi_fail(CFD_E_ARG, "batch config: cfl must be positive");
    if (c->rk != 3 && c->rk != 4) return capi_fail(CFD_E_ARG, "batch config: rk must be 3 or 4");
    if (c->ref_mode) return capi_fail(CFD_E_ARG, "batch config: the reference iteration is not batched");
    if (!(c->freeze_tol >= 0.0) || c->freeze_stall < 0)
        return capi_fail(CFD_E_ARG, "batch config: freeze_tol and freeze_stall must be >= 0");
    if (!c->lts) return capi_fail(CFD_E_ARG, "batch config: batched steady solves use local time stepping (lts = 1)");
    if (c->positivity < 0 || c->positivity > 2)
        return capi_fail(CFD_E_ARG, "batch config: positivity must be 0, 1 or 2");
    if (c->positivity && (c->order != 2 || c->viscous))
        return capi_fail(CFD_E_ARG, "batch config: the positivity fallback needs order 2, inviscid");
    for (int k = 1; k <= 5; k++)
        if (c->bc[k] < 1 || c->bc[k] > 10) return capi_fail(CFD_E_ARG, "batch config: bc types must be 1..10");
    if (!(c->p_exit >= 0.0) || !std::isfinite(c->p_exit))
        return capi_fail(CFD_E_ARG, "batch config: p_exit must be finite and >= 0");
    return capi_check_viscous(c, "batch config");
}

// Builds connectivity and tiled layouts of the instances (in parallel), concatenates them on
// the device and allocates the solver state.
int build_batch(const std::vector<InstanceInput> &in, const double *fs4, const cfd_config_t &cfg,
                int tile, int face_budget, int block, int nthreads, cfd_batch_s **out) {
    *out = nullptr;
    const int ni = (int)in.size();
    std::vector<Layout> lays(ni);
    std::vector<std::vector<int>> sts(ni);  // WENO line stencils (order 3)
    std::vector<std::string> errs(ni);
    {
        std::atomic<int> next{0};
        auto work = [&]() {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= ni) break;
                try {
                    RefMesh rm;
                    build_ref_mesh(in[i].npts, in[i].pts, in[i].ncells, in[i].cells, rm);
                    build_layout(rm, tile, 1, lays[i], face_budget);
                    if (cfg.order == 3) line_stencils(rm, lays[i], sts[i]);
                } catch (const std::exception &e) {
                    errs[i] = e.what();
                }
            }
        };
        const int nt = std::max(1, std::min(nthreads, ni));
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(work);
        work();
        for (auto &t : th) t.join();
    }
    for (int i = 0; i < ni; i++)
        if (!errs[i].empty()) return capi_fail(CFD_E_MESH, "cfd_batch_create: instance " + std::to_string(i) + ": " + errs[i]);

    cfd_batch_s *b = new cfd_batch_s();
    b->ninst = ni;
    b->cfg = cfg;
    b->tile = tile;
    b->face_budget = face_budget;
    b->block = block;
    b->cell_off.assign(ni + 1, 0);
    b->inst_tile.assign(ni + 1, 0);
    std::vector<int> fe_off(ni + 1, 0);
    for (int i = 0; i < ni; i++) {
        b->cell_off[i + 1] = b->cell_off[i] + (int)lays[i].area.size();
        b->inst_tile[i + 1] = b->inst_tile[i] + lays[i].ntiles;
        fe_off[i + 1] = fe_off[i] + (int)lays[i].fL.size();
        b->max_tile_faces = std::max(b->max_tile_faces, lays[i].max_tile_faces);
    }
    b->nc = b->cell_off[ni];
    b->ntiles = b->inst_tile[ni];
    b->nfe = fe_off[ni];
    const size_t nc = b->nc, nt = b->ntiles, nfe = b->nfe;
    hipError_t e = hipSuccess;
#define BALLOC(p, n)                                   \
    if ((e = dalloc(&(p), (n))) != hipSuccess) {       \
        delete b;                                      \
        return capi_hip_fail(e, "cfd_batch_create alloc"); \
    }
    BALLOC(b->d_tile_cstart, nt + 1);
    BALLOC(b->d_tile_fstart, nt + 1);
    BALLOC(b->d_fgeo, 6 * nfe);
    BALLOC(b->d_fw, 4 * nfe);
    BALLOC(b->d_fLR, 2 * nfe);
    BALLOC(b->d_fnx, nfe);
    BALLOC(b->d_fny, nfe);
    BALLOC(b->d_fdl, nfe);
    BALLOC(b->d_slots, 4 * nc);
    BALLOC(b->d_area, nc);
    BALLOC(b->d_inv_area, nc);
    BALLOC(b->d_tile_inst, nt);
    BALLOC(b->d_inst_tile, (size_t)ni + 1);
    BALLOC(b->d_tile_list, nt);
    BALLOC(b->d_tile_count, 1);
    BALLOC(b->d_inst_off, (size_t)ni);
    BALLOC(b->d_status, (size_t)ni);
    BALLOC(b->d_steps, (size_t)ni);
    BALLOC(b->d_inst_fs, (size_t)ni);
    BALLOC(b->d_err4, 4 * (size_t)ni);
    BALLOC(b->Un, nc);
    BALLOC(b->Ua, nc);
    BALLOC(b->Ub, nc);
    if (cfg.rk == 4) BALLOC(b->acc, nc);
    if (cfg.positivity) {
        BALLOC(b->Uc, nc);
        if (cfg.rk == 4) BALLOC(b->acc2, nc);
        BALLOC(b->pflag, nc);
        BALLOC(b->dirty, 2 * nt + 4);
        if ((e = hipMemset(b->dirty, 0, (2 * nt + 4) * sizeof(int))) != hipSuccess) {
            delete b;
            return capi_hip_fail(e, "cfd_batch_create alloc");
        }
    }
    if (cfg.order >= 2) BALLOC(b->grad, 8 * nc);
    if (cfg.order == 3) BALLOC(b->d_fst, 4 * nfe);
    BALLOC(b->hbuf, nc);
    BALLOC(b->err_partial, 4 * nt);
    if (cfg.order == 2 && (cfg.freeze_tol > 0.0 || cfg.freeze_stall > 0)) {
        BALLOC(b->d_fz, (size_t)ni);
        BALLOC(b->d_fz_since, (size_t)ni);
        BALLOC(b->d_fz_best, (size_t)ni);
        BALLOC(b->psi, 8 * nfe);
    }
#undef BALLOC
    b->iperm.resize(nc);
    for (int i = 0; i < ni; i++) {
        Layout &L = lays[i];
        const int cb = b->cell_off[i], tb = b->inst_tile[i], fb = fe_off[i];
        const size_t nci = L.area.size(), nti = L.ntiles, nfi = L.fL.size();
        std::vector<int> tc(nti), tf(nti), ti(nti, i);
        for (size_t t = 0; t < nti; t++) {
            tc[t] = L.tile_cstart[t] + cb;
            tf[t] = L.tile_fstart[t] + fb;
        }
        std::vector<double> geo, w;
        std::vector<int> lr;
        pack_face_records(L, cb, geo, w, lr);
        if (cfg.order == 3) {
            std::vector<int> &st = sts[i];
            for (int &v : st)
                if (v >= 0) v += cb;
            if ((e = h2d(b->d_fst + 4 * (size_t)fb, st.data(), 4 * nfi)) != hipSuccess) {
                delete b;
                return capi_hip_fail(e, "cfd_batch_create upload");
            }
            std::vector<int>().swap(st);
        }
        for (size_t c = 0; c < nci; c++) b->iperm[cb + c] = cb + L.iperm[c];
        if ((e = h2d(b->d_tile_cstart + tb, tc.data(), nti)) != hipSuccess ||
            (e = h2d(b->d_tile_fstart + tb, tf.data(), nti)) != hipSuccess ||
            (e = h2d(b->d_tile_inst + tb, ti.data(), nti)) != hipSuccess ||
            (e = h2d(b->d_fgeo + 6 * (size_t)fb, geo.data(), 6 * nfi)) != hipSuccess ||
            (e = h2d(b->d_fw + 4 * (size_t)fb, w.data(), 4 * nfi)) != hipSuccess ||
            (e = h2d(b->d_fLR + 2 * (size_t)fb, lr.data(), 2 * nfi)) != hipSuccess ||
            (e = h2d(b->d_fnx + fb, L.fnx.data(), nfi)) != hipSuccess ||
            (e = h2d(b->d_fny + fb, L.fny.data(), nfi)) != hipSuccess ||
            (e = h2d(b->d_fdl + fb, L.fdl.data(), nfi)) != hipSuccess ||
            (e = h2d(b->d_slots + 4 * (size_t)cb, L.slots.data(), 4 * nci)) != hipSuccess ||
            (e = h2d(b->d_area + cb, L.area.data(), nci)) != hipSuccess ||
            (e = h2d(b->d_inv_area + cb, L.inv_area.data(), nci)) != hipSuccess) {
            delete b;
            return capi_hip_fail(e, "cfd_batch_create upload");
        }
        L = Layout();  // release this instance's host copy
    }
    const int last_c = b->nc, last_f = b->nfe;
    std::vector<double4> fsv(ni);
    for (int i = 0; i < ni; i++) fsv[i] = make_double4(fs4[4 * i], fs4[4 * i + 1], fs4[4 * i + 2], fs4[4 * i + 3]);
    if ((e = h2d(b->d_tile_cstart + nt, &last_c, 1)) != hipSuccess ||
        (e = h2d(b->d_tile_fstart + nt, &last_f, 1)) != hipSuccess ||
        (e = h2d(b->d_inst_tile, b->inst_tile.data(), (size_t)ni + 1)) != hipSuccess ||
        (e = h2d(b->d_inst_fs, fsv.data(), (size_t)ni)) != hipSuccess ||
        (e = hipStreamCreate(&b->stream)) != hipSuccess) {
        delete b;
        return capi_hip_fail(e, "cfd_batch_create upload");
    }
    DevMesh &d = b->dev;
    d.nc = b->nc;
    d.ntiles = b->ntiles;
    d.tile = tile;
    d.max_tile_faces = b->max_tile_faces;
    d.nfe = b->nfe;
    d.tile_cstart = b->d_tile_cstart;
    d.tile_fstart = b->d_tile_fstart;
    d.fgeo = reinterpret_cast<const double2 *>(b->d_fgeo);
    d.fw = reinterpret_cast<const double2 *>(b->d_fw);
    d.fLR = reinterpret_cast<const int2 *>(b->d_fLR);
    d.fst = reinterpret_cast<const int4 *>(b->d_fst);
    d.fnx = b->d_fnx;
    d.fny = b->d_fny;
    d.fdl = b->d_fdl;
    d.slots = b->d_slots;
    d.area = b->d_area;
    d.inv_area = b->d_inv_area;
    d.tile_inst = b->d_tile_inst;
    d.inst_fs = b->d_inst_fs;
    d.tile_list = b->d_tile_list;
    d.tile_count = b->d_tile_count;
    for (int k = 0; k < 6; k++) b->dcfg.bc[k] = cfg.bc[k];
    for (int k = 0; k < 4; k++) b->dcfg.fs[k] = 0.0;
    b->dcfg.cfl = cfg.cfl;
    b->dcfg.roe_bdiss = cfg.roe_bdiss;
    b->dcfg.xcd_remap = cfg.xcd_remap;
    b->dcfg.grad_lds = cfg.grad_lds;
    b->dcfg.p_exit = cfg.p_exit;  // 0: each instance's prescribed p
    capi_fill_viscous(b->dcfg, cfg);
    const size_t sm = flux_smem_bytes(d, cfg.order, cfg.grad_lds);
    if (sm > 65536 || grad_smem_bytes(d, cfg.gradient) > 65536) {
        delete b;
        return capi_fail(CFD_E_ARG, "cfd_batch_create: tile too large for 64 KiB of LDS per workgroup");
    }
    *out = b;
    return 0;
}

// Marks every instance as marching and rebuilds the full tile list.
int batch_reset(cfd_batch_s *b) {
    std::vector<int> ones(b->ninst, 1), zeros(b->ninst, 0);
    BHIP(h2d(b->d_status, ones.data(), ones.size()));
    BHIP(h2d(b->d_steps, zeros.data(), zeros.size()));
    if (b->d_fz) {
        std::vector<double> big(b->ninst, 1.7976931348623157e308);
        BHIP(h2d(b->d_fz, b->fz_start.empty() ? zeros.data() : b->fz_start.data(), zeros.size()));
        BHIP(h2d(b->d_fz_since, zeros.data(), zeros.size()));
        BHIP(h2d(b->d_fz_best, big.data(), big.size()));
    }
    BHIP(launch_batch_list(b->ninst, b->ntiles, b->d_inst_tile, b->d_tile_inst, b->d_status,
                           b->d_inst_off, b->d_tile_list, b->d_tile_count, b->stream));
    return 0;
}

// The launches of one step on b->stream (captured into the graph). The tile passes' grids cover
// `grid_tiles` tiles (at least the tiles still marching: the compacted list never grows during a
// solve), so that a batch whose instances have mostly converged launches few empty workgroups.
int batch_step_launches(cfd_batch_s *b, double tol, int grid_tiles) {
    const cfd_config_t &c = b->cfg;
    const DevMesh &dm = b->dev;
    DevCfg dc = b->dcfg;
    dc.grid_tiles = grid_tiles;
    const int nst = (c.rk == 4) ? 4 : 3;
    // With the positivity fallback no stage writes an array it reads (its recomputation reads the
    // stage inputs again): the last stage writes Uc, copied into Un over the marching tiles after
    // the step, and the RK4 accumulator alternates between acc and acc2.
    const bool pos = c.positivity != 0;
    double4 *last = pos ? b->Uc : b->Un;
    const double4 *ins3[3] = {b->Un, b->Ua, b->Ub};
    double4 *outs3[3] = {b->Ua, b->Ub, last};
    const double4 *ins4[4] = {b->Un, b->Ua, b->Ub, b->Ua};
    double4 *outs4[4] = {b->Ua, b->Ub, b->Ua, last};
    double4 *acc_w[4] = {b->acc, pos ? b->acc2 : b->acc, b->acc, b->acc};
    const double4 *acc_r[4] = {b->acc, b->acc, pos ? b->acc2 : b->acc, b->acc};
    for (int k = 0; k < nst; k++) {
        const double4 *in = (nst == 4) ? ins4[k] : ins3[k];
        double4 *out = (nst == 4) ? outs4[k] : outs3[k];
        if (c.order >= 2) BHIP(launch_grad(dm, dc, c.gradient, in, b->grad, b->block, b->stream));
        StageArgs a{};
        a.rk = nst;
        a.stage = k;
        a.lts = 1;
        a.last = (k == nst - 1);
        a.Un = b->Un;
        a.Uout = out;
        a.acc_in = acc_r[k];
        a.acc_out = acc_w[k];
        a.hbuf = b->hbuf;
        a.err_partial = a.last ? b->err_partial : nullptr;
        a.freeze = b->d_fz;
        a.psi = b->psi;
        if (pos) {
            // the stages of one captured step use the marked-tile counts 0 .. nst-1 in turn
            a.pflag = b->pflag;
            a.dirty = b->dirty;
            a.pos_range = (c.positivity == 1 ? 1 : 0) | (k << 1) | (((k + 1) % nst) << 3);
            BHIP(launch_flux_pos(dm, dc, c.flux, c.limiter, 1, in, b->grad, a, b->block, b->stream));
            BHIP(launch_flux_pos(dm, dc, c.flux, c.limiter, 2, in, b->grad, a, b->block, b->stream));
        } else {
            BHIP(launch_flux(dm, dc, c.flux, c.order, c.limiter, MODE_STAGE, in, b->grad, a, b->block, b->stream));
        }
    }
    if (pos) BHIP(launch_batch_commit(dm, grid_tiles, b->Uc, b->Un, b->stream));
    BHIP(launch_batch_status(b->ninst, b->d_inst_tile, b->err_partial, tol, b->d_status, b->d_steps, b->d_err4,
                             b->d_fz, b->d_fz_best, b->d_fz_since, c.freeze_tol, c.freeze_stall, b->stream));
    BHIP(launch_batch_list(b->ninst, b->ntiles, b->d_inst_tile, b->d_tile_inst, b->d_status,
                           b->d_inst_off, b->d_tile_list, b->d_tile_count, b->stream));
    return 0;
}

int batch_graph(cfd_batch_s *b, double tol, int grid_tiles = -1) {
    if (grid_tiles < 0) grid_tiles = b->ntiles;
    if (b->exec && b->graph_tol == tol && b->graph_tiles == grid_tiles) return 0;
    if (b->exec) {
        BHIP(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        BHIP(hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_step_launches(b, tol, grid_tiles);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return capi_hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    BHIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_tol = tol;
    b->graph_tiles = grid_tiles;
    return 0;
}

// Time of `n` steps (after one warm-up step) on the current state, in milliseconds.
int batch_time_steps(cfd_batch_s *b, int n, double *ms) {
    int rc = batch_reset(b);
    if (rc) return rc;
    rc = batch_graph(b, -1.0);  // tol < 0: nothing converges while timing
    if (rc) return rc;
    BHIP(hipGraphLaunch(b->exec, b->stream));
    hipEvent_t e0, e1;
    BHIP(hipEventCreate(&e0));
    BHIP(hipEventCreate(&e1));
    BHIP(hipEventRecord(e0, b->stream));
    for (int k = 0; k < n; k++) BHIP(hipGraphLaunch(b->exec, b->stream));
    BHIP(hipEventRecord(e1, b->stream));
    BHIP(hipEventSynchronize(e1));
    float f = 0.f;
    BHIP(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    return 0;
}

// Steps the autotuner marches each candidate from freestream before timing it.
constexpr int kTuneDevelop = 200;

// `n` graph-launched steps of the current state (tolerance -1: nothing converges), then a wait.
int batch_march(cfd_batch_s *b, int n) {
    int rc = batch_reset(b);
    if (rc) return rc;
    rc = batch_graph(b, -1.0);
    if (rc) return rc;
    for (int k = 0; k < n; k++) BHIP(hipGraphLaunch(b->exec, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

// ---- autotune cache --------------------------------------------------------------------------
// One JSON object in a file, one entry per line: "key": [tile, face budget, block, ms per cell].
// The key names the GPU (architecture, compute units) and the batch's statistics (instances, total
// and largest cell counts in powers of two) and scheme, so a batch of the same shape on the same
// GPU reuses the measured choice. Path: CFD_AUTOTUNE_CACHE, else $HOME/.cache/cfdhip_autotune.json;
// an unreadable or unwritable file only disables the cache.
std::string tune_cache_path() {
    if (const char *p = std::getenv("CFD_AUTOTUNE_CACHE")) return p;
    if (const char *h = std::getenv("HOME")) return std::string(h) + "/.cache/cfdhip_autotune.json";
    return "";
}

int ilog2(long v) {
    int k = 0;
    while (v > 1) {
        v >>= 1;
        k++;
    }
    return k;
}

std::string tune_key(const cfd_config_t &c, const std::vector<InstanceInput> &in) {
    int dev = 0;
    hipDeviceProp_t p{};
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
        arch = p.gcnArchName;
        cus = p.multiProcessorCount;
    }
    long total = 0, largest = 0;
    for (const auto &x : in) {
        total += x.ncells;
        largest = std::max<long>(largest, x.ncells);
    }
    std::ostringstream k;
    k << arch << ":" << cus << "|inst2^" << ilog2((long)in.size()) << "|cells2^" << ilog2(total) << "|max2^"
      << ilog2(largest) << "|order" << c.order << "|flux" << c.flux << "|lim" << c.limiter << "|grad"
      << c.gradient << "|visc" << c.viscous << "|rk" << c.rk;
    std::string s = k.str();
    for (char &ch : s)
        if (ch == '"' || ch == '\\') ch = '_';
    return s;
}

// entries of the cache file: key -> "[tile, budget, block, per_cell]"
std::vector<std::pair<std::string, std::string>> tune_cache_load(const std::string &path) {
    std::vector<std::pair<std::string, std::string>> out;
    if (path.empty()) return out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const size_t a = line.find('"'), b = line.find("\": [");
        const size_t e = line.find(']', b == std::string::npos ? 0 : b);
        if (a == std::string::npos || b == std::string::npos || e == std::string::npos || b < a) continue;
        out.emplace_back(line.substr(a + 1, b - a - 1), line.substr(b + 3, e - b - 2));
    }
    return out;
}

bool tune_cache_get(const std::string &path, const std::string &key, int *tile, int *budget, int *block) {
    for (const auto &kv : tune_cache_load(path)) {
        if (kv.first != key) continue;
        int t = 0, bu = 0, bl = 0;
        if (std::sscanf(kv.second.c_str(), "[%d, %d, %d", &t, &bu, &bl) == 3 && t > 0 && bu > 0 &&
            (bl == 64 || bl == 128 || bl == 256) && t <= bl) {
            *tile = t;
            *budget = bu;
            *block = bl;
            return true;
        }
    }
    return false;
}

void tune_cache_put(const std::string &path, const std::string &key, int tile, int budget, int block,
                    double per_cell) {
    if (path.empty()) return;
    auto entries = tune_cache_load(path);
    char val[128];
    std::snprintf(val, sizeof(val), "[%d, %d, %d, %.6e]", tile, budget, block, per_cell);
    bool found = false;
    for (auto &kv : entries)
        if (kv.first == key) {
            kv.second = val;
            found = true;
        }
    if (!found) entries.emplace_back(key, val);
    const std::string tmp = path + ".tmp" + std::to_string((long)std::chrono::steady_clock::now().time_since_epoch().count());
    {
        std::ofstream f(tmp);
        if (!f) return;
        f << "{\n";
        for (size_t i = 0; i < entries.size(); i++)
            f << "\"" << entries[i].first << "\": " << entries[i].second << (i + 1 < entries.size() ? ",\n" : "\n");
        f << "}\n";
        if (!f) return;
    }
    (void)std::rename(tmp.c_str(), path.c_str());
}

int set_freestream(cfd_batch_s *b, const double *fs4) {
    std::vector<double4> U(b->nc);
    const double gm1 = 1.4 - 1.0;  // formed as the reference forms gamma - 1 (include/definitions.h:62-64)
    for (int i = 0; i < b->ninst; i++) {
        const double r = fs4[4 * i], u = fs4[4 * i + 1], v = fs4[4 * i + 2], p = fs4[4 * i + 3];
        const double4 q = make_double4(r, r * u, r * v, p / gm1 + 0.5 * r * (u * u + v * v));
        for (int c = b->cell_off[i]; c < b->cell_off[i + 1]; c++) U[c] = q;
    }
    BHIP(h2d(b->Un, U.data(), U.size()));
    return 0;
}

}  // namespace

extern "C" {

int cfd_batch_create(int ninst, const int *pt_off, const double *pts3, const int *cell_off,
                     const int *cells4, const double *fs4, const cfd_config_t *cfg, int nthreads,
                     cfd_batch_t **out) {
    if (!out || !pt_off || !pts3 || !cell_off || !cells4 || !fs4 || !cfg || ninst <= 0)
        return capi_fail(CFD_E_ARG, "cfd_batch_create: bad arguments");
    *out = nullptr;
    int rc = check_batch_config(cfg);
    if (rc) return rc;
    std::vector<InstanceInput> in(ninst);
    for (int i = 0; i < ninst; i++) {
        const int np = pt_off[i + 1] - pt_off[i], nc = cell_off[i + 1] - cell_off[i];
        if (np <= 0 || nc <= 0) return capi_fail(CFD_E_ARG, "cfd_batch_create: empty instance " + std::to_string(i));
        in[i] = InstanceInput{np, nc, pts3 + 3 * (size_t)pt_off[i], cells4 + 4 * (size_t)cell_off[i]};
    }
    const double t0 = now_s();
    int tile = 0, budget = 0, block = cfg->block;
    double tune_s = 0.0;
    int tune_src = 0;
    const std::string cache = (block == 0) ? tune_cache_path() : std::string();
    const std::string key = (block == 0) ? tune_key(*cfg, in) : std::string();
    if (block == 0 && tune_cache_get(cache, key, &tile, &budget, &block)) {
        tune_src = 2;
    } else if (block == 0) {
        // Autotune: time each (tile, face budget, block) candidate on a sample of the instances
        // (the largest ones, up to ~1.5M cells) and keep the fastest per cell.
        std::vector<int> order(ninst);
        for (int i = 0; i < ninst; i++) order[i] = i;
        std::sort(order.begin(), order.end(), [&](int a, int c) { return in[a].ncells > in[c].ncells; });
        std::vector<InstanceInput> sample;
        std::vector<double> sfs;
        long cells = 0;
        for (int i : order) {
            if (cells > 1500000 && !sample.empty()) break;
            sample.push_back(in[i]);
            for (int k = 0; k < 4; k++) sfs.push_back(fs4[4 * i + k]);
            cells += in[i].ncells;
        }
        const int cand[][3] = {{64, 128, 64}, {128, 256, 128}, {128, 320, 128}, {256, 512, 256}};
        double best = 1e300;
        for (const auto &cd : cand) {
            cfd_batch_s *sb = nullptr;
            cfd_config_t sc = *cfg;
            sc.block = cd[2];
            rc = build_batch(sample, sfs.data(), sc, cd[0], cd[1], cd[2], nthreads, &sb);
            if (rc) continue;  // e.g. a candidate that does not fit in LDS
            double ms = 0.0;
            // timed on a developing flow, not on the uniform start (where every limiter skips its
            // division and every face takes the same flux branch)
            rc = set_freestream(sb, sfs.data());
            if (!rc) rc = batch_march(sb, kTuneDevelop);
            if (!rc) rc = batch_time_steps(sb, 5, &ms);
            delete sb;
            if (rc) return rc;
            const double per_cell = ms / (double)cells;
            if (per_cell < best) {
                best = per_cell;
                tile = cd[0];
                budget = cd[1];
                block = cd[2];
            }
        }
        if (block == 0) return capi_fail(CFD_E_ARG, "cfd_batch_create: no autotune candidate fits");
        tune_s = now_s() - t0;
        tune_src = 1;
        tune_cache_put(cache, key, tile, budget, block, best);
    } else {
        if (block != 64 && block != 128 && block != 256)
            return capi_fail(CFD_E_ARG, "cfd_batch_create: block must be 0 (autotune), 64, 128 or 256");
        tile = block;
        budget = 2 * block;
    }
    cfd_config_t c = *cfg;
    c.block = block;
    cfd_batch_s *b = nullptr;
    rc = build_batch(in, fs4, c, tile, budget, block, nthreads, &b);
    if (rc) return rc;
    b->tune_s = tune_s;
    b->tune_src = tune_src;
    b->build_s = now_s() - t0 - tune_s;
    rc = set_freestream(b, fs4);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

void cfd_batch_destroy(cfd_batch_t *b) { delete b; }

int cfd_batch_info(const cfd_batch_t *b, int *out8, double *secs2) {
    if (!b) return capi_fail(CFD_E_ARG, "cfd_batch_info: null batch");
    if (out8) {
        const int v[8] = {b->ninst, b->nc, b->ntiles, b->nfe, b->tile, b->face_budget, b->block, b->max_tile_faces};
        std::memcpy(out8, v, sizeof(v));
    }
    if (secs2) {
        secs2[0] = b->build_s;
        secs2[1] = b->tune_s;
    }
    return 0;
}

int cfd_batch_tuning(const cfd_batch_t *b, int *source) {
    if (!b || !source) return capi_fail(CFD_E_ARG, "cfd_batch_tuning: null argument");
    *source = b->tune_src;
    return 0;
}

int cfd_batch_set_state(cfd_batch_t *b, const double *U) {
    if (!b || !U) return capi_fail(CFD_E_ARG, "cfd_batch_set_state: null argument");
    std::vector<double4> h(b->nc);
    for (int c = 0; c < b->nc; c++)
        h[b->iperm[c]] = make_double4(U[4 * (size_t)c], U[4 * (size_t)c + 1], U[4 * (size_t)c + 2], U[4 * (size_t)c + 3]);
    BHIP(h2d(b->Un, h.data(), h.size()));
    return 0;
}

int cfd_batch_get_state(cfd_batch_t *b, double *U) {
    if (!b || !U) return capi_fail(CFD_E_ARG, "cfd_batch_get_state: null argument");
    std::vector<double4> h(b->nc);
    BHIP(hipStreamSynchronize(b->stream));
    BHIP(hipMemcpy(h.data(), b->Un, h.size() * sizeof(double4), hipMemcpyDeviceToHost));
    for (int c = 0; c < b->nc; c++) {
        const double4 q = h[b->iperm[c]];
        U[4 * (size_t)c] = q.x;
        U[4 * (size_t)c + 1] = q.y;
        U[4 * (size_t)c + 2] = q.z;
        U[4 * (size_t)c + 3] = q.w;
    }
    return 0;
}

int cfd_batch_solve(cfd_batch_t *b, double tol, int max_steps, int check_every, int *steps,
                    double *err4, int *status, double *conv_s, double *wall_s, void *stream) {
    if (!b || max_steps < 0 || check_every <= 0) return capi_fail(CFD_E_ARG, "cfd_batch_solve: bad arguments");
    BHIP(hipStreamSynchronize((hipStream_t)stream));
    int rc = batch_reset(b);
    if (rc) return rc;
    rc = batch_graph(b, tol);
    if (rc) return rc;
    const int ni = b->ninst;
    std::vector<int> st(ni, 1);
    std::vector<double> conv(ni, -1.0);
    BHIP(hipStreamSynchronize(b->stream));
    const double t0 = now_s();
    int done = 0;
    bool all = false;
    std::vector<double> e4(4 * (size_t)ni);
    b->hist.clear();
    b->hist_checks = 0;
    while (done < max_steps && !all) {
        const int k = std::min(check_every, max_steps - done);
        for (int j = 0; j < k; j++) BHIP(hipGraphLaunch(b->exec, b->stream));
        done += k;
        BHIP(hipMemcpyAsync(st.data(), b->d_status, ni * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        BHIP(hipMemcpyAsync(e4.data(), b->d_err4, e4.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        BHIP(hipStreamSynchronize(b->stream));
        for (int i = 0; i < ni; i++) b->hist.push_back(e4[4 * (size_t)i]);
        b->hist_checks++;
        const double t = now_s() - t0;
        all = true;
        int marching = 0;
        for (int i = 0; i < ni; i++) {
            if (st[i] != 1 && conv[i] < 0.0) conv[i] = t;
            if (st[i] == 1) {
                all = false;
                marching++;
            }
        }
        if (b->progress && b->progress_every > 0 && b->hist_checks % b->progress_every == 0)
            b->progress(done, marching, b->progress_user);
        if (!all && !b->dcfg.persist) {
            // once the marching instances hold half the tiles the launch grids cover or fewer,
            // capture the step again with grids sized to them
            int active = 0;
            for (int i = 0; i < ni; i++)
                if (st[i] == 1) active += b->inst_tile[i + 1] - b->inst_tile[i];
            if (2 * active <= b->graph_tiles && (rc = batch_graph(b, tol, active))) return rc;
        }
    }
    const double wall = now_s() - t0;
    if (steps) BHIP(hipMemcpy(steps, b->d_steps, ni * sizeof(int), hipMemcpyDeviceToHost));
    if (err4) BHIP(hipMemcpy(err4, b->d_err4, 4 * (size_t)ni * sizeof(double), hipMemcpyDeviceToHost));
    if (status) std::memcpy(status, st.data(), ni * sizeof(int));
    if (conv_s) std::memcpy(conv_s, conv.data(), ni * sizeof(double));
    if (wall_s) *wall_s = wall;
    for (int i = 0; i < ni; i++)
        if (st[i] != 0) return CFD_E_NOTCONV;
    return 0;
}

int cfd_batch_set_freeze(cfd_batch_t *b, const int *mode) {
    if (!b) return capi_fail(CFD_E_ARG, "cfd_batch_set_freeze: null batch");
    if (!b->d_fz) return capi_fail(CFD_E_ARG, "cfd_batch_set_freeze: limiter freezing is off (freeze_tol, freeze_stall)");
    if (!mode) {
        b->fz_start.clear();
        return 0;
    }
    for (int i = 0; i < b->ninst; i++)
        if (mode[i] != 0 && mode[i] != 1) return capi_fail(CFD_E_ARG, "cfd_batch_set_freeze: mode must be 0 or 1");
    b->fz_start.assign(mode, mode + b->ninst);
    return 0;
}

int cfd_batch_set_progress(cfd_batch_t *b, cfd_progress_fn fn, void *user, int every_checks) {
    if (!b || every_checks < 0) return capi_fail(CFD_E_ARG, "cfd_batch_set_progress: bad arguments");
    b->progress = fn;
    b->progress_user = user;
    b->progress_every = every_checks;
    return 0;
}

int cfd_batch_set_cfl(cfd_batch_t *b, double cfl) {
    if (!b || !(cfl > 0.0) || !std::isfinite(cfl)) return capi_fail(CFD_E_ARG, "cfd_batch_set_cfl: cfl must be positive");
    b->cfg.cfl = cfl;
    b->dcfg.cfl = cfl;
    b->graph_tol = std::nan("");  // the captured step holds the old value: capture again
    return 0;
}

int cfd_batch_set_launch(cfd_batch_t *b, int mode) {
    if (!b || (mode != 0 && mode != 1)) return capi_fail(CFD_E_ARG, "cfd_batch_set_launch: mode must be 0 or 1");
    if (mode == 1 && (b->cfg.order > 2 || b->cfg.viscous || b->cfg.positivity))
        return capi_fail(CFD_E_ARG, "cfd_batch_set_launch: persistent workgroups need order 1 or 2, inviscid, "
                                    "without the positivity fallback");
    if (b->dcfg.persist != mode) {
        b->dcfg.persist = mode;
        b->graph_tol = std::nan("");  // the captured step holds the old launches: capture again
    }
    return 0;
}

int cfd_batch_history(const cfd_batch_t *b, int *nchecks, double *e0) {
    if (!b || !nchecks) return capi_fail(CFD_E_ARG, "cfd_batch_history: null argument");
    *nchecks = b->hist_checks;
    if (e0) std::memcpy(e0, b->hist.data(), b->hist.size() * sizeof(double));
    return 0;
}

int cfd_batch_time_steps(cfd_batch_t *b, int nsteps, float *ms) {
    if (!b || !ms || nsteps <= 0) return capi_fail(CFD_E_ARG, "cfd_batch_time_steps: bad arguments");
    double m = 0.0;
    int rc = batch_time_steps(b, nsteps, &m);
    *ms = (float)m;
    return rc;
}

}  // extern "C"
