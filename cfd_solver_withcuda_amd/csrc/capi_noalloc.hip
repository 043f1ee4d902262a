// Allocation-free residual entry point (cfd_residual_noalloc, include/cfd_hip.h).
//
// It runs the kernels of cfd_residual (gradient pass, flux pass in residual mode, minimum of the
// local time steps) on a caller's device state. The gradients and the per-tile time-step minima
// those kernels need live in `scratch`, a file-scope device array of fixed size, so the call
// allocates nothing: no hipMalloc, no hipFree, no allocating library routine. Each device that
// loads the library holds its own copy of the array until the process ends; its address on a
// device is looked up once, with that device current, on the first call there.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_internal.h"
#include "fv_kernels.h"

using namespace cfd;

namespace {

// Capacity: gradients (8 doubles per cell) for CFD_NOALLOC_MAX_CELLS cells, then one time-step
// minimum per tile (a tile holds at least one cell, so at most CFD_NOALLOC_MAX_CELLS tiles).
constexpr size_t kMaxCells = CFD_NOALLOC_MAX_CELLS;
constexpr size_t kGradDoubles = 8 * kMaxCells;
constexpr size_t kScratchDoubles = kGradDoubles + kMaxCells;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

// This device's address of `scratch` (looked up on the first call with the device current).
int scratch_on_current_device(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return capi_hip_fail(e, "cfd_residual_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return capi_fail(CFD_E_ARG, "cfd_residual_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return capi_hip_fail(e, "cfd_residual_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) return capi_fail(CFD_E_ARG, "cfd_residual_noalloc: no device address for the scratch array");
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" int cfd_residual_noalloc(const cfd_mesh_t *m, const cfd_config_t *cfg, const void *d_U,
                                    void *d_R, void *d_dt, void *d_dtmin, void *stream) {
    if (!m || !cfg || !d_U || !d_R) return capi_fail(CFD_E_ARG, "cfd_residual_noalloc: null argument");
    if (!m->on_device) return capi_fail(CFD_E_MESH, "cfd_residual_noalloc: mesh was built with cfd_mesh_create_host");
    if ((size_t)m->dev.nc > kMaxCells || (size_t)m->dev.ntiles > kMaxCells)
        return capi_fail(CFD_E_ARG, "cfd_residual_noalloc: the mesh has " + std::to_string(m->dev.nc) +
                                        " cells; the scratch array holds at most " +
                                        std::to_string(kMaxCells) + " (CFD_NOALLOC_MAX_CELLS)");
    if (int rc = capi_check_config(m, cfg)) return rc;
    if (cfg->order > 2 || cfg->ref_mode)
        return capi_fail(CFD_E_ARG, "cfd_residual_noalloc: order 1 or 2 in the native iteration only");
    double *base = nullptr;
    if (int rc = scratch_on_current_device(&base)) return rc;
    double *grad = base;
    double *dt_partial = base + kGradDoubles;
    DevMesh dm;
    {
        std::lock_guard<std::mutex> lock(m->lazy);  // other solvers may add lazily built arrays
        dm = m->dev;
    }
    DevCfg dc{};
    capi_fill_devcfg(dc, *cfg);
    hipStream_t st = (hipStream_t)stream;
    const double4 *U = static_cast<const double4 *>(d_U);
    hipError_t e;
    if (cfg->order >= 2 && (e = launch_grad(dm, dc, cfg->gradient, U, grad, cfg->block, st)) != hipSuccess)
        return capi_hip_fail(e, "cfd_residual_noalloc: gradient pass");
    StageArgs a{};
    a.R_out = static_cast<double *>(d_R);
    a.dt_out = static_cast<double *>(d_dt);
    a.dt_partial = d_dtmin ? dt_partial : nullptr;
    if ((e = launch_flux(dm, dc, cfg->flux, cfg->order, cfg->limiter, MODE_RESID, U, grad, a, cfg->block, st)) != hipSuccess)
        return capi_hip_fail(e, "cfd_residual_noalloc: flux pass");
    if (d_dtmin && (e = launch_min_reduce(dt_partial, dm.ntiles, static_cast<double *>(d_dtmin), st)) != hipSuccess)
        return capi_hip_fail(e, "cfd_residual_noalloc: time-step minimum");
    return 0;
}
