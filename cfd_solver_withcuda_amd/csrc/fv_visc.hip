// Flux pass with the laminar viscous terms (second order, every flux and limiter), compiled
// apart from the inviscid variants so that the two translation units build in parallel.
#include "fv_flux.h"

namespace cfd {

hipError_t launch_flux_visc(const DevMesh &m, const DevCfg &c, int flux, int lim, int mode,
                            const double4 *Us, const double *grad, const StageArgs &st, int block,
                            size_t sm, hipStream_t s) {
    return launch_flux_family<2, true>(m, c, flux, lim, mode, Us, grad, st, block, sm, s);
}

}  // namespace cfd
