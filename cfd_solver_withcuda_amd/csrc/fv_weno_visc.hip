// Flux pass with WENO5 reconstruction and laminar viscous terms (order 3, every flux), compiled
// apart from the other variants so that the translation units build in parallel.
#include "fv_flux.h"

namespace cfd {

hipError_t launch_flux_weno_visc(const DevMesh &m, const DevCfg &c, int flux, int lim, int mode,
                                 const double4 *Us, const double *grad, const StageArgs &st, int block,
                                 size_t sm, hipStream_t s) {
    return launch_flux_family<3, true>(m, c, flux, lim, mode, Us, grad, st, block, sm, s);
}

}  // namespace cfd
