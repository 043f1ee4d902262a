// Order-2 stages of the flux pass with limiter freezing (every flux and limiter, inviscid and
// viscous): the variants that read the freezing state and record or use the frozen limiter
// factors. The other stage kernels hold none of that code; this translation unit builds in
// parallel with them.
#include "fv_flux.h"

namespace cfd {

namespace {

template <int FLUX, int LIM, bool VISC>
void launch_fz_t(const DevMesh &m, const DevCfg &c, const double4 *Us, const double *grad,
                 const StageArgs &st, int block, size_t sm, hipStream_t s) {
    hipLaunchKernelGGL((flux_tile_kernel<FLUX, 2, LIM, MODE_STAGE, VISC, 0, true>), dim3(grid_of(m, c)),
                       dim3(block), sm, s, m, c, Us, grad, st);
}

template <bool VISC>
void launch_fz(const DevMesh &m, const DevCfg &c, int flux, int lim, const double4 *Us, const double *grad,
               const StageArgs &st, int block, size_t sm, hipStream_t s) {
    auto by_lim = [&](auto fluxc) {
        constexpr int F = decltype(fluxc)::value;
        switch (lim) {
        case LIM_NONE: launch_fz_t<F, LIM_NONE, VISC>(m, c, Us, grad, st, block, sm, s); break;
        case LIM_MINMOD: launch_fz_t<F, LIM_MINMOD, VISC>(m, c, Us, grad, st, block, sm, s); break;
        case LIM_VANLEER: launch_fz_t<F, LIM_VANLEER, VISC>(m, c, Us, grad, st, block, sm, s); break;
        case LIM_VANALBADA: launch_fz_t<F, LIM_VANALBADA, VISC>(m, c, Us, grad, st, block, sm, s); break;
        default: launch_fz_t<F, LIM_SUPERBEE, VISC>(m, c, Us, grad, st, block, sm, s); break;
        }
    };
    switch (flux) {
    case FLUX_ROE: by_lim(std::integral_constant<int, FLUX_ROE>{}); break;
    case FLUX_HLLC: by_lim(std::integral_constant<int, FLUX_HLLC>{}); break;
    case FLUX_LLF: by_lim(std::integral_constant<int, FLUX_LLF>{}); break;
    default: by_lim(std::integral_constant<int, FLUX_AUSM>{}); break;
    }
}

}  // namespace

hipError_t launch_flux_frozen(const DevMesh &m, const DevCfg &c, int flux, int lim, bool visc,
                              const double4 *Us, const double *grad, const StageArgs &st, int block,
                              size_t sm, hipStream_t s) {
    if (visc) launch_fz<true>(m, c, flux, lim, Us, grad, st, block, sm, s);
    else launch_fz<false>(m, c, flux, lim, Us, grad, st, block, sm, s);
    return hipGetLastError();
}

}  // namespace cfd
