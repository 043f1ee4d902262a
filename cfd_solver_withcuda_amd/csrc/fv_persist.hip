// Stage passes of batched solves with persistent workgroups (order 1 or 2, inviscid, every flux
// and limiter, with or without limiter freezing): the grid holds what the GPU runs at once and
// each workgroup walks the tiles still marching. Compiled apart so that it builds in parallel with
// the one-tile-per-workgroup variants.
#include "fv_flux.h"

namespace cfd {

namespace {

template <int FLUX, int ORDER, int LIM, bool FZ>
void launch_persist_t(const DevMesh &m, const DevCfg &c, const double4 *Us, const double *grad,
                      const StageArgs &st, int block, size_t sm, hipStream_t s) {
    auto k = flux_tile_kernel<FLUX, ORDER, LIM, MODE_STAGE, false, 0, FZ, true>;
    hipLaunchKernelGGL(k, dim3(grid_persist(k, m, c, block, sm)), dim3(block), sm, s, m, c, Us, grad, st);
}

template <int FLUX>
void launch_persist_lim(const DevMesh &m, const DevCfg &c, int order, int lim, const double4 *Us,
                        const double *grad, const StageArgs &st, int block, size_t sm, hipStream_t s) {
    if (order == 1) {
        launch_persist_t<FLUX, 1, LIM_NONE, false>(m, c, Us, grad, st, block, sm, s);
        return;
    }
    auto by_fz = [&](auto limc) {
        constexpr int L = decltype(limc)::value;
        if (st.freeze)
            launch_persist_t<FLUX, 2, L, true>(m, c, Us, grad, st, block, sm, s);
        else
            launch_persist_t<FLUX, 2, L, false>(m, c, Us, grad, st, block, sm, s);
    };
    switch (lim) {
    case LIM_NONE: by_fz(std::integral_constant<int, LIM_NONE>{}); break;
    case LIM_MINMOD: by_fz(std::integral_constant<int, LIM_MINMOD>{}); break;
    case LIM_VANLEER: by_fz(std::integral_constant<int, LIM_VANLEER>{}); break;
    case LIM_VANALBADA: by_fz(std::integral_constant<int, LIM_VANALBADA>{}); break;
    default: by_fz(std::integral_constant<int, LIM_SUPERBEE>{}); break;
    }
}

}  // namespace

hipError_t launch_flux_persist(const DevMesh &m, const DevCfg &c, int flux, int order, int lim,
                               const double4 *Us, const double *grad, const StageArgs &st, int block,
                               size_t sm, hipStream_t s) {
    switch (flux) {
    case FLUX_ROE: launch_persist_lim<FLUX_ROE>(m, c, order, lim, Us, grad, st, block, sm, s); break;
    case FLUX_HLLC: launch_persist_lim<FLUX_HLLC>(m, c, order, lim, Us, grad, st, block, sm, s); break;
    case FLUX_LLF: launch_persist_lim<FLUX_LLF>(m, c, order, lim, Us, grad, st, block, sm, s); break;
    default: launch_persist_lim<FLUX_AUSM>(m, c, order, lim, Us, grad, st, block, sm, s); break;
    }
    return hipGetLastError();
}

}  // namespace cfd
