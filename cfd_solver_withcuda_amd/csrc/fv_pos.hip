// Flux pass of the native MUSCL stages with the positivity fallback (order 2, inviscid, every
// flux and limiter): pass 1 runs the stage and flags the cells whose new state fails the
// admissibility check, pass 2 recomputes the tiles holding a face of a flagged cell with
// first-order states on those faces. Compiled apart so that it builds in parallel with the other
// variants.
#include "fv_flux.h"

namespace cfd {

namespace {

// The marked tiles of a detecting launch, listed (in no particular order: each is recomputed on its
// own) after the marks; one thread per tile.
__global__ void __launch_bounds__(256) list_marked_tiles_kernel(int *dirty, int ntiles, int cur, int next) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int *count = dirty + 2 * (size_t)ntiles;
    if (t == 0) count[next] = 0;
    if (t < ntiles && dirty[t]) dirty[ntiles + atomicAdd(&count[cur], 1)] = t;
}

// Pass 1 runs one workgroup per tile; pass 2 runs persistent workgroups over the list of marked
// tiles (usually few or none), so that it does not launch one workgroup per tile.
template <int FLUX, int LIM, int PASS, bool FZ>
void launch_pos_fz(const DevMesh &m, const DevCfg &c, const double4 *Us, const double *grad,
                   const StageArgs &st, int block, size_t sm, hipStream_t s) {
    if constexpr (PASS == 1) {
        hipLaunchKernelGGL((flux_tile_kernel<FLUX, 2, LIM, MODE_STAGE, false, 1, FZ>), dim3(grid_of(m, c)),
                           dim3(block), sm, s, m, c, Us, grad, st);
    } else {
        hipLaunchKernelGGL(list_marked_tiles_kernel, dim3((m.ntiles + 255) / 256), dim3(256), 0, s, st.dirty,
                           m.ntiles, (st.pos_range >> 1) & 3, (st.pos_range >> 3) & 3);
        auto k = flux_tile_kernel<FLUX, 2, LIM, MODE_STAGE, false, 2, FZ, true>;
        hipLaunchKernelGGL(k, dim3(grid_persist(k, m, c, block, sm)), dim3(block), sm, s, m, c, Us, grad, st);
    }
}

template <int FLUX, int LIM, int PASS>
void launch_pos_t(const DevMesh &m, const DevCfg &c, const double4 *Us, const double *grad,
                  const StageArgs &st, int block, size_t sm, hipStream_t s) {
    // the freezing code only where a freezing state is set: it costs registers
    if (st.freeze)
        launch_pos_fz<FLUX, LIM, PASS, true>(m, c, Us, grad, st, block, sm, s);
    else
        launch_pos_fz<FLUX, LIM, PASS, false>(m, c, Us, grad, st, block, sm, s);
}

template <int PASS>
hipError_t launch_pos_pass(const DevMesh &m, const DevCfg &c, int flux, int lim, const double4 *Us,
                           const double *grad, const StageArgs &st, int block, size_t sm, hipStream_t s) {
    auto by_lim = [&](auto fluxc) {
        constexpr int F = decltype(fluxc)::value;
        switch (lim) {
        case LIM_NONE: launch_pos_t<F, LIM_NONE, PASS>(m, c, Us, grad, st, block, sm, s); break;
        case LIM_MINMOD: launch_pos_t<F, LIM_MINMOD, PASS>(m, c, Us, grad, st, block, sm, s); break;
        case LIM_VANLEER: launch_pos_t<F, LIM_VANLEER, PASS>(m, c, Us, grad, st, block, sm, s); break;
        case LIM_VANALBADA: launch_pos_t<F, LIM_VANALBADA, PASS>(m, c, Us, grad, st, block, sm, s); break;
        default: launch_pos_t<F, LIM_SUPERBEE, PASS>(m, c, Us, grad, st, block, sm, s); break;
        }
    };
    switch (flux) {
    case FLUX_ROE: by_lim(std::integral_constant<int, FLUX_ROE>{}); break;
    case FLUX_HLLC: by_lim(std::integral_constant<int, FLUX_HLLC>{}); break;
    case FLUX_LLF: by_lim(std::integral_constant<int, FLUX_LLF>{}); break;
    default: by_lim(std::integral_constant<int, FLUX_AUSM>{}); break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_flux_pos(const DevMesh &m, const DevCfg &c, int flux, int lim, int pass,
                           const double4 *Us, const double *grad, const StageArgs &st, int block,
                           hipStream_t s) {
    const bool geom = (st.lts && st.stage == 0) || (st.last && st.dt_partial);
    size_t sm = flux_launch_smem(m, 2, c.grad_lds, block, geom);
    if (pass == 1) {
        // the detection's per-face densities and pressures (4 floats a face) follow the face
        // geometry in LDS
        sm += sizeof(float) * 4 * (size_t)m.max_tile_faces;
        return launch_pos_pass<1>(m, c, flux, lim, Us, grad, st, block, sm, s);
    }
    return launch_pos_pass<2>(m, c, flux, lim, Us, grad, st, block, sm, s);
}

}  // namespace cfd
