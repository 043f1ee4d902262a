// Internal declarations of the reference-layout kernels (ref_kernels.hip).
#pragma once
#include "fv_device.h"

namespace cfd {

enum { REF_LIM_MINMOD = 0, REF_LIM_VANLEER = 1, REF_LIM_SUPERBEE = 2, REF_LIM_VANALBADA = 3 };
enum {
    REF_BC_SUBSONIC_INLET = 0,
    REF_BC_SUPERSONIC_INLET = 1,
    REF_BC_SUBSONIC_EXIT = 2,
    REF_BC_SUPERSONIC_EXIT = 3,
    REF_BC_VISCOUS_WALL = 4,
    REF_BC_INVISCID_WALL = 5,
    REF_BC_SYMMETRY = 6,
    REF_BC_FARFIELD = 7
};

// boundary data passed by value to the BC kernel
struct RefBcData {
    double rho, u, v, p;
};

}  // namespace cfd
