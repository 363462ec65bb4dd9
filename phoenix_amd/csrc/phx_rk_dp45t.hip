/*
 * phx_rk_dp45t.hip -- the TE-TM instantiations of the adaptive Dormand-Prince 5(4) kernel in FP32
 * with 32-row tiles, in a translation unit of their own so that they compile next to the others.  The
 * tile height is chosen by launch_rk_generic (phx_rk_generic.hip).
 */
#include "phx_rk_generic.h"

namespace phx {

/* The reservoir variants come in two forms, with and without the potential: without it, the 8 scalar
   registers of its descriptors and the potential's row window are free, and the kernel spills fewer
   registers (the kernel's register count is over 256 VGPRs + AGPRs either way). */
template <int TY>
static hipError_t launch_tetm(const RkGenArgs<float>& a, hipStream_t s) {
    const bool res = a.n_in[0] != nullptr, st = a.stochastic;
    const bool pot = a.pot[0] != nullptr || a.pot[1] != nullptr;
    if (res && st)
        return pot ? launch_gen<float, TabDP45Adaptive, TY, true, true, true>(a, s)
                   : launch_gen<float, TabDP45Adaptive, TY, true, true, true, false>(a, s);
    if (res)
        return pot ? launch_gen<float, TabDP45Adaptive, TY, true, true, false>(a, s)
                   : launch_gen<float, TabDP45Adaptive, TY, true, true, false, false>(a, s);
    if (st) return launch_gen<float, TabDP45Adaptive, TY, true, false, true>(a, s);
    return launch_gen<float, TabDP45Adaptive, TY, true, false, false>(a, s);
}

hipError_t launch_dp45_tetm_tall(const RkGenArgs<float>& a, int ty, hipStream_t s) {
    if (!a.tetm) return hipErrorInvalidValue;
    if (ty == 32) return launch_tetm<32>(a, s);
    return hipErrorInvalidValue;
}

} // namespace phx
