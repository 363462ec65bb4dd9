/*
 * phx_rk_tab6.hip -- instantiations of the fused Runge-Kutta kernel (phx_rk_generic.h) for the
 * six-stage iterators Nystroem and Cash-Karp 5(4), fixed and adaptive.
 */
#include "phx_rk_generic.h"

namespace phx {

template <typename real>
hipError_t launch_rk_tab6(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s) {
    switch (tab) {
        case RkTableau::NYSTROEM: return launch_tab<real, TabNystroem>(a, s);
        case RkTableau::CASHKARP: return launch_tab<real, TabCashKarp>(a, s);
        case RkTableau::CASHKARP_ADAPTIVE: return launch_tab<real, TabCashKarpAdaptive>(a, s);
        default: break;
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_rk_tab6<float>(const RkGenArgs<float>&, RkTableau, hipStream_t);
template hipError_t launch_rk_tab6<double>(const RkGenArgs<double>&, RkTableau, hipStream_t);

} // namespace phx
