/*
 * phx_rk_tab6f.hip -- instantiations of the fused Runge-Kutta kernel (phx_rk_generic.h) for the
 * six-stage Fehlberg 5(4) iterator, fixed and adaptive.
 */
#include "phx_rk_generic.h"

namespace phx {

template <typename real>
hipError_t launch_rk_tab6f(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s) {
    switch (tab) {
        case RkTableau::FEHLBERG5: return launch_tab<real, TabFehlberg5>(a, s);
        case RkTableau::FEHLBERG5_ADAPTIVE: return launch_tab<real, TabFehlberg5Adaptive>(a, s);
        default: break;
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_rk_tab6f<float>(const RkGenArgs<float>&, RkTableau, hipStream_t);
template hipError_t launch_rk_tab6f<double>(const RkGenArgs<double>&, RkTableau, hipStream_t);

} // namespace phx
