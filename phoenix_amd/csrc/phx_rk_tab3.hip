/*
 * phx_rk_tab3.hip -- instantiations of the fused Runge-Kutta kernel (phx_rk_generic.h) for the
 * one- to three-stage fixed-step iterators (Newton, MP, Heun2, Ralston, Heun3, Ralston3, RK3, SSPRK3, HW).
 */
#include "phx_rk_generic.h"

namespace phx {

template <typename real>
hipError_t launch_rk_tab3(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s) {
    switch (tab) {
        case RkTableau::NEWTON: return launch_tab<real, TabNewton>(a, s);
        case RkTableau::MP: return launch_tab<real, TabMP>(a, s);
        case RkTableau::HEUN2: return launch_tab<real, TabHeun2>(a, s);
        case RkTableau::RALSTON: return launch_tab<real, TabRalston>(a, s);
        case RkTableau::HEUN3: return launch_tab<real, TabHeun3>(a, s);
        case RkTableau::RALSTON3: return launch_tab<real, TabRalston3>(a, s);
        case RkTableau::RK3: return launch_tab<real, TabRK3>(a, s);
        case RkTableau::SSPRK3: return launch_tab<real, TabSSPRK3>(a, s);
        case RkTableau::HW: return launch_tab<real, TabHW>(a, s);
        default: break;
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_rk_tab3<float>(const RkGenArgs<float>&, RkTableau, hipStream_t);
template hipError_t launch_rk_tab3<double>(const RkGenArgs<double>&, RkTableau, hipStream_t);

} // namespace phx
