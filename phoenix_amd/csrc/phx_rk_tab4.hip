/*
 * phx_rk_tab4.hip -- instantiations of the fused Runge-Kutta kernel (phx_rk_generic.h) for the
 * three- and four-stage iterators (Fehlberg 2(1) fixed and adaptive, Ralston4, Rule38, Bogacki-Shampine 3(2) fixed and adaptive).
 */
#include "phx_rk_generic.h"

namespace phx {

template <typename real>
hipError_t launch_rk_tab4(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s) {
    switch (tab) {
        case RkTableau::FEHLBERG2: return launch_tab<real, TabFehlberg2>(a, s);
        case RkTableau::FEHLBERG2_ADAPTIVE: return launch_tab<real, TabFehlberg2Adaptive>(a, s);
        case RkTableau::RALSTON4: return launch_tab<real, TabRalston4>(a, s);
        case RkTableau::RULE38: return launch_tab<real, TabRule38>(a, s);
        case RkTableau::BOSHA: return launch_tab<real, TabBoSha>(a, s);
        case RkTableau::BOSHA_ADAPTIVE: return launch_tab<real, TabBoShaAdaptive>(a, s);
        default: break;
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_rk_tab4<float>(const RkGenArgs<float>&, RkTableau, hipStream_t);
template hipError_t launch_rk_tab4<double>(const RkGenArgs<double>&, RkTableau, hipStream_t);

} // namespace phx
