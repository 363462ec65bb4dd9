/*
 * phx_rk4_flow_f64.hip -- instantiation unit: the FP64 many-steps-per-launch RK4
 * kernels (rk4_scalar_flow, phx_rk4.h).
 */
#include "phx_rk4.h"

namespace phx {
template hipError_t launch_rk4_flow<double>(const Rk4Args<double>&, const Rk4Tune&, const Rk4Flow&, hipStream_t);
} // namespace phx
