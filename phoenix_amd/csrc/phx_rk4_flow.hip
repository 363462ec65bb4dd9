/*
 * phx_rk4_flow.hip -- instantiation unit: the FP32 many-steps-per-launch RK4
 * kernels (rk4_scalar_flow, phx_rk4.h), compiled apart from the one-step kernels
 * so that the units build side by side.
 */
#include "phx_rk4.h"

namespace phx {
template hipError_t launch_rk4_flow<float>(const Rk4Args<float>&, const Rk4Tune&, const Rk4Flow&, hipStream_t);
} // namespace phx
