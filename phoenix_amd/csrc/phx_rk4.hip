/*
 * phx_rk4.hip -- FP32 instantiations of the fused RK4 kernels of phx_rk4.h (the two
 * precisions live in separate translation units so that they compile in parallel).
 */
#include "phx_rk4.h"

namespace phx {

template hipError_t launch_rk4_scalar<float>(const Rk4Args<float>&, const Rk4Tune&, hipStream_t);
template hipError_t launch_rk4_batched<float>(const Rk4BatchArgs<float>&, int, bool, bool, bool, hipStream_t);

} // namespace phx
