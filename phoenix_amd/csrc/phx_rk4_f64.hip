/*
 * phx_rk4_f64.hip -- FP64 instantiations of the fused RK4 kernels of phx_rk4.h (the two
 * precisions live in separate translation units so that they compile in parallel).
 */
#include "phx_rk4.h"

namespace phx {

template hipError_t launch_rk4_scalar<double>(const Rk4Args<double>&, const Rk4Tune&, hipStream_t);
template hipError_t launch_rk4_batched<double>(const Rk4BatchArgs<double>&, int, bool, bool, bool, hipStream_t);

} // namespace phx
