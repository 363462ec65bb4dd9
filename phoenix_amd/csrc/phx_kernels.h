/*
 * phx_kernels.h -- host-side launch interface of the gfx950 kernels.
 * Device layout: full grid, row-major, psi interleaved complex (float2/double2),
 * n, P, V real.  No padding: boundaries are resolved inside the kernels.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"

namespace phx {

template <typename real>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

template <typename real>
struct Rk4Args {
    const typename Vec2<real>::type* psi_in;
    const real* n_in; /* nullptr: no reservoir */
    typename Vec2<real>::type* psi_out;
    real* n_out;
    const real* pump; /* nullptr: no pump */
    const real* pot;  /* nullptr: no potential */
    int Nc, Nr;
    int periodic_x, periodic_y;
    int nstrips, nbands; /* set by the launcher */
    real dt;
    real pamp[4], vamp[4]; /* temporal amplitude per stage: a(t)(1-w)+w a(t+dt) */
    PhxPhys<real> ph;
};

/* ty: output rows per wavefront tile (8, 16 or 32). */
template <typename real>
hipError_t launch_rk4_scalar(const Rk4Args<real>& a, int ty, hipStream_t s);

} // namespace phx
