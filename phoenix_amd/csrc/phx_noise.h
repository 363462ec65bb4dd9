/*
 * phx_noise.h -- counter-based Gaussian noise for the stochastic GP term.
 *
 * The reference draws its per-step noise with cuRAND (curand_normal2,
 * kernel_random_numbers.cu:17-25) or std::normal_distribution on the CPU; both
 * are stateful.  This build uses Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) with
 *   key = 64-bit seed, counter = {cell, step lo, step hi, component},
 * followed by the Box-Muller transform, so each cell's draw is a pure function
 * of (seed, step, cell) and needs no state buffer.
 */
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

__host__ __device__ __forceinline__ void phx_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

/* library math (HIP device math on the GPU, libm on the host) */
__host__ __device__ __forceinline__ float phx_log(float x) { return logf(x); }
__host__ __device__ __forceinline__ double phx_log(double x) { return log(x); }
__host__ __device__ __forceinline__ float phx_sqrt(float x) { return sqrtf(x); }
__host__ __device__ __forceinline__ double phx_sqrt(double x) { return sqrt(x); }

/* principal square root of x + i y: u = sqrt((|z| + |x|) / 2), then (u, y / 2u) for x >= 0 and
   (|y| / 2u, u sign y) for x < 0 (no cancellation in either branch) */
template <typename real>
__device__ __forceinline__ void csqrt_principal(real x, real y, real& zr, real& zi) {
    const real r = hypot(x, y);
    if (r == real(0)) {
        zr = real(0);
        zi = real(0);
        return;
    }
    const real u = phx_sqrt((r + fabs(x)) * real(0.5));
    if (x >= real(0)) {
        zr = u;
        zi = y / (real(2) * u);
    } else {
        zr = fabs(y) / (real(2) * u);
        zi = copysign(u, y);
    }
}
__host__ __device__ __forceinline__ void phx_sincos(float x, float& s, float& c) {
    s = sinf(x);
    c = cosf(x);
}
__host__ __device__ __forceinline__ void phx_sincos(double x, double& s, double& c) {
    s = sin(x);
    c = cos(x);
}

/* complex standard normal (each part N(0,1)) for (seed, step, cell, component) */
template <typename real>
__host__ __device__ __forceinline__ void phx_noise_normal2(uint64_t seed, uint64_t step, uint32_t cell, uint32_t component, real& z0, real& z1) {
    uint32_t c[4] = {cell, (uint32_t)step, (uint32_t)(step >> 32), component};
    phx_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const real u1 = ((real)c[0] + real(0.5)) * real(1.0 / 4294967296.0);
    const real u2 = ((real)c[1] + real(0.5)) * real(1.0 / 4294967296.0);
    const real r = phx_sqrt(real(-2) * phx_log(u1));
    real s, co;
    phx_sincos(real(6.283185307179586476925286766559) * u2, s, co);
    z0 = r * co;
    z1 = r * s;
}
