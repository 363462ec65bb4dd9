/*
 * phx_ssfm.hip -- pointwise parts of the split-step Fourier iterator (SSFM::step,
 * split_step_fourier_method.cu:25-71); the linear half steps run through the LDS FFT
 * (launch_fft_propagate, phx_fft.hip).
 *
 *   ssfm_nonlinear    : gp_scalar_nonlinear (kernel_gp_compute.cuh:264-382): per cell
 *                       psi <- psi exp(dt (R n - gamma_c) / 2 + dt B / hbar) exp(-i dt A / hbar),
 *                       A = g_c |psi|^2 + sum_g V_g Re a_g + g_r n (+ g_pm |psi_other|^2 with TE-TM)
 *                       (- g_c / dV with noise), B = sum_g V_g Im a_g, and the explicit Euler
 *                       reservoir step n' = n + dt (-(gamma_r + R |psi|^2) n + sum_g P_g Re a_g
 *                       (+ R n / dV with noise)); the TE-TM splitting delta_LT is not part of it
 *                       (kernel_gp_compute.cuh:320,357).
 *   ssfm_independent  : gp_scalar_independent (kernel_gp_compute.cuh:387-440): psi += the pulse
 *                       sum_g F_g a_g times dt / hbar in the scalar model and times -i dt / hbar
 *                       in the TE-TM model (the reference writes the two differently, :392-396
 *                       and :409-413), plus the noise xi sqrt((R n + gamma_c) / (4 dV)); the scalar
 *                       model takes the reservoir of the step start and then installs the new one;
 *                       the TE-TM model takes the new reservoir of each component and keeps the old
 *                       one (the reference never copies it back, so its TE-TM reservoirs do not
 *                       evolve under SSFM).
 * A complex reservoir (n as {re, im} cells, under a complex temporal pump amplitude) carries Im n
 * through the same formulas: the exponent gains dt g_r Im n / hbar in its real part and
 * dt R Im n / 2 in its phase, the reservoir step takes sum_g P_g a_g whole, and the noise
 * amplitude is the principal complex square root of (R n + gamma_c) / (4 dV).
 * One thread per cell, coalesced; the fields are the solver's full-grid arrays.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_kernels.h"
#include "phx_noise.h"

namespace phx {

namespace {
__device__ __forceinline__ void phx_sincos_dev(float x, float& s, float& c) { sincosf(x, &s, &c); }
__device__ __forceinline__ void phx_sincos_dev(double x, double& s, double& c) { sincos(x, &s, &c); }
__device__ __forceinline__ float phx_exp_dev(float x) { return expf(x); }
__device__ __forceinline__ double phx_exp_dev(double x) { return exp(x); }
} // namespace

template <typename real, bool TETM, bool CRES>
__global__ __launch_bounds__(256) void ssfm_nonlinear(SsfmArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= a.cells) return;
    constexpr int NC = TETM ? 2 : 1;
    const PhxPhys<real>& ph = a.ph;
    v2 psi[NC];
    real nrm[NC], n[NC], ni[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        psi[k] = a.psi[k][i];
        nrm[k] = phx_fma(psi[k].x, psi[k].x, psi[k].y * psi[k].y);
        if constexpr (CRES) {
            const v2 z = reinterpret_cast<const v2*>(a.n_cur[k])[i];
            n[k] = z.x;
            ni[k] = z.y;
        } else {
            n[k] = a.n_cur[k] ? a.n_cur[k][i] : real(0);
            ni[k] = real(0);
        }
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
        real A = ph.g_c * nrm[k], B = 0;
        for (int g = 0; g < a.npot; g++) {
            const real v = a.pot[k][g][i];
            A += v * a.vamp[g].x;
            B += v * a.vamp[g].y;
        }
        A += ph.g_r * n[k];
        if (TETM) A += ph.g_pm * nrm[1 - k];
        if (a.stochastic) A -= ph.g_c / ph.dV;
        /* Im n in g_r n + i hbar/2 R n moves g_r Im n into the gain and hbar/2 R Im n out of the phase */
        if constexpr (CRES) B += ph.g_r * ni[k];
        const real er = phx_fma(a.dt * ph.one_over_h_bar_s, B, a.dt * (real(0.5) * (ph.R * n[k] - ph.gamma_c)));
        real ei = -(A * ph.one_over_h_bar_s * a.dt);
        if constexpr (CRES) ei = phx_fma(a.dt * real(0.5) * ph.R, ni[k], ei);
        real sn, cs;
        phx_sincos_dev(ei, sn, cs);
        const real m = phx_exp_dev(er);
        const real fr = m * cs, fi = m * sn;
        v2 o;
        o.x = phx_fma(psi[k].x, fr, -(psi[k].y * fi));
        o.y = phx_fma(psi[k].x, fi, psi[k].y * fr);
        a.psi[k][i] = o;
        if (a.n_nxt[k]) {
            real rn = -(ph.gamma_r * n[k]);
            rn -= ph.R * nrm[k] * n[k];
            for (int g = 0; g < a.npump; g++) rn += a.pump[k][g][i] * a.pamp[g];
            if (a.stochastic) rn += ph.R * n[k] / ph.dV;
            if constexpr (CRES) {
                real rni = -(ph.gamma_r * ni[k]);
                rni -= ph.R * nrm[k] * ni[k];
                for (int g = 0; g < a.npump; g++) rni += a.pump[k][g][i] * a.pamp_im[g];
                if (a.stochastic) rni += ph.R * ni[k] / ph.dV;
                v2 z;
                z.x = phx_fma(rn, a.dt, n[k]);
                z.y = phx_fma(rni, a.dt, ni[k]);
                reinterpret_cast<v2*>(a.n_nxt[k])[i] = z;
            } else {
                a.n_nxt[k][i] = phx_fma(rn, a.dt, n[k]);
            }
        }
    }
}

template <typename real, bool TETM, bool CRES>
__global__ __launch_bounds__(256) void ssfm_independent(SsfmArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= a.cells) return;
    constexpr int NC = TETM ? 2 : 1;
    const PhxPhys<real>& ph = a.ph;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        if (a.npulse == 0 && !a.stochastic) continue;
        real dr = 0, di = 0;
        const real f = a.dt * ph.one_over_h_bar_s;
        for (int g = 0; g < a.npulse; g++) {
            const auto F = a.pulse[k][g][i];
            const auto am = a.lamp[g];
            /* (dt / hbar) F a_g */
            const real zr = f * phx_fma(F.x, am.x, -(F.y * am.y)), zi = f * phx_fma(F.x, am.y, F.y * am.x);
            if (TETM) { /* -i (dt / hbar) F a_g */
                dr += zi;
                di -= zr;
            } else {
                dr += zr;
                di += zi;
            }
        }
        if (a.stochastic) {
            /* scalar: the reservoir of the step start; TE-TM: the new one of the component */
            const real* nsrc = TETM ? a.n_nxt[k] : a.n_cur[0];
            const auto xi = a.noise[i];
            if constexpr (CRES) {
                /* principal complex square root of (R n + gamma_c) / (4 dV), complex product with xi */
                const v2 z = reinterpret_cast<const v2*>(nsrc)[i];
                const real d = real(4) * ph.dV;
                real sr, si;
                csqrt_principal<real>(phx_fma(ph.R, z.x, ph.gamma_c) / d, (ph.R * z.y) / d, sr, si);
                dr += xi.x * sr - xi.y * si;
                di += xi.x * si + xi.y * sr;
            } else {
                const real nr = nsrc ? nsrc[i] : real(0);
                const real sg = phx_sqrt(phx_fma(ph.R, nr, ph.gamma_c) / (real(4) * ph.dV));
                dr = phx_fma(xi.x, sg, dr);
                di = phx_fma(xi.y, sg, di);
            }
        }
        auto p = a.psi[k][i];
        p.x += dr;
        p.y += di;
        a.psi[k][i] = p;
    }
    /* scalar model: the new reservoir becomes current (TE-TM keeps the old one) */
    if (!TETM && a.n_nxt[0]) {
        if constexpr (CRES)
            reinterpret_cast<v2*>(a.n_cur[0])[i] = reinterpret_cast<const v2*>(a.n_nxt[0])[i];
        else
            a.n_cur[0][i] = a.n_nxt[0][i];
    }
}

template <typename real>
hipError_t launch_ssfm_nonlinear(const SsfmArgs<real>& a, hipStream_t s) {
    const dim3 grid((unsigned)((a.cells + 255) / 256));
    if (a.cres && !a.n_cur[0]) return hipErrorInvalidValue;
    if (a.tetm && a.cres)
        hipLaunchKernelGGL((ssfm_nonlinear<real, true, true>), grid, dim3(256), 0, s, a);
    else if (a.tetm)
        hipLaunchKernelGGL((ssfm_nonlinear<real, true, false>), grid, dim3(256), 0, s, a);
    else if (a.cres)
        hipLaunchKernelGGL((ssfm_nonlinear<real, false, true>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((ssfm_nonlinear<real, false, false>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_ssfm_independent(const SsfmArgs<real>& a, hipStream_t s) {
    const dim3 grid((unsigned)((a.cells + 255) / 256));
    if (a.cres && (!a.n_cur[0] || !a.n_nxt[0] || (a.tetm && !a.n_nxt[1]))) return hipErrorInvalidValue;
    if (a.tetm && a.cres)
        hipLaunchKernelGGL((ssfm_independent<real, true, true>), grid, dim3(256), 0, s, a);
    else if (a.tetm)
        hipLaunchKernelGGL((ssfm_independent<real, true, false>), grid, dim3(256), 0, s, a);
    else if (a.cres)
        hipLaunchKernelGGL((ssfm_independent<real, false, true>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((ssfm_independent<real, false, false>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_ssfm_nonlinear<float>(const SsfmArgs<float>&, hipStream_t);
template hipError_t launch_ssfm_nonlinear<double>(const SsfmArgs<double>&, hipStream_t);
template hipError_t launch_ssfm_independent<float>(const SsfmArgs<float>&, hipStream_t);
template hipError_t launch_ssfm_independent<double>(const SsfmArgs<double>&, hipStream_t);

} // namespace phx
