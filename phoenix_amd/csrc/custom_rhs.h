/*
 * custom_rhs.h -- the per-cell right-hand side of the GP + reservoir model.
 *
 * This is the build's "custom kernel" hook, the counterpart of editing
 * Kernel::Compute::gp_scalar (include/kernel/kernel_gp_compute.cuh:18-236) as
 * described in custom_kernel.md:1-171.  The fused RK kernels compute the
 * stencil (the 5-point Laplacian sum, and for TE-TM the 9-point cross term)
 * in registers and call the functions below once per cell and stage, so an
 * edit here changes every iterator, FP32 and FP64, with no second code path.
 *
 * Custom parameters: add X(name, default) to PHX_CUSTOM_PARAMETERS.  Each entry
 * becomes a `real name` member of PhxPhys (available below as p.name) and a
 * `--name <value>` command line flag, like the reference's "Custom Parameters
 * go here!" (system_parameters.hpp:59-62) and "Custom Read-Ins go here!"
 * (system_initialization.cpp:222-224).
 */
#pragma once

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

#define PHX_CUSTOM_PARAMETERS(X) /* X(custom_var, 0.0) */

template <typename real>
struct PhxPhys {
    real m_eff_scaled;      /* -hbar_s^2 / (2 m_eff)                  (system.cpp:151) */
    real one_over_h_bar_s;  /* 1 / hbar_s                              (system.cpp:157) */
    real m2_over_dx2_p_dy2; /* -2 (1/dx^2 + 1/dy^2)                    (system.cpp:139) */
    real one_over_dx2, one_over_dy2;
    real dx, dy, dV;
    real gamma_c, gamma_r, g_c, g_r, R, g_pm, delta_LT;
#define PHX_X(name, def) real name;
    PHX_CUSTOM_PARAMETERS(PHX_X)
#undef PHX_X
};

/*
 * Scalar model (kernel_gp_compute.cuh:50-110, non-stochastic branch):
 *   k_psi = -i/hbar [ m~ lap + (g_c|psi|^2 + g_r n + V) psi ] + 1/2 (R n - gamma_c) psi
 *   k_n   = -(gamma_r + R|psi|^2) n + P
 * (pr, pi) = psi, (lr, li) = C psi + (N+S)/dy^2 + (E+W)/dx^2, n reservoir,
 * pump = sum_k P_k a_k(w), pot = sum_k V_k a_k(w).
 * Without a reservoir the caller passes n = 0 and ignores kn.
 */
template <typename real>
__host__ __device__ __forceinline__ void phx_rhs_scalar(const PhxPhys<real>& p, real pr, real pi, real lr, real li,
                                                        real n, real pump, real pot, real& kr, real& ki, real& kn) {
    const real nrm = pr * pr + pi * pi;
    const real hk = p.m_eff_scaled * p.one_over_h_bar_s;
    /* real potential felt by psi, divided by hbar */
    const real a = p.one_over_h_bar_s * (p.g_c * nrm + p.g_r * n + pot);
    /* gain - loss */
    const real b = real(0.5) * (p.R * n - p.gamma_c);
    /* -i (hk*lap + a*psi) + b*psi */
    kr = hk * li + a * pi + b * pr;
    ki = -(hk * lr) - a * pr + b * pi;
    kn = -(p.gamma_r + p.R * nrm) * n + pump;
}

/*
 * TE-TM spinor model, one component (kernel_gp_compute.cuh:137-234):
 *   k_psi = -i/hbar [ m~ lap_s + (g_c|psi_s|^2 + g_r n_s + V_s + g_pm|psi_o|^2) psi_s + delta_LT cross_s ]
 *           + 1/2 (R n_s - gamma_c) psi_s
 *   k_n   = -(gamma_r + R|psi_s|^2) n_s + P_s
 * (cr, ci) = cross_s = (E+W)_o/dx^2 - (N+S)_o/dy^2 +- (i/2)/(dx dy) D_o, computed by the caller
 * (the centre terms cancel only for dx == dy; the reference omits them, and so does this hook).
 */
template <typename real>
__host__ __device__ __forceinline__ void phx_rhs_tetm(const PhxPhys<real>& p, real pr, real pi, real lr, real li,
                                                      real cr, real ci, real other_norm, real n, real pump, real pot,
                                                      real& kr, real& ki, real& kn) {
    const real nrm = pr * pr + pi * pi;
    const real hk = p.m_eff_scaled * p.one_over_h_bar_s;
    const real hl = p.delta_LT * p.one_over_h_bar_s;
    const real a = p.one_over_h_bar_s * (p.g_c * nrm + p.g_r * n + pot + p.g_pm * other_norm);
    const real b = real(0.5) * (p.R * n - p.gamma_c);
    kr = hk * li + hl * ci + a * pi + b * pr;
    ki = -(hk * lr) - hl * cr - a * pr + b * pi;
    kn = -(p.gamma_r + p.R * nrm) * n + pump;
}
