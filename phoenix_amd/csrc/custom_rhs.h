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

/* Fused multiply-add written out: the kernels are compiled without automatic
   contraction, so every cell rounds identically whatever the tile shape. */
__host__ __device__ __forceinline__ float phx_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ __forceinline__ double phx_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename real>
struct PhxPhys {
    real m_eff_scaled;      /* -hbar_s^2 / (2 m_eff)                  (system.cpp:151) */
    real one_over_h_bar_s;  /* 1 / hbar_s                              (system.cpp:157) */
    real m2_over_dx2_p_dy2; /* -2 (1/dx^2 + 1/dy^2)                    (system.cpp:139) */
    real one_over_dx2, one_over_dy2;
    real dx, dy, dV;
    real gamma_c, gamma_r, g_c, g_r, R, g_pm, delta_LT;
    /* folded constants, filled on the host from the members above (phx_fold) */
    real kin_c, kin_x, kin_y; /* (m~/hbar) * {C, 1/dx^2, 1/dy^2}: kinetic stencil weights */
    real kin_yx;              /* kin_y / kin_x = dx^2 / dy^2 (exactly 1 for dx == dy) */
    real gc_h, gr_h, gpm_h;   /* g_c/hbar, g_r/hbar, g_pm/hbar */
    real lt_h, lt_hxy;        /* delta_LT/hbar, delta_LT/hbar * 1/2 /dx/dy */
    real half_R, mhalf_gamma_c; /* R/2, -gamma_c/2 */
#define PHX_X(name, def) real name;
    PHX_CUSTOM_PARAMETERS(PHX_X)
#undef PHX_X
};

/* Derive the folded constants (host side, in double before rounding to real). */
template <typename real>
inline void phx_fold(PhxPhys<real>& p, double m_eff_scaled, double one_over_h_bar_s, double m2, double idx2, double idy2,
                     double dx, double dy) {
    const double hk = m_eff_scaled * one_over_h_bar_s;
    p.kin_c = (real)(hk * m2);
    p.kin_x = (real)(hk * idx2);
    p.kin_y = (real)(hk * idy2);
    p.kin_yx = (real)(idy2 / idx2);
    p.gc_h = (real)((double)p.g_c * one_over_h_bar_s);
    p.gr_h = (real)((double)p.g_r * one_over_h_bar_s);
    p.gpm_h = (real)((double)p.g_pm * one_over_h_bar_s);
    p.lt_h = (real)((double)p.delta_LT * one_over_h_bar_s);
    p.lt_hxy = (real)((double)p.delta_LT * one_over_h_bar_s * 0.5 / dx / dy);
    p.half_R = (real)(0.5 * (double)p.R);
    p.mhalf_gamma_c = (real)(-0.5 * (double)p.gamma_c);
}

/*
 * Scalar model (kernel_gp_compute.cuh:50-110, non-stochastic branch):
 *   k_psi = -i/hbar [ m~ lap + (g_c|psi|^2 + g_r n + V) psi ] + 1/2 (R n - gamma_c) psi
 *   k_n   = -(gamma_r + R|psi|^2) n + P
 * Inputs: psi = (pr, pi); (tr, ti) = the neighbour part of (m~/hbar) lap psi,
 * kin_x (E+W) + kin_y (N+S) with lap psi = C psi + (N+S)/dy^2 + (E+W)/dx^2
 * (the centre term kin_c psi is added here, folded into the potential:
 * -i (kin_c psi + T_nb + a psi) = -i ((a + kin_c) psi + T_nb)); n reservoir;
 * pump = sum_k P_k a_k(w); pot_h = sum_k V_k a_k(w) / hbar.
 * Without a reservoir the caller passes n = 0 and ignores kn.
 */
template <typename real>
__host__ __device__ __forceinline__ void phx_rhs_scalar(const PhxPhys<real>& p, real pr, real pi, real tr, real ti,
                                                        real n, real pump, real pot_h, real& kr, real& ki, real& kn) {
    const real nrm = phx_fma(pr, pr, pi * pi);
    /* real potential felt by psi, over hbar, plus the stencil centre */
    const real a = phx_fma(p.gc_h, nrm, phx_fma(p.gr_h, n, pot_h + p.kin_c));
    /* gain - loss */
    const real b = phx_fma(p.half_R, n, p.mhalf_gamma_c);
    /* -i (T_nb + a psi) + b psi */
    kr = phx_fma(b, pr, phx_fma(a, pi, ti));
    ki = phx_fma(b, pi, phx_fma(-a, pr, -tr));
    kn = phx_fma(phx_fma(-p.R, nrm, -p.gamma_r), n, pump);
}

/*
 * Transpose of the Jacobian of phx_rhs_scalar (the adjoint RK4 kernel calls it
 * once per cell and stage; an edit of phx_rhs_scalar needs the matching edit
 * here).  Given the cotangent (cr, ci, cn) of (kr, ki, kn) at the state
 * (pr, pi, n) and the neighbour part of the kinetic stencil applied to the
 * cotangent, (Kr, Ki) = kin_x (E+W) + kin_y (N+S) of c (the 5-point stencil is
 * symmetric, so it is its own transpose; its centre term kin_c c is added here,
 * folded into a as in phx_rhs_scalar), returns the cotangent (gr, gi, gn) of
 * (pr, pi, n):
 *   gr = cr (b + 2 g uv) - ci (a + 2 g u^2) - cn 2 R n u - Ki
 *   gi = cr (a + 2 g v^2) + ci (b - 2 g uv) - cn 2 R n v + Kr
 *   gn = cr (R/2 u + g_r/hbar v) + ci (R/2 v - g_r/hbar u) - cn (gamma_r + R |psi|^2)
 * with u = pr, v = pi, g = g_c/hbar, a = g |psi|^2 + g_r/hbar n + pot_h + kin_c
 * and b as in phx_rhs_scalar.
 */
template <typename real>
__host__ __device__ __forceinline__ void phx_rhs_scalar_vjp(const PhxPhys<real>& p, real pr, real pi, real n, real pot_h,
                                                            real cr, real ci, real cn, real Kr, real Ki, real& gr,
                                                            real& gi, real& gn) {
    const real nrm = phx_fma(pr, pr, pi * pi);
    const real a = phx_fma(p.gc_h, nrm, phx_fma(p.gr_h, n, pot_h + p.kin_c));
    const real b = phx_fma(p.half_R, n, p.mhalf_gamma_c);
    const real g2 = p.gc_h + p.gc_h;
    const real uv2 = g2 * (pr * pi);
    const real rn2 = (p.R + p.R) * n;
    gr = phx_fma(cr, b + uv2, phx_fma(-ci, phx_fma(g2 * pr, pr, a), phx_fma(-cn * rn2, pr, -Ki)));
    gi = phx_fma(cr, phx_fma(g2 * pi, pi, a), phx_fma(ci, b - uv2, phx_fma(-cn * rn2, pi, Kr)));
    gn = phx_fma(cr, phx_fma(p.half_R, pr, p.gr_h * pi), phx_fma(ci, phx_fma(p.half_R, pi, -(p.gr_h * pr)), -(cn * phx_fma(p.R, nrm, p.gamma_r))));
}

/* Cotangents of the pump and potential values of one RHS evaluation:
   dk/dP = pump amplitude on kn, dk/dV = amplitude/hbar * (pi, -pr) on (kr, ki). */
template <typename real>
__host__ __device__ __forceinline__ void phx_rhs_scalar_param_vjp(real pr, real pi, real cr, real ci, real cn, real pamp,
                                                                  real vamp_h, real& gp, real& gv) {
    gp = pamp * cn;
    gv = vamp_h * phx_fma(cr, pi, -(ci * pr));
}

/*
 * TE-TM spinor model, one component (kernel_gp_compute.cuh:137-234):
 *   k_psi = -i/hbar [ m~ lap_s + (g_c|psi_s|^2 + g_r n_s + V_s + g_pm|psi_o|^2) psi_s + delta_LT cross_s ]
 *           + 1/2 (R n_s - gamma_c) psi_s
 *   k_n   = -(gamma_r + R|psi_s|^2) n_s + P_s
 * (cr, ci) = delta_LT/hbar * cross_s with
 * cross_s = (E+W)_o/dx^2 - (N+S)_o/dy^2 +- (i/2)/(dx dy) D_o (o = other component),
 * evaluated by the caller; the centre terms cancel only for dx == dy, and the
 * reference omits them, as does this hook.
 */
template <typename real>
__host__ __device__ __forceinline__ void phx_rhs_tetm(const PhxPhys<real>& p, real pr, real pi, real tr, real ti,
                                                      real cr, real ci, real other_norm, real n, real pump, real pot_h,
                                                      real& kr, real& ki, real& kn) {
    const real nrm = phx_fma(pr, pr, pi * pi);
    const real a = phx_fma(p.gpm_h, other_norm, phx_fma(p.gc_h, nrm, phx_fma(p.gr_h, n, pot_h)));
    const real b = phx_fma(p.half_R, n, p.mhalf_gamma_c);
    kr = phx_fma(b, pr, phx_fma(a, pi, ti + ci));
    ki = phx_fma(b, pi, phx_fma(-a, pr, -tr - cr));
    kn = phx_fma(phx_fma(-p.R, nrm, -p.gamma_r), n, pump);
}
