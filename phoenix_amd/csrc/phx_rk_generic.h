/*
 * phx_rk_generic.h -- the fused explicit Runge-Kutta kernel for any lower-triangular
 * Butcher tableau (phx_tableaux.h), scalar or TE-TM spinor model, with optional
 * stochastic noise and embedded error estimate.  Included by the translation units
 * that instantiate it for groups of tableaux (phx_rk_generic.hip, phx_rk_tab*.hip),
 * so that the instantiations compile in parallel; every (tableau, feature set)
 * instantiation lives in exactly one of them.
 *
 * Reference: the stage sequence of SOLVER_SEQUENCE with CALCULATE_K /
 * INTERMEDIATE_SUM_K / FINAL_SUM_K / ERROR_K (cuda_macro.cuh:110-250,306-353),
 * the iterators' step() functions (solver/iterator/<name>.cu), the RHS gp_scalar
 * (kernel_gp_compute.cuh:18-236) and the sums of kernel_summation.cuh:195-289.
 *
 * Mapping: one wavefront streams a strip of 64 extended columns top to bottom; with
 * S stages the halo is S cells, stage s is evaluated at row i-s after row i is
 * loaded, and stage s+1's input
 *   Y_s = psi0 + dt * sum_{j<=s} a[s][j] K_j (+ noise)
 * is formed as soon as K_s is known, so every K stays in registers until its last
 * use.  East/west neighbours come through DPP wave shifts.
 *
 * Outputs: FINAL mode writes psi0 + dt * sum b_j K_j; ADAPTIVE mode writes the same
 * and in addition the per-wavefront FP64 partial sums of |dt sum_j e_j K_j(psi+)|^2
 * and |psi+_new|^2 over the output cells, for a deterministic second-pass reduction.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_device.h"
#include "phx_kernels.h"
#include "phx_noise.h"
#include "phx_tableaux.h"

namespace phx {

__device__ __forceinline__ double rk_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

/* deterministic reduction of the per-wave partials (one workgroup, fixed order) into sums[0..1] */
hipError_t launch_reduce_partials(const double* partials, int nwaves, double* sums, hipStream_t s);

/* ---------------------------------------------------------------- kernel */
/* Template flags: TETM (two coupled components), RES (reservoir), STOCH (noise).  A pump or
   potential that is absent reads as 0 through a descriptor of size 0.  No lane-divergent
   branches: halo lanes, lanes and rows outside a zero-BC domain load through out-of-range
   offsets (0) and store through them (dropped); rows past the last grid row skip their store
   with a wave-uniform branch. */
/* POT = false: no potential (the potential's loads and descriptors compiled out; a.pot must be null) */
template <typename real, typename TAB, int TY, bool TETM, bool RES, bool STOCH, bool POT = true>
__global__ __launch_bounds__(256) void rk_generic_fused(RkGenArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    constexpr int S = TAB::S;
    constexpr int H = S;          /* halo = number of K evaluations */
    constexpr int E = TY + 2 * H; /* extended rows */
    constexpr int NC = TETM ? 2 : 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const int strip = tile % a.nstrips;
    const int band = (tile / a.nstrips) * kWavesPerBlock + wave;
    if (band >= a.nbands) return;
    const int outw = kWave - 2 * H;
    const int x0 = strip * outw;
    const int y0 = band * TY;
    const int Nc = a.Nc, Nr = a.Nr;

    /* column of this lane: inside, wrapped (periodic x) or outside a zero-BC domain */
    const int xcol = x0 - H + lane;
    const bool inside = xcol >= 0 && xcol < Nc;
    const bool col_in = inside || a.periodic_x;
    const int gcol = inside ? xcol : ((xcol % Nc) + Nc) % Nc;
    const uint32_t off_r = col_in ? (uint32_t)gcol * (uint32_t)sizeof(real) : kOutOfRange;
    const uint32_t off_c = col_in ? (uint32_t)gcol * (uint32_t)sizeof(v2) : kOutOfRange;
    /* store offsets: halo lanes and lanes past the last column write nothing */
    const bool out_lane = lane >= H && lane < kWave - H && xcol < Nc;
    const uint32_t so_r = out_lane ? (uint32_t)xcol * (uint32_t)sizeof(real) : kOutOfRange;
    const uint32_t so_c = out_lane ? (uint32_t)xcol * (uint32_t)sizeof(v2) : kOutOfRange;
    /* one descriptor per array for the whole kernel; the row is part of the lane offset, and
       a row outside a zero-BC domain gets the out-of-range offset (the sum of an in-range row
       offset and kOutOfRange stays out of range: every array is below 2 GiB) */
    const uint32_t row_r = (uint32_t)Nc * (uint32_t)sizeof(real);
    const uint32_t row_c = (uint32_t)Nc * (uint32_t)sizeof(v2);
    const uint32_t all_r = row_r * (uint32_t)Nr, all_c = row_c * (uint32_t)Nr;
    rsrc_t d_psi[NC], d_n[NC], d_p[NC], d_v[NC], o_psi[NC], o_n[NC];
    const rsrc_t d_xi = row_rsrc(a.noise, STOCH ? all_c : 0u);
    const RowWrap wrap = make_row_wrap(Nr, H);
#pragma unroll
    for (int k = 0; k < NC; k++) {
        d_psi[k] = row_rsrc(a.psi_in[k], all_c);
        d_n[k] = row_rsrc(a.n_in[k], RES ? all_r : 0u);
        d_p[k] = row_rsrc(a.pump[k], a.pump[k] ? all_r : 0u);
        d_v[k] = row_rsrc(a.pot[k], a.pot[k] ? all_r : 0u);
        o_psi[k] = row_rsrc(a.psi_out[k], all_c);
        o_n[k] = row_rsrc(a.n_out[k], RES ? all_r : 0u);
    }

    PhxPhys<real> ph = a.ph;
    ph.mhalf_gamma_c = in_vgpr(ph.mhalf_gamma_c);
    ph.gamma_r = in_vgpr(ph.gamma_r);
    const real dt = a.dt;
    /* psi advances by dt (real time) or by -i dt (imaginary time, --imagTime):
       psi0 + (dtr + i dti) S with (dtr, dti) = (dt, 0) or (0, -dt); the reservoir always by dt */
    const real dtr = a.imag ? real(0) : dt;
    const real dti = a.imag ? -dt : real(0);

    /* per component, per extended row */
    real p0r[NC][E], p0i[NC][E], n0[NC][E], pp[NC][E], vv[NC][E];
    real Kr[NC][S][E], Ki[NC][S][E], Kn[NC][S][E];
    real Yr[NC][S][E], Yi[NC][S][E], Yn[NC][S][E]; /* Y[s] = input of stage s+1, s = 0..S-2 */
    real xr[E], xi[E], sg[E];                      /* noise xi and its amplitude sqrt((R n0 + gc)/(4 dV)) */
    double err_acc = 0.0, nrm_acc = 0.0;

    auto load_row = [&](int r) {
        bool in;
        const uint32_t grow = tile_row(y0 - H + r, wrap, a.periodic_y, in); /* wave-uniform */
        const uint32_t vc = in ? off_c + grow * row_c : kOutOfRange;
        const uint32_t vr = in ? off_r + grow * row_r : kOutOfRange;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const v2 v = load_cplx_s<real>(d_psi[k], vc, 0u);
            p0r[k][r] = v.x;
            p0i[k][r] = v.y;
            n0[k][r] = RES ? load_real_s<real>(d_n[k], vr, 0u) : real(0);
            pp[k][r] = load_real_s<real>(d_p[k], vr, 0u);
            vv[k][r] = POT ? load_real_s<real>(d_v[k], vr, 0u) : real(0);
        }
        if (STOCH) {
            /* this step's noise field (launch_noise_field); 0 outside a zero-BC domain */
            const v2 z = load_cplx_s<real>(d_xi, vc, 0u);
            xr[r] = z.x;
            xi[r] = z.y;
            const real nplus = RES ? n0[0][r] : real(0);
            sg[r] = phx_sqrt((RES ? phx_fma(ph.R, nplus, ph.gamma_c) : ph.gamma_c) / (real(4) * ph.dV));
        }
    };

    /* input of stage s (0-based) at row e, component k */
    auto in_r = [&](int k, int s, int e) -> real { return s == 0 ? p0r[k][e] : Yr[k][s - 1][e]; };
    auto in_i = [&](int k, int s, int e) -> real { return s == 0 ? p0i[k][e] : Yi[k][s - 1][e]; };
    auto in_n = [&](int k, int s, int e) -> real { return s == 0 ? n0[k][e] : Yn[k][s - 1][e]; };

    /* stage s at row e: K_s from the stage input at rows e-1..e+1 */
    auto stage = [&](int s, int e) {
        const real w = real(TAB::c[s]);
        const real pamp = phx_fma(a.amp_next_pump - a.amp_pump, w, a.amp_pump);
        const real vamp_h = phx_fma(a.amp_next_pot_h - a.amp_pot_h, w, a.amp_pot_h);
        real ewr[NC], ewi[NC], nsr[NC], nsi[NC], cr_[NC], ci_[NC], nrm[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const real cr = in_r(k, s, e), ci = in_i(k, s, e);
            ewr[k] = east(cr) + west(cr);
            ewi[k] = east(ci) + west(ci);
            nsr[k] = in_r(k, s, e - 1) + in_r(k, s, e + 1);
            nsi[k] = in_i(k, s, e - 1) + in_i(k, s, e + 1);
            cr_[k] = cr;
            ci_[k] = ci;
            nrm[k] = phx_fma(cr, cr, ci * ci);
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const real nv = RES ? in_n(k, s, e) : real(0);
            const real pmp = pp[k][e] * pamp;
            const real poth = POT ? vv[k][e] * vamp_h : real(0);
            real kr, ki, kn;
            if constexpr (TETM) {
                /* cross term of the other component o (kernel_gp_compute.cuh:123-130):
                   (E+W)_o/dx^2 - (N+S)_o/dy^2 +- (i/2)/(dx dy) (SE + NW - SW - NE)_o */
                const int o = 1 - k;
                const real dr = in_r(o, s, e + 1) - in_r(o, s, e - 1);
                const real di = in_i(o, s, e + 1) - in_i(o, s, e - 1);
                const real Dr = east(dr) - west(dr), Di = east(di) - west(di);
                const real sgn = k == 0 ? real(1) : real(-1);
                /* delta_LT/hbar * cross, with (i c D) = (-c D_i, c D_r) */
                const real hx = ph.lt_h * ph.one_over_dx2, hy = ph.lt_h * ph.one_over_dy2;
                const real ccr = phx_fma(hx, ewr[o], phx_fma(-hy, nsr[o], -sgn * ph.lt_hxy * Di));
                const real cci = phx_fma(hx, ewi[o], phx_fma(-hy, nsi[o], sgn * ph.lt_hxy * Dr));
                const real tr = phx_fma(ph.kin_x, ewr[k], phx_fma(ph.kin_y, nsr[k], ph.kin_c * cr_[k]));
                const real ti = phx_fma(ph.kin_x, ewi[k], phx_fma(ph.kin_y, nsi[k], ph.kin_c * ci_[k]));
                phx_rhs_tetm<real>(ph, cr_[k], ci_[k], tr, ti, ccr, cci, nrm[o], nv, pmp, poth, kr, ki, kn);
            } else {
                /* neighbour part only: the hook adds the centre term */
                const real tr = phx_fma(ph.kin_x, ewr[k], ph.kin_y * nsr[k]);
                const real ti = phx_fma(ph.kin_x, ewi[k], ph.kin_y * nsi[k]);
                phx_rhs_scalar<real>(ph, cr_[k], ci_[k], tr, ti, nv, pmp, poth, kr, ki, kn);
            }
            if (STOCH) {
                /* stochastic corrections of the RHS (kernel_gp_compute.cuh:83-85,106-108):
                   k_psi -= g_c/hbar (-i psi)/dV ; k_n += R n / dV */
                const real g = ph.gc_h / ph.dV;
                kr -= g * ci_[k];
                ki += g * cr_[k];
                if (RES) kn = phx_fma(ph.R / ph.dV, nv, kn);
            }
            Kr[k][s][e] = kr;
            Ki[k][s][e] = ki;
            Kn[k][s][e] = kn;
        }
        /* input of the next stage: Y_s = psi0 + dt sum_{j<=s} a[s+1][j] K_j (+ noise) */
        if (s + 1 < S) {
            real wsum = 0;
#pragma unroll
            for (int j = 0; j <= s; j++) wsum += real(TAB::a[s + 1][j]);
#pragma unroll
            for (int k = 0; k < NC; k++) {
                real sr = 0, si = 0, sn = 0;
                bool first = true;
#pragma unroll
                for (int j = 0; j <= s; j++) {
                    const real wj = real(TAB::a[s + 1][j]);
                    if (TAB::a[s + 1][j] == 0.0) continue;
                    if (first) {
                        sr = wj * Kr[k][j][e];
                        si = wj * Ki[k][j][e];
                        sn = wj * Kn[k][j][e];
                        first = false;
                    } else {
                        sr = phx_fma(wj, Kr[k][j][e], sr);
                        si = phx_fma(wj, Ki[k][j][e], si);
                        sn = phx_fma(wj, Kn[k][j][e], sn);
                    }
                }
                real yr = phx_fma(dtr, sr, phx_fma(-dti, si, p0r[k][e]));
                real yi = phx_fma(dtr, si, phx_fma(dti, sr, p0i[k][e]));
                if (STOCH) {
                    yr = phx_fma(wsum * xr[e], sg[e], yr);
                    yi = phx_fma(wsum * xi[e], sg[e], yi);
                }
                Yr[k][s][e] = yr;
                Yi[k][s][e] = yi;
                Yn[k][s][e] = RES ? phx_fma(dt, sn, n0[k][e]) : real(0);
            }
        }
    };

    /* after the last stage at row e: final sum, store, error partials */
    auto finish = [&](int e) {
        const int grow = y0 + e - H; /* wave-uniform */
        real out0r = 0, out0i = 0;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            real sr = 0, si = 0, sn = 0;
            bool first = true;
            real bsum = 0;
#pragma unroll
            for (int j = 0; j < TAB::NB; j++) {
                bsum += real(TAB::b[j]);
                if (TAB::b[j] == 0.0) continue;
                const real wj = real(TAB::b[j]);
                if (first) {
                    sr = wj * Kr[k][j][e];
                    si = wj * Ki[k][j][e];
                    sn = wj * Kn[k][j][e];
                    first = false;
                } else {
                    sr = phx_fma(wj, Kr[k][j][e], sr);
                    si = phx_fma(wj, Ki[k][j][e], si);
                    sn = phx_fma(wj, Kn[k][j][e], sn);
                }
            }
            real o_r = phx_fma(dtr, sr, phx_fma(-dti, si, p0r[k][e]));
            real o_i = phx_fma(dtr, si, phx_fma(dti, sr, p0i[k][e]));
            if (STOCH) {
                o_r = phx_fma(bsum * xr[e], sg[e], o_r);
                o_i = phx_fma(bsum * xi[e], sg[e], o_i);
            }
            if (k == 0) {
                out0r = o_r;
                out0i = o_i;
            }
            /* a row past the grid end has an offset beyond the array (launch_gen checks that
               it does not wrap): the store is dropped, no branch */
            v2 o;
            o.x = o_r;
            o.y = o_i;
            store_cplx_s(o_psi[k], so_c + (uint32_t)grow * row_c, 0u, o);
            if (RES) store_real_s(o_n[k], so_r + (uint32_t)grow * row_r, 0u, phx_fma(dt, sn, n0[k][e]));
        }
        if constexpr (TAB::ERR) {
            /* runge_sum_to_error (kernel_summation.cuh:270-289): |dt sum_j e_j K_j(psi+)|^2;
               with -tetm the reference adds the psi+ error twice (cuda_macro.cuh:244) */
            real sr = 0, si = 0;
#pragma unroll
            for (int j = 0; j < S; j++) {
                if (TAB::e[j] == 0.0) continue;
                sr = phx_fma(real(TAB::e[j]), Kr[0][j][e], sr);
                si = phx_fma(real(TAB::e[j]), Ki[0][j][e], si);
            }
            const double er = (double)(dt * sr), ei = (double)(dt * si);
            const double ee = er * er + ei * ei;
            const bool cell = out_lane && grow < Nr;
            err_acc += cell ? (TETM ? 2.0 * ee : ee) : 0.0;
            nrm_acc += cell ? (double)out0r * (double)out0r + (double)out0i * (double)out0i : 0.0;
        }
    };

#pragma unroll
    for (int it = -2; it < E; ++it) {
        if (it + 2 < E) load_row(it + 2);
        if (it < 0) continue;
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = it - (s + 1);
            if (e >= s + 1 && e <= E - 2 - s) stage(s, e);
        }
        const int ef = it - S;
        if (ef >= H && ef <= E - 1 - H) finish(ef);
    }

    if constexpr (TAB::ERR) {
        const double es = rk_wave_sum(err_acc), ns = rk_wave_sum(nrm_acc);
        if (lane == 0) {
            const int gw = band * a.nstrips + strip;
            a.partials[2 * gw] = es;
            a.partials[2 * gw + 1] = ns;
        }
    }
}

/* ---------------------------------------------------------------- launch */
template <typename real, typename TAB, int TY, bool TETM, bool RES, bool STOCH, bool POT = true>
static hipError_t launch_gen(const RkGenArgs<real>& a0, hipStream_t s) {
    RkGenArgs<real> a = a0;
    /* whole-array descriptors: every field must stay below the 32-bit offset range */
    /* (rows up to TY + S past the grid end are addressed, and dropped, by the stores) */
    if ((size_t)a.Nc * (size_t)(a.Nr + TY + TAB::S) * sizeof(typename Vec2<real>::type) > kMaxArrayBytes)
        return hipErrorInvalidValue;
    const int outw = kWave - 2 * TAB::S;
    a.nstrips = (a.Nc + outw - 1) / outw;
    a.nbands = (a.Nr + TY - 1) / TY;
    const int bandgroups = (a.nbands + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((rk_generic_fused<real, TAB, TY, TETM, RES, STOCH, POT>), dim3(a.nstrips * bandgroups),
                       dim3(kWave * kWavesPerBlock), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !TAB::ERR) return e;
    return launch_reduce_partials(a.partials, a.nstrips * a.nbands, a.sums, s);
}

/* every feature set: (TE-TM, reservoir, noise); pump and potential are runtime inputs */
template <typename real, typename TAB, int TY>
static hipError_t dispatch_flags(const RkGenArgs<real>& a, hipStream_t s) {
    const bool tetm = a.tetm, res = a.n_in[0] != nullptr, st = a.stochastic;
#define PHX_GEN(T, R, N) \
    if (tetm == T && res == R && st == N) return launch_gen<real, TAB, TY, T, R, N>(a, s);
    PHX_GEN(false, false, false)
    PHX_GEN(false, true, false)
    PHX_GEN(false, false, true)
    PHX_GEN(false, true, true)
    PHX_GEN(true, false, false)
    PHX_GEN(true, true, false)
    PHX_GEN(true, false, true)
    PHX_GEN(true, true, true)
#undef PHX_GEN
    return hipErrorNotSupported;
}

/* the kernel for one tableau, every feature set, tile height 16 */
template <typename real, typename TAB>
hipError_t launch_tab(const RkGenArgs<real>& a, hipStream_t s) {
    return dispatch_flags<real, TAB, 16>(a, s);
}

} // namespace phx
