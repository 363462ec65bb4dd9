/*
 * phx_rk_generic.hip -- fused explicit Runge-Kutta step for any lower-triangular
 * Butcher tableau, scalar or TE-TM spinor model, with optional stochastic noise
 * and embedded error estimate.  Used for Dormand-Prince 5(4) (fixed and
 * adaptive) and RK4 with noise or TE-TM; the plain scalar RK4 has its own
 * hand-tuned kernel (phx_rk4.hip).
 *
 * Reference: the stage sequence of SOLVER_SEQUENCE with CALCULATE_K /
 * INTERMEDIATE_SUM_K / FINAL_SUM_K / ERROR_K (cuda_macro.cuh:110-250,306-353),
 * DormandPrince45::step (dormand_prince.cu:37-95), the RHS gp_scalar
 * (kernel_gp_compute.cuh:18-236) and the sums of kernel_summation.cuh:195-289.
 *
 * Mapping as in phx_rk4.hip: one wavefront streams a strip of 64 extended
 * columns top to bottom; with S stages the halo is S cells, stage s is
 * evaluated at row i-s after row i is loaded, and stage s+1's input
 *   Y_s = psi0 + dt * sum_{j<=s} a[s][j] K_j (+ noise)
 * is formed as soon as K_s is known, so every K stays in registers until its
 * last use.  East/west neighbours come through DPP wave shifts.
 *
 * Outputs: FINAL mode writes psi0 + dt * sum b_j K_j; ADAPTIVE mode writes the
 * same (the 5th-order solution, = input of the last stage) and in addition the
 * per-wavefront FP64 partial sums of |dt sum_j e_j K_j(psi+)|^2 and |psi+_new|^2
 * over the output cells, for a deterministic second-pass reduction.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_device.h"
#include "phx_kernels.h"
#include "phx_noise.h"

namespace phx {

namespace {
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
} // namespace

/* ---------------------------------------------------------------- tableaux */
/* Coefficients as in the reference iterators; the kernel rounds them to real,
   like the Type::real template arguments of the summation kernels. */
struct TabRK4 {
    static constexpr int S = 4, NB = 4;
    static constexpr bool ERR = false;
    static constexpr double a[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0}};
    static constexpr double c[4] = {0.0, 0.5, 0.5, 1.0};
    static constexpr double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    static constexpr double e[4] = {0, 0, 0, 0};
};

/* Dormand-Prince 5(4) (dormand_prince.cu:44-80); with ERR the 7th stage is evaluated
   on the 5th-order solution for the embedded error. */
struct TabDP45Base {
    static constexpr double a[7][7] = {
        {0, 0, 0, 0, 0, 0, 0},
        {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
        {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
        {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
        {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
        {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
        {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
    static constexpr double c[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
    static constexpr double b[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0};
    static constexpr double e[7] = {35.0 / 384.0 - 5179.0 / 57600.0, 0.0, 500.0 / 1113.0 - 7571.0 / 16695.0, 125.0 / 192.0 - 393.0 / 640.0,
                                    -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0, -1.0 / 40.0};
};
struct TabDP45Fixed : TabDP45Base {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = false;
};
struct TabDP45Adaptive : TabDP45Base {
    static constexpr int S = 7, NB = 6;
    static constexpr bool ERR = true;
};

/* ---------------------------------------------------------------- kernel */
template <typename real, typename TAB, int TY, bool TETM, bool RES, bool PUMP, bool POT, bool STOCH>
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

    int gcol = x0 - H + lane;
    bool col_in = true;
    if (gcol < 0 || gcol >= Nc) {
        if (a.periodic_x)
            gcol = ((gcol % Nc) + Nc) % Nc;
        else
            col_in = false;
    }
    const uint32_t off_r = col_in ? (uint32_t)gcol * (uint32_t)sizeof(real) : kOutOfRange;
    const uint32_t off_c = col_in ? (uint32_t)gcol * (uint32_t)sizeof(v2) : kOutOfRange;
    const bool out_lane = lane >= H && lane < kWave - H && (x0 - H + lane) < Nc;
    const uint32_t rbytes = (uint32_t)Nc * (uint32_t)sizeof(real);
    const uint32_t cbytes = (uint32_t)Nc * (uint32_t)sizeof(v2);

    PhxPhys<real> ph = a.ph;
    ph.mhalf_gamma_c = in_vgpr(ph.mhalf_gamma_c);
    ph.gamma_r = in_vgpr(ph.gamma_r);
    const real dt = a.dt;

    /* per component, per extended row */
    real p0r[NC][E], p0i[NC][E], n0[NC][E], pp[NC][E], vv[NC][E];
    real Kr[NC][S][E], Ki[NC][S][E], Kn[NC][S][E];
    real Yr[NC][S][E], Yi[NC][S][E], Yn[NC][S][E]; /* Y[s] = input of stage s+1, s = 0..S-2 */
    real xr[E], xi[E], sg[E];                      /* noise xi and its amplitude sqrt((R n0 + gc)/(4 dV)) */
    double err_acc = 0.0, nrm_acc = 0.0;

    auto load_row = [&](int r) {
        int grow = y0 - H + r;
        bool in = true;
        if (grow < 0 || grow >= Nr) {
            if (a.periodic_y)
                grow = ((grow % Nr) + Nr) % Nr;
            else
                in = false;
        }
        const size_t row = (size_t)(in ? grow : 0) * (size_t)Nc;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const v2 v = load_cplx<real>(row_rsrc(a.psi_in[k] + row, in ? cbytes : 0u), off_c);
            p0r[k][r] = v.x;
            p0i[k][r] = v.y;
            n0[k][r] = RES ? load_real<real>(row_rsrc(a.n_in[k] + row, in ? rbytes : 0u), off_r) : real(0);
            pp[k][r] = PUMP ? load_real<real>(row_rsrc(a.pump[k] + row, in ? rbytes : 0u), off_r) : real(0);
            vv[k][r] = POT ? load_real<real>(row_rsrc(a.pot[k] + row, in ? rbytes : 0u), off_r) : real(0);
        }
        if (STOCH) {
            /* own noise stream: Philox4x32-10 keyed by the seed, counter (cell, step); 0 outside a zero-BC domain */
            real zr = 0, zi = 0;
            if (in && col_in) phx_noise_normal2<real>(a.noise_seed, a.noise_step, (uint32_t)(grow * Nc + gcol), 0, zr, zi);
            xr[r] = zr * a.noise_amp;
            xi[r] = zi * a.noise_amp;
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
            const real nv = RES || TETM ? in_n(k, s, e) : real(0);
            const real pmp = PUMP ? pp[k][e] * pamp : real(0);
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
                real yr = phx_fma(dt, sr, p0r[k][e]);
                real yi = phx_fma(dt, si, p0i[k][e]);
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
        const int grow = y0 + e - H;
        const bool store = out_lane && grow < Nr;
        const size_t idx = (size_t)(store ? grow : 0) * (size_t)Nc + (size_t)(x0 - H + lane);
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
            real o_r = phx_fma(dt, sr, p0r[k][e]);
            real o_i = phx_fma(dt, si, p0i[k][e]);
            if (STOCH) {
                o_r = phx_fma(bsum * xr[e], sg[e], o_r);
                o_i = phx_fma(bsum * xi[e], sg[e], o_i);
            }
            if (k == 0) {
                out0r = o_r;
                out0i = o_i;
            }
            if (store) {
                v2 o;
                o.x = o_r;
                o.y = o_i;
                a.psi_out[k][idx] = o;
                if (RES) a.n_out[k][idx] = phx_fma(dt, sn, n0[k][e]);
            }
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
            if (store) {
                err_acc += TETM ? 2.0 * ee : ee;
                nrm_acc += (double)out0r * (double)out0r + (double)out0i * (double)out0i;
            }
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
        const double es = wave_sum(err_acc), ns = wave_sum(nrm_acc);
        if (lane == 0) {
            const int gw = band * a.nstrips + strip;
            a.partials[2 * gw] = es;
            a.partials[2 * gw + 1] = ns;
        }
    }
}

/* deterministic reduction of the per-wave partials: one workgroup, fixed order */
__global__ __launch_bounds__(256) void reduce_partials(const double* __restrict__ partials, int nwaves, double* __restrict__ out) {
    __shared__ double se[256], sn[256];
    double e = 0, n = 0;
    for (int i = threadIdx.x; i < nwaves; i += 256) {
        e += partials[2 * i];
        n += partials[2 * i + 1];
    }
    se[threadIdx.x] = e;
    sn[threadIdx.x] = n;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sn[threadIdx.x] += sn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = se[0];
        out[1] = sn[0];
    }
}

/* ---------------------------------------------------------------- launch */
template <typename real, typename TAB, int TY, bool TETM, bool RES, bool PUMP, bool POT, bool STOCH>
static hipError_t launch_gen(const RkGenArgs<real>& a0, hipStream_t s) {
    RkGenArgs<real> a = a0;
    const int outw = kWave - 2 * TAB::S;
    a.nstrips = (a.Nc + outw - 1) / outw;
    a.nbands = (a.Nr + TY - 1) / TY;
    const int bandgroups = (a.nbands + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((rk_generic_fused<real, TAB, TY, TETM, RES, PUMP, POT, STOCH>), dim3(a.nstrips * bandgroups),
                       dim3(kWave * kWavesPerBlock), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !TAB::ERR) return e;
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, s, a.partials, a.nstrips * a.nbands, a.sums);
    return hipGetLastError();
}

template <typename real, typename TAB, int TY>
static hipError_t dispatch_flags(const RkGenArgs<real>& a, hipStream_t s) {
    const bool tetm = a.tetm, res = a.n_in[0] != nullptr, pump = res && a.pump[0] != nullptr, pot = a.pot[0] != nullptr,
               st = a.stochastic;
    /* instantiated combinations: (tetm, res) x pump x pot x stoch with pump => res */
#define PHX_GEN(T, R, P, V, N)                                                                         \
    if (tetm == T && res == R && pump == P && pot == V && st == N) return launch_gen<real, TAB, TY, T, R, P, V, N>(a, s);
    PHX_GEN(false, false, false, false, false)
    PHX_GEN(false, false, false, true, false)
    PHX_GEN(false, true, false, false, false)
    PHX_GEN(false, true, true, false, false)
    PHX_GEN(false, true, true, true, false)
    PHX_GEN(false, true, false, true, false)
    PHX_GEN(false, true, true, false, true)
    PHX_GEN(false, false, false, false, true)
    PHX_GEN(true, true, true, false, false)
    PHX_GEN(true, true, true, true, false)
    PHX_GEN(true, true, false, false, false)
    PHX_GEN(true, true, true, false, true)
#undef PHX_GEN
    return hipErrorNotSupported;
}

template <typename real>
hipError_t launch_rk_generic(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s) {
    switch (tab) {
        case RkTableau::RK4: return dispatch_flags<real, TabRK4, 16>(a, s);
        case RkTableau::DP45_FIXED: return dispatch_flags<real, TabDP45Fixed, 16>(a, s);
        case RkTableau::DP45_ADAPTIVE: return dispatch_flags<real, TabDP45Adaptive, 16>(a, s);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_rk_generic<float>(const RkGenArgs<float>&, RkTableau, hipStream_t);
template hipError_t launch_rk_generic<double>(const RkGenArgs<double>&, RkTableau, hipStream_t);

} // namespace phx
