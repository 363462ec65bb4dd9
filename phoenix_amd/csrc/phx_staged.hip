/*
 * phx_staged.hip -- stage-by-stage explicit Runge-Kutta step with the K's in HBM.
 *
 * The fused kernels (phx_rk4.h, phx_rk_generic.h) keep every stage in registers, which
 * bounds them to at most seven stages and to one temporal group per source.  This path
 * serves everything else of the reference's step(), one launch per macro of its
 * sequence (cuda_macro.cuh:110-250):
 *   - CALCULATE_K (staged_k): K_s = f(Y_s) on the extent halo - s;
 *   - INTERMEDIATE_SUM_K (staged_sum): Y_{s+1} = X + dt sum_j a[s+1][j] K_j (+ noise);
 *   - FINAL_SUM_K / ERROR_K (staged_final): X' = X + dt sum_j b_j K_j (+ noise) on the grid,
 *     and the FP64 per-block sums of |dt sum_j e_j K_j(psi+)|^2 and |psi+'|^2.
 * It covers any tableau (the 12- and 13-stage DP85 and NSRK78 included), pump, potential
 * and pulse chains with any number of temporal groups (kernel_gp_compute.cuh:68-100),
 * complex temporal envelopes on the potential and the pulse, TE-TM, noise, imaginary
 * time and the complex reservoir (n as {re, im} cells, which imaginary time with a reservoir
 * and complex temporal pump amplitudes need).
 *
 * Layout: the step-start state X and the output are the solver's full-grid arrays; the
 * stage inputs Y and the K slots live on a grid padded by H = number of stages on every
 * side ((N_r + 2H) x (N_c + 2H), row pitch P = N_c + 2H), the reference's halo layout with
 * one subgrid.  X, the sources and the noise are read through the boundary map (periodic
 * wrap, or 0 outside a zero-boundary domain), so with a zero boundary the ring outside the
 * grid evolves within the step exactly as the reference's halo does.
 * Traffic per stage and cell (FP32 scalar with reservoir): Y read 12 B (+ L2-served
 * neighbours), K write 12 B, sum: (s + 1) K reads + Y write -- HBM-bound, O(S^2) in the
 * stage count; it is not on the timed path of the benchmark workloads.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_device.h"
#include "phx_kernels.h"
#include "phx_noise.h"

namespace phx {

namespace {

constexpr int kBx = 64, kBy = 4; /* 256 threads: one wavefront per padded row segment */

template <typename real>
struct Map {
    int Nc, Nr, H, px, py;
    /* full-grid index of padded cell (pr, pc), or -1 outside a zero-boundary domain */
    __device__ __forceinline__ long idx(int pr, int pc) const {
        int r = pr - H, c = pc - H;
        if (r < 0 || r >= Nr) {
            if (!py) return -1;
            r = ((r % Nr) + Nr) % Nr;
        }
        if (c < 0 || c >= Nc) {
            if (!px) return -1;
            c = ((c % Nc) + Nc) % Nc;
        }
        return (long)r * Nc + c;
    }
};

template <typename real>
__device__ __forceinline__ typename Vec2<real>::type cz() {
    typename Vec2<real>::type z;
    z.x = real(0);
    z.y = real(0);
    return z;
}

template <typename real>
__device__ __forceinline__ Map<real> map_of(const StagedArgs<real>& a) {
    return Map<real>{a.Nc, a.Nr, a.H, a.periodic_x, a.periodic_y};
}

/* noise amplitude sqrt((R n+ + gamma_c) / (4 dV)) with the step-start reservoir+ (kernel_summation.cuh:209-225) */
template <typename real, bool RES>
__device__ __forceinline__ real noise_sigma(const StagedArgs<real>& a, long gi) {
    const real n = (RES && gi >= 0) ? a.x_n[0][gi] : real(0);
    return phx_sqrt((RES ? phx_fma(a.ph.R, n, a.ph.gamma_c) : a.ph.gamma_c) / (real(4) * a.ph.dV));
}

/* the noise increment w_sum xi s of a stage sum, s = sqrt((R n+ + gamma_c) / (4 dV)); with a complex
   reservoir s is the principal complex square root and the product complex */
template <typename real, bool RES, bool CRES>
__device__ __forceinline__ void noise_inc(const StagedArgs<real>& a, long gi, real& nzr, real& nzi) {
    using v2 = typename Vec2<real>::type;
    const real xr = a.wsum * a.noise[gi].x, xi = a.wsum * a.noise[gi].y;
    if constexpr (CRES) {
        const v2 n = reinterpret_cast<const v2*>(a.x_n[0])[gi];
        const real d = real(4) * a.ph.dV;
        v2 sq;
        csqrt_principal<real>(phx_fma(a.ph.R, n.x, a.ph.gamma_c) / d, (a.ph.R * n.y) / d, sq.x, sq.y);
        nzr = xr * sq.x - xi * sq.y;
        nzi = xr * sq.y + xi * sq.x;
    } else {
        const real sg = noise_sigma<real, RES>(a, gi);
        nzr = xr * sg;
        nzi = xi * sg;
    }
}

/* a complex stage sum of the reservoir: x + (dtr + i dti) (sr + i si), as for psi */
template <typename real>
__device__ __forceinline__ typename Vec2<real>::type cres_advance(const StagedArgs<real>& a, typename Vec2<real>::type x, real sr,
                                                                  real si) {
    typename Vec2<real>::type y;
    y.x = phx_fma(a.dtr, sr, phx_fma(-a.dti, si, x.x));
    y.y = phx_fma(a.dtr, si, phx_fma(a.dti, sr, x.y));
    return y;
}

} // namespace

/* ------------------------------------------------------------------ CALCULATE_K */
template <typename real, bool TETM, bool RES, bool CRES>
__global__ __launch_bounds__(kBx * kBy) void staged_k(StagedArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    constexpr int NC = TETM ? 2 : 1;
    const int lo = a.H - a.ext;
    const int pc = lo + (int)(blockIdx.x * kBx + threadIdx.x);
    const int pr = lo + (int)(blockIdx.y * kBy + threadIdx.y);
    if (pc >= a.H + a.Nc + a.ext || pr >= a.H + a.Nr + a.ext) return;
    const Map<real> m = map_of(a);
    const size_t P = (size_t)a.P;
    const bool first = a.stage == 0; /* stage 1 reads the step-start state through the boundary map */
    auto psi = [&](int k, int r, int c) -> v2 {
        if (first) {
            const long g = m.idx(r, c);
            return g >= 0 ? a.x_psi[k][g] : cz<real>();
        }
        return a.y_psi[k][(size_t)r * P + c];
    };
    /* the reservoir at a cell: (Re n, Im n); Im n = 0 for a real reservoir */
    auto res = [&](int k, int r, int c) -> v2 {
        v2 z = cz<real>();
        if (!RES) return z;
        if constexpr (CRES) {
            if (first) {
                const long g = m.idx(r, c);
                return g >= 0 ? reinterpret_cast<const v2*>(a.x_n[k])[g] : z;
            }
            return reinterpret_cast<const v2*>(a.y_n[k])[(size_t)r * P + c];
        }
        if (first) {
            const long g = m.idx(r, c);
            z.x = g >= 0 ? a.x_n[k][g] : real(0);
            return z;
        }
        z.x = a.y_n[k][(size_t)r * P + c];
        return z;
    };
    const long gi = m.idx(pr, pc); /* sources and noise: 0 outside a zero-boundary domain */
    const PhxPhys<real>& ph = a.ph;
    v2 cen[NC];
    real ewr[NC], ewi[NC], nsr[NC], nsi[NC], nrm[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        cen[k] = psi(k, pr, pc);
        const v2 e = psi(k, pr, pc + 1), w = psi(k, pr, pc - 1), n = psi(k, pr - 1, pc), s = psi(k, pr + 1, pc);
        ewr[k] = e.x + w.x;
        ewi[k] = e.y + w.y;
        nsr[k] = n.x + s.x;
        nsi[k] = n.y + s.y;
        nrm[k] = phx_fma(cen[k].x, cen[k].x, cen[k].y * cen[k].y);
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const v2 nc = res(k, pr, pc);
        const real nv = nc.x;
        /* sources: sum over temporal groups g of field_g(cell) * a_g(w) */
        real pmp = 0, pmi = 0, vre = 0, vim = 0, fre = 0, fim = 0;
        if (gi >= 0) {
            for (int g = 0; g < a.npump; g++) {
                const real pv = a.pump[k][(size_t)g * a.N + gi];
                pmp = phx_fma(pv, a.pump_amp[g], pmp);
                if (CRES) pmi = phx_fma(pv, a.pump_amp_im[g], pmi);
            }
            for (int g = 0; g < a.npot; g++) {
                const real v = a.pot[k][(size_t)g * a.N + gi];
                vre = phx_fma(v, a.pot_amp[g].x, vre);
                vim = phx_fma(v, a.pot_amp[g].y, vim);
            }
            for (int g = 0; g < a.npulse; g++) {
                const v2 f = a.pulse[k][(size_t)g * a.N + gi];
                const v2 am = a.pulse_amp[g];
                fre += f.x * am.x - f.y * am.y;
                fim += f.x * am.y + f.y * am.x;
            }
        }
        const real poth = vre * ph.one_over_h_bar_s;
        real kr, ki, kn;
        if constexpr (TETM) {
            /* cross term of the other component (kernel_gp_compute.cuh:123-130):
               (E+W)_o/dx^2 - (N+S)_o/dy^2 +- (i/2)/(dx dy) (SE + NW - SW - NE)_o */
            const int o = 1 - k;
            const v2 se = psi(o, pr + 1, pc + 1), nw = psi(o, pr - 1, pc - 1), sw = psi(o, pr + 1, pc - 1),
                     ne = psi(o, pr - 1, pc + 1);
            const real Dr = (se.x - ne.x) - (sw.x - nw.x), Di = (se.y - ne.y) - (sw.y - nw.y);
            const real sgn = k == 0 ? real(1) : real(-1);
            const real hx = ph.lt_h * ph.one_over_dx2, hy = ph.lt_h * ph.one_over_dy2;
            const real ccr = phx_fma(hx, ewr[o], phx_fma(-hy, nsr[o], -sgn * ph.lt_hxy * Di));
            const real cci = phx_fma(hx, ewi[o], phx_fma(-hy, nsi[o], sgn * ph.lt_hxy * Dr));
            const real tr = phx_fma(ph.kin_x, ewr[k], phx_fma(ph.kin_y, nsr[k], ph.kin_c * cen[k].x));
            const real ti = phx_fma(ph.kin_x, ewi[k], phx_fma(ph.kin_y, nsi[k], ph.kin_c * cen[k].y));
            phx_rhs_tetm<real>(ph, cen[k].x, cen[k].y, tr, ti, ccr, cci, nrm[o], nv, pmp, poth, kr, ki, kn);
        } else {
            const real tr = phx_fma(ph.kin_x, ewr[k], ph.kin_y * nsr[k]);
            const real ti = phx_fma(ph.kin_x, ewi[k], ph.kin_y * nsi[k]);
            phx_rhs_scalar<real>(ph, cen[k].x, cen[k].y, tr, ti, nv, pmp, poth, kr, ki, kn);
        }
        if constexpr (CRES) {
            /* Im n terms of (1/hbar) g_r n (-i psi) + 1/2 R n psi and of k_n (kernel_gp_compute.cuh:66-82):
               the hooks above took n = Re n */
            const real ni = nc.y;
            const real hr = real(0.5) * ph.R;
            kr = phx_fma(ph.gr_h * ni, cen[k].x, phx_fma(-(hr * ni), cen[k].y, kr));
            ki = phx_fma(ph.gr_h * ni, cen[k].y, phx_fma(hr * ni, cen[k].x, ki));
        }
        /* imaginary part of a complex potential: (1/hbar)(i V_im)(-i psi) = V_im/hbar psi */
        const real vih = vim * ph.one_over_h_bar_s;
        kr = phx_fma(vih, cen[k].x, kr);
        ki = phx_fma(vih, cen[k].y, ki);
        /* pulse: + (1/hbar) sum_g F_g a_g(w) (kernel_gp_compute.cuh:94-100) */
        kr = phx_fma(ph.one_over_h_bar_s, fre, kr);
        ki = phx_fma(ph.one_over_h_bar_s, fim, ki);
        if (a.stochastic) {
            /* kernel_gp_compute.cuh:83-85,106-108: k_psi -= g_c/hbar (-i psi)/dV ; k_n += R n / dV */
            const real g = ph.gc_h / ph.dV;
            kr -= g * cen[k].y;
            ki += g * cen[k].x;
            if (RES) kn = phx_fma(ph.R / ph.dV, nv, kn);
        }
        const size_t o = (size_t)a.stage * a.padN + (size_t)pr * P + pc;
        v2 kv;
        kv.x = kr;
        kv.y = ki;
        a.k_psi[k][o] = kv;
        if constexpr (CRES) {
            /* Im k_n = -(gamma_r + R |psi|^2) Im n + sum_g P_g Im a_g (+ R Im n / dV with noise) */
            real kni = phx_fma(phx_fma(-ph.R, nrm[k], -ph.gamma_r), nc.y, pmi);
            if (a.stochastic) kni = phx_fma(ph.R / ph.dV, nc.y, kni);
            v2 knv;
            knv.x = kn;
            knv.y = kni;
            reinterpret_cast<v2*>(a.k_n[k])[o] = knv;
        } else if (RES) {
            a.k_n[k][o] = kn;
        }
    }
}

/* ----------------------------------------------------------- INTERMEDIATE_SUM_K */
template <typename real, bool TETM, bool RES, bool CRES>
__global__ __launch_bounds__(kBx * kBy) void staged_sum(StagedArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    constexpr int NC = TETM ? 2 : 1;
    const int lo = a.H - a.ext;
    const int pc = lo + (int)(blockIdx.x * kBx + threadIdx.x);
    const int pr = lo + (int)(blockIdx.y * kBy + threadIdx.y);
    if (pc >= a.H + a.Nc + a.ext || pr >= a.H + a.Nr + a.ext) return;
    const Map<real> m = map_of(a);
    const long gi = m.idx(pr, pc);
    const size_t pi = (size_t)pr * a.P + pc;
    real nzr = 0, nzi = 0;
    if (a.stochastic && gi >= 0) noise_inc<real, RES, CRES>(a, gi, nzr, nzi);
#pragma unroll
    for (int k = 0; k < NC; k++) {
        real sr = 0, si = 0, sn = 0, sni = 0;
        for (int j = 0; j < a.nw; j++) {
            const size_t o = (size_t)a.slot[j] * a.padN + pi;
            const v2 kv = a.k_psi[k][o];
            sr = phx_fma(a.w[j], kv.x, sr);
            si = phx_fma(a.w[j], kv.y, si);
            if constexpr (CRES) {
                const v2 kn = reinterpret_cast<const v2*>(a.k_n[k])[o];
                sn = phx_fma(a.w[j], kn.x, sn);
                sni = phx_fma(a.w[j], kn.y, sni);
            } else if (RES) {
                sn = phx_fma(a.w[j], a.k_n[k][o], sn);
            }
        }
        const v2 x = gi >= 0 ? a.x_psi[k][gi] : cz<real>();
        v2 y;
        y.x = phx_fma(a.dtr, sr, phx_fma(-a.dti, si, x.x)) + nzr;
        y.y = phx_fma(a.dtr, si, phx_fma(a.dti, sr, x.y)) + nzi;
        a.y_psi[k][pi] = y;
        if constexpr (CRES) {
            const v2 xn = gi >= 0 ? reinterpret_cast<const v2*>(a.x_n[k])[gi] : cz<real>();
            reinterpret_cast<v2*>(a.y_n[k])[pi] = cres_advance<real>(a, xn, sn, sni);
        } else if (RES) {
            a.y_n[k][pi] = phx_fma(a.dt, sn, gi >= 0 ? a.x_n[k][gi] : real(0));
        }
    }
}

/* --------------------------------------------------------- FINAL_SUM_K / ERROR_K */
template <typename real, bool TETM, bool RES, bool CRES>
__global__ __launch_bounds__(kBx * kBy) void staged_final(StagedArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    constexpr int NC = TETM ? 2 : 1;
    const int c = (int)(blockIdx.x * kBx + threadIdx.x);
    const int r = (int)(blockIdx.y * kBy + threadIdx.y);
    const bool inside = c < a.Nc && r < a.Nr;
    double ee = 0.0, nn = 0.0;
    if (inside) {
        const long gi = (long)r * a.Nc + c;
        const size_t pi = (size_t)(r + a.H) * a.P + (c + a.H);
        real nzr = 0, nzi = 0;
        if (a.stochastic) noise_inc<real, RES, CRES>(a, gi, nzr, nzi);
#pragma unroll
        for (int k = 0; k < NC; k++) {
            real sr = 0, si = 0, sn = 0, sni = 0;
            for (int j = 0; j < a.nw; j++) {
                const size_t o = (size_t)a.slot[j] * a.padN + pi;
                const v2 kv = a.k_psi[k][o];
                sr = phx_fma(a.w[j], kv.x, sr);
                si = phx_fma(a.w[j], kv.y, si);
                if constexpr (CRES) {
                    const v2 kn = reinterpret_cast<const v2*>(a.k_n[k])[o];
                    sn = phx_fma(a.w[j], kn.x, sn);
                    sni = phx_fma(a.w[j], kn.y, sni);
                } else if (RES) {
                    sn = phx_fma(a.w[j], a.k_n[k][o], sn);
                }
            }
            const v2 x = a.x_psi[k][gi];
            v2 y;
            y.x = phx_fma(a.dtr, sr, phx_fma(-a.dti, si, x.x)) + nzr;
            y.y = phx_fma(a.dtr, si, phx_fma(a.dti, sr, x.y)) + nzi;
            a.out_psi[k][gi] = y;
            if constexpr (CRES) {
                reinterpret_cast<v2*>(a.out_n[k])[gi] = cres_advance<real>(a, reinterpret_cast<const v2*>(a.x_n[k])[gi], sn, sni);
            } else if (RES) {
                a.out_n[k][gi] = phx_fma(a.dt, sn, a.x_n[k][gi]);
            }
            if (k == 0 && a.ne > 0) {
                /* runge_sum_to_error (kernel_summation.cuh:270-289); with -tetm the psi+
                   error counts twice (cuda_macro.cuh:244) */
                real er = 0, ei = 0;
                for (int j = 0; j < a.ne; j++) {
                    const v2 kv = a.k_psi[0][(size_t)a.eslot[j] * a.padN + pi];
                    er = phx_fma(a.e[j], kv.x, er);
                    ei = phx_fma(a.e[j], kv.y, ei);
                }
                const double dr = (double)(a.dt * er), di = (double)(a.dt * ei);
                ee = (TETM ? 2.0 : 1.0) * (dr * dr + di * di);
                nn = (double)y.x * (double)y.x + (double)y.y * (double)y.y;
            }
        }
    }
    if (a.ne > 0) {
        /* per-block FP64 partial sums in a fixed order: wavefront shuffles, then LDS */
        __shared__ double red[2][kBy];
        for (int o = 32; o > 0; o >>= 1) {
            ee += __shfl_xor(ee, o, 64);
            nn += __shfl_xor(nn, o, 64);
        }
        if (threadIdx.x == 0) {
            red[0][threadIdx.y] = ee;
            red[1][threadIdx.y] = nn;
        }
        __syncthreads();
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            double se = 0.0, sn = 0.0;
            for (int w = 0; w < kBy; w++) {
                se += red[0][w];
                sn += red[1][w];
            }
            const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
            a.partials[2 * b] = se;
            a.partials[2 * b + 1] = sn;
        }
    }
}

/* ---------------------------------------------------------------- launchers */
namespace {

template <typename real>
dim3 ext_grid(const StagedArgs<real>& a) {
    const int w = a.Nc + 2 * a.ext, h = a.Nr + 2 * a.ext;
    return dim3((unsigned)((w + kBx - 1) / kBx), (unsigned)((h + kBy - 1) / kBy));
}

template <typename real, bool TETM, bool RES, bool CRES>
hipError_t launch_kind(const StagedArgs<real>& a, StagedOp op, hipStream_t s) {
    const dim3 blk(kBx, kBy);
    switch (op) {
        case StagedOp::K: hipLaunchKernelGGL((staged_k<real, TETM, RES, CRES>), ext_grid(a), blk, 0, s, a); break;
        case StagedOp::SUM: hipLaunchKernelGGL((staged_sum<real, TETM, RES, CRES>), ext_grid(a), blk, 0, s, a); break;
        case StagedOp::FINAL: {
            const dim3 g((unsigned)((a.Nc + kBx - 1) / kBx), (unsigned)((a.Nr + kBy - 1) / kBy));
            hipLaunchKernelGGL((staged_final<real, TETM, RES, CRES>), g, blk, 0, s, a);
            break;
        }
    }
    return hipGetLastError();
}

} // namespace

template <typename real>
hipError_t launch_staged(const StagedArgs<real>& a, StagedOp op, hipStream_t s) {
    if (a.ext < 0 || a.ext > a.H || a.nw > kStagedMaxStages || a.ne > kStagedMaxStages || a.npump > kStagedMaxGroups ||
        a.npot > kStagedMaxGroups || a.npulse > kStagedMaxGroups)
        return hipErrorInvalidValue;
    const bool res = a.x_n[0] != nullptr;
    if (a.cres && !res) return hipErrorInvalidValue;
    if (a.cres) return a.tetm ? launch_kind<real, true, true, true>(a, op, s) : launch_kind<real, false, true, true>(a, op, s);
    if (a.tetm) return res ? launch_kind<real, true, true, false>(a, op, s) : launch_kind<real, true, false, false>(a, op, s);
    return res ? launch_kind<real, false, true, false>(a, op, s) : launch_kind<real, false, false, false>(a, op, s);
}

template hipError_t launch_staged<float>(const StagedArgs<float>&, StagedOp, hipStream_t);
template hipError_t launch_staged<double>(const StagedArgs<double>&, StagedOp, hipStream_t);

} // namespace phx
