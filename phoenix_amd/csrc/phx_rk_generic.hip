/*
 * phx_rk_generic.hip -- instantiations of the fused Runge-Kutta kernel
 * (phx_rk_generic.h) for RK4 (with noise or TE-TM; the plain scalar RK4 has its own
 * hand-tuned kernel in phx_rk4.h) and Dormand-Prince 5(4) fixed and adaptive, the
 * deterministic reduction of the adaptive error partials, the per-step noise field, the
 * imaginary-time normalisation, and the dispatch over all tableaux (the other groups are instantiated in phx_rk_tab3.hip,
 * phx_rk_tab4.hip, phx_rk_tab6.hip and phx_rk_tab6f.hip).
 */
#include <algorithm>
#include <cstdlib>

#include "phx_rk_generic.h"

namespace phx {

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

hipError_t launch_reduce_partials(const double* partials, int nwaves, double* sums, hipStream_t s) {
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, s, partials, nwaves, sums);
    return hipGetLastError();
}

/* per-256-cell FP64 sums of |psi+|^2 and |psi-|^2, fixed order inside the block */
template <typename real>
__global__ __launch_bounds__(256) void sumsq_blocks(const typename Vec2<real>::type* __restrict__ a,
                                                    const typename Vec2<real>::type* __restrict__ b, size_t cells,
                                                    double* __restrict__ partials) {
    __shared__ double sa[256], sb[256];
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    double x = 0.0, y = 0.0;
    if (i < cells) {
        const auto u = a[i];
        x = (double)u.x * (double)u.x + (double)u.y * (double)u.y;
        if (b) {
            const auto v = b[i];
            y = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        }
    }
    sa[threadIdx.x] = x;
    sb[threadIdx.x] = y;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sa[0];
        partials[2 * blockIdx.x + 1] = sb[0];
    }
}

/* psi_k *= sqrt(amp / (S_k dV)), S_k < 1e-10 -> 1 (solver_normalize_imag_time.cu:13-45) */
template <typename real>
__global__ __launch_bounds__(256) void imag_scale(typename Vec2<real>::type* __restrict__ a, typename Vec2<real>::type* __restrict__ b,
                                                  size_t cells, const double* __restrict__ sums, double amp, double dV) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= cells) return;
    const double s0 = sums[0] < 1e-10 ? 1.0 : sums[0];
    const real f0 = (real)sqrt(amp / (s0 * dV));
    auto u = a[i];
    u.x *= f0;
    u.y *= f0;
    a[i] = u;
    if (b) {
        const double s1 = sums[1] < 1e-10 ? 1.0 : sums[1];
        const real f1 = (real)sqrt(amp / (s1 * dV));
        auto v = b[i];
        v.x *= f1;
        v.y *= f1;
        b[i] = v;
    }
}

template <typename real>
hipError_t launch_imag_normalize(typename Vec2<real>::type* psi_plus, typename Vec2<real>::type* psi_minus, size_t cells,
                                 double amp, double dV, double* partials, double* sums, hipStream_t s) {
    const size_t nb = imag_norm_blocks(cells);
    hipLaunchKernelGGL(sumsq_blocks<real>, dim3((unsigned)nb), dim3(256), 0, s, psi_plus, psi_minus, cells, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_reduce_partials(partials, (int)nb, sums, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(imag_scale<real>, dim3((unsigned)nb), dim3(256), 0, s, psi_plus, psi_minus, cells, sums, amp, dV);
    return hipGetLastError();
}

template hipError_t launch_imag_normalize<float>(float2*, float2*, size_t, double, double, double*, double*, hipStream_t);
template hipError_t launch_imag_normalize<double>(double2*, double2*, size_t, double, double, double*, double*, hipStream_t);

/* one complex normal pair per cell, scaled by amp (own Philox stream, see phx_noise.h) */
template <typename real>
__global__ __launch_bounds__(256) void noise_field(typename Vec2<real>::type* __restrict__ xi, uint32_t cells, uint64_t seed,
                                                   uint64_t step, real amp) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= cells) return;
    real zr, zi;
    phx_noise_normal2<real>(seed, step, i, 0, zr, zi);
    typename Vec2<real>::type z;
    z.x = zr * amp;
    z.y = zi * amp;
    xi[i] = z;
}

template <typename real>
hipError_t launch_noise_field(typename Vec2<real>::type* xi, int Nc, int Nr, uint64_t seed, uint64_t step, real amp,
                              hipStream_t s) {
    const size_t cells = (size_t)Nc * (size_t)Nr;
    if (cells * sizeof(typename Vec2<real>::type) > kMaxArrayBytes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(noise_field<real>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, xi, (uint32_t)cells, seed,
                       step, amp);
    return hipGetLastError();
}

template hipError_t launch_noise_field<float>(float2*, int, int, uint64_t, uint64_t, float, hipStream_t);
template hipError_t launch_noise_field<double>(double2*, int, int, uint64_t, uint64_t, double, hipStream_t);

int dp45_adaptive_tile_rows(bool tetm, bool fp32, int Nc, int Nr) {
    if (!(tetm && fp32)) return kRkGenMinTileRows; /* the tall tiles exist for the FP32 spinor kernel */
    const char* e = std::getenv("PHX_GEN_TY"); /* read at every launch: tests switch it within a process */
    const int env_ty = e ? std::atoi(e) : 0;
    if (env_ty == 16 || env_ty == 32) return env_ty;
    static const int simds = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        return 4 * std::max(cus, 1);
    }();
    /* One wave per SIMD at the kernel's register count (over 256 VGPRs + AGPRs at every height),
       and about the same VALU work per row evaluation at every height (145-148 instructions): a
       SIMD's time is its number of tiles (rounded up) times the row evaluations of one tile,
       S (TY + 2 S) - S (S + 1) for S = 7.  At 2048^2 this predicts 32 rows 1.16x faster than 16
       (1.21x measured).  Other heights that the model ranks near or above 32 measured slower and
       are not built: 42 rows (2009 tiles, 2 rounds) 1.24x slower, 28 rows (3034 tiles, 3 full
       rounds of 238 row evaluations against 3 x 266) 1.19x slower. */
    constexpr int S = 7;
    int best = kRkGenMinTileRows;
    long best_cost = -1;
    for (int ty : {16, 32}) {
        const long waves = (long)((Nc + 63 - 2 * S) / (64 - 2 * S)) * ((Nr + ty - 1) / ty);
        const long cost = ((waves + simds - 1) / simds) * (long)(S * (ty + 2 * S) - S * (S + 1));
        if (best_cost < 0 || cost < best_cost) best = ty, best_cost = cost;
    }
    return best;
}

template <typename real>
hipError_t launch_rk_generic(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s) {
    switch (tab) {
        case RkTableau::RK4: return launch_tab<real, TabRK4>(a, s);
        case RkTableau::DP45_FIXED: return launch_tab<real, TabDP45Fixed>(a, s);
        case RkTableau::DP45_ADAPTIVE: {
            const int ty = dp45_adaptive_tile_rows(a.tetm, sizeof(real) == 4, a.Nc, a.Nr);
            if constexpr (sizeof(real) == 4)
                if (ty != kRkGenMinTileRows && a.tetm) return launch_dp45_tetm_tall(a, ty, s);
            return launch_tab<real, TabDP45Adaptive>(a, s);
        }
        case RkTableau::NEWTON:
        case RkTableau::MP:
        case RkTableau::HEUN2:
        case RkTableau::RALSTON:
        case RkTableau::HEUN3:
        case RkTableau::RALSTON3:
        case RkTableau::RK3:
        case RkTableau::SSPRK3:
        case RkTableau::HW: return launch_rk_tab3<real>(a, tab, s);
        case RkTableau::FEHLBERG2:
        case RkTableau::FEHLBERG2_ADAPTIVE:
        case RkTableau::RALSTON4:
        case RkTableau::RULE38:
        case RkTableau::BOSHA:
        case RkTableau::BOSHA_ADAPTIVE: return launch_rk_tab4<real>(a, tab, s);
        case RkTableau::NYSTROEM:
        case RkTableau::CASHKARP:
        case RkTableau::CASHKARP_ADAPTIVE: return launch_rk_tab6<real>(a, tab, s);
        case RkTableau::FEHLBERG5:
        case RkTableau::FEHLBERG5_ADAPTIVE: return launch_rk_tab6f<real>(a, tab, s);
        /* more stages than the fused kernel keeps in registers: staged path (launch_staged) */
        case RkTableau::NSRK78:
        case RkTableau::NSRK78_ADAPTIVE:
        case RkTableau::DP85:
        case RkTableau::DP85_ADAPTIVE: return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_rk_generic<float>(const RkGenArgs<float>&, RkTableau, hipStream_t);
template hipError_t launch_rk_generic<double>(const RkGenArgs<double>&, RkTableau, hipStream_t);

} // namespace phx
