/*
 * phx_rk4.hip -- fused RK4 step of the scalar GP + reservoir model for gfx950.
 *
 * One launch = one full RungeKutta4::step (runge_kutta.cu:74-94): halo sync,
 * K1..K4 on shrinking extents, three intermediate sums and the final sum,
 * which the reference runs as 2 + 12*S separate kernels over global memory
 * (cuda_macro.cuh:306-353).  Here every intermediate value stays in registers.
 *
 * Mapping (one wavefront = one tile, no inter-wave synchronisation):
 *   - a wave owns a strip of 64 extended columns (56 output columns plus a
 *     4-cell halo either side), one lane per column;
 *   - it streams TY+8 extended rows top to bottom; after loading row i it
 *     evaluates stage 1 at row i-1, stage 2 at i-2, stage 3 at i-3 and stage 4
 *     at i-4 and writes output row i-4.  Stage s is therefore evaluated on rows
 *     [s, TY+7-s] and (implicitly, through the lanes that later read it) on
 *     columns [s, 63-s]: the extents of GENERATE_SUBGRID_INDEX with
 *     current_halo = 4 - s (cuda_macro.cuh:43-49,112-113);
 *   - north/south neighbours come from the register window of the previous
 *     stage, east/west neighbours from the adjacent lanes (cross-lane shuffle);
 *   - out-of-domain cells of a zero boundary are loaded as 0 (psi, n, P, V),
 *     which is what synchronize_halos writes into the reference's halos
 *     (kernel_halo.cuh:66-69); periodic boundaries wrap the load index.
 * HBM traffic per step and cell: read psi, n, P (, V) once, write psi, n once
 * (plus the 8-row / 8-column halo re-read, served from L2).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_kernels.h"

namespace phx {

constexpr int kWave = 64;
constexpr int kHalo = 4;
constexpr int kOutW = kWave - 2 * kHalo; /* 56 */
constexpr int kWavesPerBlock = 4;

/* Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
   Remap so that each XCD receives a contiguous range of tiles: horizontally
   adjacent tiles then share one L2 for their halo columns. */
__device__ __forceinline__ int xcd_contiguous(int b, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int x = b & 7;
    const int start = x * q + (x < r ? x : r);
    return start + (b >> 3);
}

template <typename real>
__device__ __forceinline__ real lane_from(real v, int src) {
    return __shfl(v, src, kWave);
}

template <typename real, int TY, bool RES, bool PUMP, bool POT>
__global__ __launch_bounds__(256) void rk4_scalar_fused(Rk4Args<real> a) {
    using v2 = typename Vec2<real>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const int strip = tile % a.nstrips;
    const int band = (tile / a.nstrips) * kWavesPerBlock + wave;
    if (band >= a.nbands) return;
    const int x0 = strip * kOutW;
    const int y0 = band * TY;
    const PhxPhys<real>& ph = a.ph;

    int gcol = x0 - kHalo + lane;
    bool col_in = true;
    if (gcol < 0 || gcol >= a.Nc) {
        if (a.periodic_x)
            gcol = ((gcol % a.Nc) + a.Nc) % a.Nc;
        else {
            col_in = false;
            gcol = 0;
        }
    }
    const bool out_lane = lane >= kHalo && lane < kWave - kHalo && (x0 - kHalo + lane) < a.Nc;

    constexpr int E = TY + 2 * kHalo;
    real p0r[E], p0i[E], n0[E], pp[E], vv[E];
    real y1r[E], y1i[E], y2r[E], y2i[E], y3r[E], y3i[E];
    real yn1[E], yn2[E], yn3[E];
    real acr[E], aci[E], acn[E];

    const real dth = a.dt * real(0.5);
    const real dt1 = a.dt;
    const real w16 = real(1.0 / 6.0), w13 = real(1.0 / 3.0);

    /* stage evaluation: centre (cr,ci), north+south sum, reservoir value */
    auto rhs = [&](real cr, real ci, real nsr, real nsi, real nv, real pump, real pot, real& kr, real& ki, real& kn) {
        const real er = lane_from(cr, lane + 1), ei = lane_from(ci, lane + 1);
        const real wr = lane_from(cr, lane - 1), wi = lane_from(ci, lane - 1);
        const real lr = ph.m2_over_dx2_p_dy2 * cr + nsr * ph.one_over_dy2 + (er + wr) * ph.one_over_dx2;
        const real li = ph.m2_over_dx2_p_dy2 * ci + nsi * ph.one_over_dy2 + (ei + wi) * ph.one_over_dx2;
        phx_rhs_scalar<real>(ph, cr, ci, lr, li, RES ? nv : real(0), PUMP ? pump : real(0), POT ? pot : real(0), kr, ki, kn);
    };

#pragma unroll
    for (int it = 0; it < E; ++it) {
        /* ---- load extended row `it` */
        {
            int grow = y0 - kHalo + it;
            bool in = col_in;
            if (grow < 0 || grow >= a.Nr) {
                if (a.periodic_y)
                    grow = ((grow % a.Nr) + a.Nr) % a.Nr;
                else
                    in = false;
            }
            real lr = 0, li = 0, ln = 0, lp = 0, lv = 0;
            if (in) {
                const size_t idx = (size_t)grow * (size_t)a.Nc + (size_t)gcol;
                const v2 v = a.psi_in[idx];
                lr = v.x;
                li = v.y;
                if (RES) ln = a.n_in[idx];
                if (PUMP) lp = a.pump[idx];
                if (POT) lv = a.pot[idx];
            }
            p0r[it] = lr;
            p0i[it] = li;
            n0[it] = ln;
            pp[it] = lp;
            vv[it] = lv;
        }
        /* ---- stage 1 at row e = it-1, input psi0 */
        {
            const int e = it - 1;
            if (e >= 1 && e <= E - 2) {
                real kr, ki, kn;
                rhs(p0r[e], p0i[e], p0r[e - 1] + p0r[e + 1], p0i[e - 1] + p0i[e + 1], n0[e], pp[e] * a.pamp[0], vv[e] * a.vamp[0], kr, ki, kn);
                y1r[e] = p0r[e] + dth * kr;
                y1i[e] = p0i[e] + dth * ki;
                acr[e] = w16 * kr;
                aci[e] = w16 * ki;
                if (RES) {
                    yn1[e] = n0[e] + dth * kn;
                    acn[e] = w16 * kn;
                }
            }
        }
        /* ---- stage 2 at row e = it-2, input y1 */
        {
            const int e = it - 2;
            if (e >= 2 && e <= E - 3) {
                real kr, ki, kn;
                rhs(y1r[e], y1i[e], y1r[e - 1] + y1r[e + 1], y1i[e - 1] + y1i[e + 1], RES ? yn1[e] : real(0), pp[e] * a.pamp[1], vv[e] * a.vamp[1], kr, ki, kn);
                y2r[e] = p0r[e] + dth * kr;
                y2i[e] = p0i[e] + dth * ki;
                acr[e] += w13 * kr;
                aci[e] += w13 * ki;
                if (RES) {
                    yn2[e] = n0[e] + dth * kn;
                    acn[e] += w13 * kn;
                }
            }
        }
        /* ---- stage 3 at row e = it-3, input y2 */
        {
            const int e = it - 3;
            if (e >= 3 && e <= E - 4) {
                real kr, ki, kn;
                rhs(y2r[e], y2i[e], y2r[e - 1] + y2r[e + 1], y2i[e - 1] + y2i[e + 1], RES ? yn2[e] : real(0), pp[e] * a.pamp[2], vv[e] * a.vamp[2], kr, ki, kn);
                y3r[e] = p0r[e] + dt1 * kr;
                y3i[e] = p0i[e] + dt1 * ki;
                acr[e] += w13 * kr;
                aci[e] += w13 * ki;
                if (RES) {
                    yn3[e] = n0[e] + dt1 * kn;
                    acn[e] += w13 * kn;
                }
            }
        }
        /* ---- stage 4 at row e = it-4, input y3, final sum and store */
        {
            const int e = it - 4;
            if (e >= 4 && e <= E - 5) {
                real kr, ki, kn;
                rhs(y3r[e], y3i[e], y3r[e - 1] + y3r[e + 1], y3i[e - 1] + y3i[e + 1], RES ? yn3[e] : real(0), pp[e] * a.pamp[3], vv[e] * a.vamp[3], kr, ki, kn);
                const int grow = y0 + e - kHalo;
                if (out_lane && grow < a.Nr) {
                    const size_t idx = (size_t)grow * (size_t)a.Nc + (size_t)(x0 - kHalo + lane);
                    v2 o;
                    o.x = p0r[e] + dt1 * (acr[e] + w16 * kr);
                    o.y = p0i[e] + dt1 * (aci[e] + w16 * ki);
                    a.psi_out[idx] = o;
                    if (RES) a.n_out[idx] = n0[e] + dt1 * (acn[e] + w16 * kn);
                }
            }
        }
    }
}

template <typename real, int TY, bool RES, bool PUMP, bool POT>
static hipError_t launch_ty(const Rk4Args<real>& a0, hipStream_t s) {
    Rk4Args<real> a = a0;
    a.nstrips = (a.Nc + kOutW - 1) / kOutW;
    a.nbands = (a.Nr + TY - 1) / TY;
    const int bandgroups = (a.nbands + kWavesPerBlock - 1) / kWavesPerBlock;
    const int nblocks = a.nstrips * bandgroups;
    hipLaunchKernelGGL((rk4_scalar_fused<real, TY, RES, PUMP, POT>), dim3(nblocks), dim3(kWave * kWavesPerBlock), 0, s, a);
    return hipGetLastError();
}

template <typename real, bool RES, bool PUMP, bool POT>
static hipError_t launch_flags(const Rk4Args<real>& a, int ty, hipStream_t s) {
    switch (ty) {
        case 8: return launch_ty<real, 8, RES, PUMP, POT>(a, s);
        case 32: return launch_ty<real, 32, RES, PUMP, POT>(a, s);
        default: return launch_ty<real, 16, RES, PUMP, POT>(a, s);
    }
}

template <typename real>
hipError_t launch_rk4_scalar(const Rk4Args<real>& a, int ty, hipStream_t s) {
    const bool res = a.n_in != nullptr;
    const bool pump = res && a.pump != nullptr;
    const bool pot = a.pot != nullptr;
    if (!res) return pot ? launch_flags<real, false, false, true>(a, ty, s) : launch_flags<real, false, false, false>(a, ty, s);
    if (pump) return pot ? launch_flags<real, true, true, true>(a, ty, s) : launch_flags<real, true, true, false>(a, ty, s);
    return pot ? launch_flags<real, true, false, true>(a, ty, s) : launch_flags<real, true, false, false>(a, ty, s);
}

template hipError_t launch_rk4_scalar<float>(const Rk4Args<float>&, int, hipStream_t);
template hipError_t launch_rk4_scalar<double>(const Rk4Args<double>&, int, hipStream_t);

} // namespace phx
