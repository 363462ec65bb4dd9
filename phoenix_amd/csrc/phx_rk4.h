/*
 * phx_rk4.h -- fused RK4 step of the scalar GP + reservoir model for gfx950.
 *
 * One launch = one full RungeKutta4::step (runge_kutta.cu:74-94): halo sync,
 * K1..K4 on shrinking extents, three intermediate sums and the final sum,
 * which the reference runs as 2 + 12*S separate kernels over global memory
 * (cuda_macro.cuh:306-353).  Here every intermediate value stays in registers.
 *
 * Mapping (one wavefront = one tile, no inter-wave synchronisation):
 *   - a wave owns a strip of 64 extended columns (56 output columns plus a
 *     4-cell halo either side), one lane per column;
 *   - it streams TY+8 extended rows top to bottom; after row i is available it
 *     evaluates stage 1 at row i-1, stage 2 at i-2, stage 3 at i-3 and stage 4
 *     at i-4 and writes output row i-4.  Stage s is therefore evaluated on rows
 *     [s, TY+7-s] and is valid on columns [s, 63-s]: the extents that
 *     GENERATE_SUBGRID_INDEX gives with current_halo = 4 - s
 *     (cuda_macro.cuh:43-49,112-113);
 *   - north/south neighbours come from the register window of the previous
 *     stage, east/west neighbours from the adjacent lanes through DPP wave
 *     shifts (one VALU op, no LDS round trip);
 *   - rows are read with raw buffer loads PF rows ahead of their first use.
 *     Each row gets its own buffer descriptor (scalar registers); a row outside
 *     a zero-boundary domain gets an empty descriptor and a column outside it an
 *     out-of-range offset, so the hardware returns 0 for psi, n, P, V -- which is
 *     what synchronize_halos writes into the reference's halos
 *     (kernel_halo.cuh:66-69).  Periodic boundaries wrap the row / column index.
 * Arithmetic uses explicit fused multiply-adds and no automatic contraction, so
 * a cell's result does not depend on the tile shape or its place in the tile.
 * HBM traffic per step and cell: read psi, n, P (, V) once, write psi, n once;
 * the 8-row / 8-column halo re-reads are served from L2.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_device.h"
#include "phx_kernels.h"

namespace phx {

constexpr int kHalo = 4;
constexpr int kOutW = kWave - 2 * kHalo; /* 56 */

/* Neighbour part of the stencil: kin_x (W + (E + r (N+S))) with r = kin_y / kin_x,
   both east/west moves folded into DPP adds.  ISO (dx == dy, r == 1 exactly) drops
   the multiply by r; the two forms give identical bits for dx == dy, so a grid gets
   the same result from either kernel (an ensemble mixing both uses the general one). */
/* CAMP: the pump's temporal amplitude is 1 at every stage (a constant envelope), so P enters as loaded and
   the multiply by the amplitude drops out (x * 1 is exact: the same bits as the general form). */
template <typename real, int TY, bool ISO, int PF, bool RES, bool PUMP, bool POT, bool PERY, int CP = 0, int CPS = CP,
          bool CAMP = false>
__device__ __forceinline__ void rk4_tile(const Rk4Args<real>& a, int strip, int band) {
    using v2 = typename Vec2<real>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int x0 = strip * kOutW;
    const int y0 = band * TY;
    const PhxPhys<real>& ph = a.ph;
    const int Nc = a.Nc, Nr = a.Nr;

    /* column of this lane (wrapped for periodic x); out-of-range offset outside a zero-BC domain */
    int gcol = x0 - kHalo + lane;
    bool col_in = true;
    if (gcol < 0 || gcol >= Nc) {
        if (a.periodic_x)
            gcol = ((gcol % Nc) + Nc) % Nc;
        else
            col_in = false;
    }
    const uint32_t off_r = col_in ? (uint32_t)gcol * (uint32_t)sizeof(real) : kOutOfRange;
    const uint32_t off_c = col_in ? (uint32_t)gcol * (uint32_t)sizeof(v2) : kOutOfRange;
    /* store offsets: halo lanes and lanes past the last column write nothing */
    const bool out_lane = lane >= kHalo && lane < kWave - kHalo && (x0 - kHalo + lane) < Nc;
    const uint32_t so_r = out_lane ? (uint32_t)(x0 - kHalo + lane) * (uint32_t)sizeof(real) : kOutOfRange;
    const uint32_t so_c = out_lane ? (uint32_t)(x0 - kHalo + lane) * (uint32_t)sizeof(v2) : kOutOfRange;
    /* one descriptor per array; a row is an SGPR offset.  Rows outside a zero-BC
       domain read through a descriptor of size 0 (zeros). */
    const uint32_t row_r = (uint32_t)Nc * (uint32_t)sizeof(real);
    const uint32_t row_c = (uint32_t)Nc * (uint32_t)sizeof(v2);
    const uint32_t all_r = row_r * (uint32_t)Nr, all_c = row_c * (uint32_t)Nr;
    const rsrc_t w_psi = row_rsrc(a.psi_out, all_c), w_n = row_rsrc(a.n_out, all_r);
    /* the same arrays with no records: output rows past the grid end store through them (dropped) */
    const rsrc_t w_psi_none = row_rsrc(a.psi_out, 0u), w_n_none = row_rsrc(a.n_out, 0u);
    const RowWrap wrap = make_row_wrap(Nr, kHalo);
    /* tile row r is grid row row0 + rdir * r.  With alt_rows, odd bands stream bottom to top: the 8
       halo rows that two vertically adjacent tiles share are then read by both at the same end of
       their sweeps, close in time, so the second read finds them in L2.  North and south only swap
       places in a sum, so every cell's result is the same either way. */
    const bool up = a.alt_rows && (band & 1);
    const int row0 = up ? y0 + TY + kHalo - 1 : y0 - kHalo, rdir = up ? -1 : 1;

    constexpr int E = TY + 2 * kHalo;
    real p0r[E], p0i[E], n0[E], pp[E], vv[E];
    real y1r[E], y1i[E], y2r[E], y2i[E], y3r[E], y3i[E];
    real yn1[E], yn2[E], yn3[E];
    real acr[E], aci[E], acn[E]; /* psi0 + dt/6 k1 + dt/3 k2 + dt/3 k3, built stage by stage */

    const real dth = a.dt * real(0.5);
    const real dt1 = a.dt;
    const real dt6 = a.dt / real(6);
    const real dt3 = a.dt / real(3);
    PhxPhys<real> phv = ph;
    phv.mhalf_gamma_c = in_vgpr(ph.mhalf_gamma_c);
    phv.gamma_r = in_vgpr(ph.gamma_r);
    const real kin_x = ph.kin_x, kin_yx = ph.kin_yx;

    auto load_row = [&](int r) {
        bool in;
        const uint32_t grow = tile_row(row0 + rdir * r, wrap, PERY, in); /* wave-uniform */
        const uint32_t sc = grow * row_c, sr = grow * row_r;
        const uint32_t nc = in ? all_c : 0u, nr = in ? all_r : 0u;
        const auto x = load_cplx_s<real, CP>(row_rsrc(a.psi_in, nc), off_c, sc);
        p0r[r] = x.x;
        p0i[r] = x.y;
        if (RES) n0[r] = load_real_s<real, CP>(row_rsrc(a.n_in, nr), off_r, sr);
        if (PUMP) pp[r] = load_real_s<real>(row_rsrc(a.pump, nr), off_r, sr);
        if (POT) vv[r] = load_real_s<real>(row_rsrc(a.pot, nr), off_r, sr);
    };

    /* neighbour part of the kinetic stencil: centre c, north+south ns */
    auto nb = [&](real c, real ns) {
        if constexpr (ISO)
            return kin_x * (west(c) + (east(c) + ns));
        else
            return kin_x * (west(c) + (east(c) + kin_yx * ns));
    };
    /* one RHS evaluation: centre (cr, ci), north+south sums, reservoir, pump*amp, V*amp/hbar */
    auto rhs = [&](real cr, real ci, real nsr, real nsi, real nv, real pmp, real poth, real& kr, real& ki, real& kn) {
        phx_rhs_scalar<real>(phv, cr, ci, nb(cr, nsr), nb(ci, nsi), RES ? nv : real(0), PUMP ? pmp : real(0), POT ? poth : real(0),
                             kr, ki, kn);
    };

#pragma unroll
    for (int it = -PF; it < E; ++it) {
        if (it + PF < E) load_row(it + PF);
        if (it < 0) continue;
        /* ---- stage 1 at row e = it-1, input psi0 */
        {
            const int e = it - 1;
            if (e >= 1 && e <= E - 2) {
                real kr, ki, kn;
                rhs(p0r[e], p0i[e], p0r[e - 1] + p0r[e + 1], p0i[e - 1] + p0i[e + 1], n0[e], (CAMP ? pp[e] : pp[e] * a.pamp[0]),
                    vv[e] * a.vamp_h[0], kr, ki, kn);
                y1r[e] = phx_fma(dth, kr, p0r[e]);
                y1i[e] = phx_fma(dth, ki, p0i[e]);
                acr[e] = phx_fma(dt6, kr, p0r[e]);
                aci[e] = phx_fma(dt6, ki, p0i[e]);
                if (RES) {
                    yn1[e] = phx_fma(dth, kn, n0[e]);
                    acn[e] = phx_fma(dt6, kn, n0[e]);
                }
            }
        }
        /* ---- stage 2 at row e = it-2, input y1 */
        {
            const int e = it - 2;
            if (e >= 2 && e <= E - 3) {
                real kr, ki, kn;
                rhs(y1r[e], y1i[e], y1r[e - 1] + y1r[e + 1], y1i[e - 1] + y1i[e + 1], RES ? yn1[e] : real(0),
                    (CAMP ? pp[e] : pp[e] * a.pamp[1]), vv[e] * a.vamp_h[1], kr, ki, kn);
                y2r[e] = phx_fma(dth, kr, p0r[e]);
                y2i[e] = phx_fma(dth, ki, p0i[e]);
                acr[e] = phx_fma(dt3, kr, acr[e]);
                aci[e] = phx_fma(dt3, ki, aci[e]);
                if (RES) {
                    yn2[e] = phx_fma(dth, kn, n0[e]);
                    acn[e] = phx_fma(dt3, kn, acn[e]);
                }
            }
        }
        /* ---- stage 3 at row e = it-3, input y2 */
        {
            const int e = it - 3;
            if (e >= 3 && e <= E - 4) {
                real kr, ki, kn;
                rhs(y2r[e], y2i[e], y2r[e - 1] + y2r[e + 1], y2i[e - 1] + y2i[e + 1], RES ? yn2[e] : real(0),
                    (CAMP ? pp[e] : pp[e] * a.pamp[2]), vv[e] * a.vamp_h[2], kr, ki, kn);
                y3r[e] = phx_fma(dt1, kr, p0r[e]);
                y3i[e] = phx_fma(dt1, ki, p0i[e]);
                acr[e] = phx_fma(dt3, kr, acr[e]);
                aci[e] = phx_fma(dt3, ki, aci[e]);
                if (RES) {
                    yn3[e] = phx_fma(dt1, kn, n0[e]);
                    acn[e] = phx_fma(dt3, kn, acn[e]);
                }
            }
        }
        /* ---- stage 4 at row e = it-4, input y3; psi0 + dt/6 (k1 + 2k2 + 2k3 + k4) */
        {
            const int e = it - 4;
            if (e >= 4 && e <= E - 5) {
                real kr, ki, kn;
                rhs(y3r[e], y3i[e], y3r[e - 1] + y3r[e + 1], y3i[e - 1] + y3i[e + 1], RES ? yn3[e] : real(0),
                    (CAMP ? pp[e] : pp[e] * a.pamp[3]), vv[e] * a.vamp_h[3], kr, ki, kn);
                /* rows past the last grid row store through an empty descriptor: no branch */
                const int grow = row0 + rdir * e;
                const bool row_ok = grow < Nr;
                v2 o;
                o.x = phx_fma(dt6, kr, acr[e]);
                o.y = phx_fma(dt6, ki, aci[e]);
                store_cplx_s<CPS>(row_ok ? w_psi : w_psi_none, so_c, (uint32_t)grow * row_c, o);
                if (RES) store_real_s<CPS>(row_ok ? w_n : w_n_none, so_r, (uint32_t)grow * row_r, phx_fma(dt6, kn, acn[e]));
            }
        }
    }
}

/* one grid: tile (strip, band) from the block index, 4 vertically stacked bands per block */
template <typename real, int TY, bool ISO, int PF, bool RES, bool PUMP, bool POT, bool PERY>
__global__ __launch_bounds__(256) void rk4_scalar_fused(Rk4Args<real> a) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const int strip = tile % a.nstrips;
    const int band = (tile / a.nstrips) * kWavesPerBlock + wave;
    if (band >= a.nbands) return;
    rk4_tile<real, TY, ISO, PF, RES, PUMP, POT, PERY>(a, strip, band);
}

/* Many steps in one launch, one resident wavefront per tile.  Before step k a tile waits
   until the eight tiles whose cells its 4-cell halo covers have finished step k-1 (their
   step counters reach base + k): their outputs of step k-1 are its inputs, and their reads
   of step k-1, which covered its cells in the buffer it is about to overwrite, are done.
   psi and n go through device-scope loads and stores (no stale line of another XCD's
   write is used); a counter moves only after all stores of the step have completed.  Each
   counter has a 128-byte line of its own: packed 32 to a line, the polls of neighbouring
   tiles contended for the same lines and a step took 21 instead of 8 us at 1024^2.  A
   wait longer than spin_limit polls sets *error and ends the wave (its neighbours then
   time out too, so every wave exits); the state is then undefined.  Tiles of the last
   band and strip must be at least kHalo cells deep/wide, so that a halo never reaches
   past the neighbouring tile (rk4_flow_shape_ok). */
template <typename real, int TY, bool ISO, int PF, bool RES, bool PUMP, bool POT, bool PERY, bool CAMP = false>
__global__ __launch_bounds__(256) void rk4_scalar_flow(Rk4Args<real> a, Rk4Flow f) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const int strip = tile % a.nstrips;
    const int band = (tile / a.nstrips) * kWavesPerBlock + wave;
    if (band >= a.nbands) return; /* no tile: nobody watches this wave's counter */
    const int me = band * a.nstrips + strip;
    /* lanes 0..7 watch the neighbours (d = 0..8 without the centre 4) */
    int watch = -1;
    if (lane < 8) {
        const int d = lane + (lane >= 4 ? 1 : 0);
        int ns = strip + d % 3 - 1, nb = band + d / 3 - 1;
        bool ok = true;
        if (ns < 0 || ns >= a.nstrips) {
            if (a.periodic_x)
                ns = (ns + a.nstrips) % a.nstrips;
            else
                ok = false;
        }
        if (nb < 0 || nb >= a.nbands) {
            if (PERY)
                nb = (nb + a.nbands) % a.nbands;
            else
                ok = false;
        }
        if (ok) watch = nb * a.nstrips + ns;
    }
    int* const mine = f.counters + (size_t)me * kFlowCounterStride;
    int* const theirs = watch < 0 ? nullptr : f.counters + (size_t)watch * kFlowCounterStride;
    Rk4Args<real> s = a;
    for (int k = 0; k < f.steps; k++) {
        if (k > 0) {
            const int need = f.base + k;
            int polls = 0;
            bool ready = theirs == nullptr;
            for (;;) {
                if (!ready) ready = __hip_atomic_load(theirs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need;
                if (__builtin_amdgcn_ballot_w64(!ready) == 0) break;
                if (++polls > f.spin_limit) {
                    if (lane == 0) __hip_atomic_store(f.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            asm volatile("" ::: "memory");
        }
        const bool even = (k & 1) == 0;
        s.psi_in = even ? a.psi_in : a.psi_out;
        s.psi_out = even ? (typename Vec2<real>::type*)a.psi_out : (typename Vec2<real>::type*)a.psi_in;
        s.n_in = even ? a.n_in : a.n_out;
        s.n_out = even ? a.n_out : (real*)a.n_in;
        rk4_tile<real, TY, ISO, PF, RES, PUMP, POT, PERY, kDeviceScope, kDeviceScope, CAMP>(s, strip, band);
        /* all of this step's stores have completed before the counter moves */
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0) __hip_atomic_store(mine, f.base + k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

/* ensemble: one flattened list of wavefront tiles over all instances; each wave looks up
   its instance and steps one tile of it (instances of any size in the same launch).
   Without an instance table: `items` grids of the common size at a fixed element stride. */
template <typename real, int TY, bool ISO, int PF, bool RES, bool PUMP, bool POT, bool PERY>
__global__ __launch_bounds__(256) void rk4_batched_fused(Rk4BatchArgs<real> b) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int gw = xcd_contiguous(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
    if (gw >= b.ntiles) return;
    Rk4Args<real> a = b.common;
    int local;
    if (b.entries) {
        const int inst = __builtin_amdgcn_readfirstlane(b.tile_inst[gw]);
        const Rk4BatchEntry<real>& en = b.entries[inst];
        a.psi_in = en.psi_in;
        a.n_in = en.n_in;
        a.psi_out = en.psi_out;
        a.n_out = en.n_out;
        a.pump = en.pump;
        a.pot = en.pot;
        a.Nc = en.Nc;
        a.Nr = en.Nr;
        a.nstrips = en.nstrips;
        a.ph.kin_c = en.kin_c;
        a.ph.kin_x = en.kin_x;
        a.ph.kin_y = en.kin_y;
        a.ph.kin_yx = en.kin_yx;
        local = gw - en.tile0;
    } else {
        const int inst = gw / b.tiles_per_item;
        const size_t so = (size_t)inst * b.item_stride, po = (size_t)inst * b.param_stride;
        a.psi_in += so;
        a.psi_out += so;
        if (RES) {
            a.n_in += so;
            a.n_out += so;
        }
        if (PUMP) a.pump += po;
        if (POT) a.pot += po;
        a.nstrips = rk4_strips(a.Nc);
        local = gw - inst * b.tiles_per_item;
    }
    rk4_tile<real, TY, ISO, PF, RES, PUMP, POT, PERY>(a, local % a.nstrips, local / a.nstrips);
}

template <typename real, int TY, bool ISO, int PF, bool RES, bool PUMP, bool POT>
inline hipError_t launch_ty(const Rk4Args<real>& a0, hipStream_t s) {
    Rk4Args<real> a = a0;
    a.nstrips = (a.Nc + kOutW - 1) / kOutW;
    a.nbands = (a.Nr + TY - 1) / TY;
    const int bandgroups = (a.nbands + kWavesPerBlock - 1) / kWavesPerBlock;
    const int nblocks = a.nstrips * bandgroups;
    /* periodic rows: a separate kernel, whose row descriptors are loop invariant (fewer
       scalar registers than the zero-boundary form, which selects an empty descriptor for
       rows outside the grid) */
    if (a.periodic_y)
        hipLaunchKernelGGL((rk4_scalar_fused<real, TY, ISO, PF, RES, PUMP, POT, true>), dim3(nblocks),
                           dim3(kWave * kWavesPerBlock), 0, s, a);
    else
        hipLaunchKernelGGL((rk4_scalar_fused<real, TY, ISO, PF, RES, PUMP, POT, false>), dim3(nblocks),
                           dim3(kWave * kWavesPerBlock), 0, s, a);
    return hipGetLastError();
}

/* Tile height TY in {8, 10, 12, 16, 20, 24} (others: 16). */
template <typename real, bool ISO, int PF, bool RES, bool PUMP, bool POT>
inline hipError_t launch_tys(const Rk4Args<real>& a, const Rk4Tune& t, hipStream_t s) {
    switch (t.ty) {
        case 8: return launch_ty<real, 8, ISO, PF, RES, PUMP, POT>(a, s);
        case 10: return launch_ty<real, 10, ISO, PF, RES, PUMP, POT>(a, s);
        case 12: return launch_ty<real, 12, ISO, PF, RES, PUMP, POT>(a, s);
        case 20: return launch_ty<real, 20, ISO, PF, RES, PUMP, POT>(a, s);
        case 24: return launch_ty<real, 24, ISO, PF, RES, PUMP, POT>(a, s);
        default: return launch_ty<real, 16, ISO, PF, RES, PUMP, POT>(a, s);
    }
}

/* Rows loaded ahead: 2, or 3, 4, 6 or 8 for the pumped reservoir model with dx == dy
   (the benchmark configuration; chosen through Rk4Tune::pf). */
template <typename real, bool RES, bool PUMP, bool POT>
inline hipError_t launch_flags(const Rk4Args<real>& a, const Rk4Tune& t, hipStream_t s) {
    const bool iso = a.ph.kin_yx == real(1);
    if constexpr (RES && PUMP && !POT) {
        if (iso && t.pf == 3) return launch_tys<real, true, 3, RES, PUMP, POT>(a, t, s);
        if (iso && (t.pf == 4 || t.pf == 5)) return launch_tys<real, true, 4, RES, PUMP, POT>(a, t, s);
        if (iso && (t.pf == 6 || t.pf == 7)) return launch_tys<real, true, 6, RES, PUMP, POT>(a, t, s);
        if (iso && t.pf >= 8) return launch_tys<real, true, 8, RES, PUMP, POT>(a, t, s);
    }
    return iso ? launch_tys<real, true, 2, RES, PUMP, POT>(a, t, s) : launch_tys<real, false, 2, RES, PUMP, POT>(a, t, s);
}

template <typename real>
hipError_t launch_rk4_scalar(const Rk4Args<real>& a, const Rk4Tune& t, hipStream_t s) {
    if ((size_t)a.Nc * (size_t)a.Nr * sizeof(typename Vec2<real>::type) > kMaxArrayBytes) return hipErrorInvalidValue;
    const bool res = a.n_in != nullptr;
    const bool pump = res && a.pump != nullptr;
    const bool pot = a.pot != nullptr;
    if (!res) return pot ? launch_flags<real, false, false, true>(a, t, s) : launch_flags<real, false, false, false>(a, t, s);
    if (pump) return pot ? launch_flags<real, true, true, true>(a, t, s) : launch_flags<real, true, true, false>(a, t, s);
    return pot ? launch_flags<real, true, false, true>(a, t, s) : launch_flags<real, true, false, false>(a, t, s);
}

/* ---------------------------------------------------------------- many steps per launch */
inline bool rk4_flow_shape_ok(int Nc, int Nr, int ty) {
    const int last_rows = Nr - ((Nr - 1) / ty) * ty, last_cols = Nc - ((Nc - 1) / kOutW) * kOutW;
    return Nr >= ty && Nc >= kOutW && last_rows >= kHalo && last_cols >= kHalo;
}

template <typename real, int TY, bool ISO, int PF, bool RES, bool PUMP, bool POT, bool CAMP = false>
inline hipError_t launch_flow_ty(const Rk4Args<real>& a0, const Rk4Flow& f, hipStream_t s) {
    Rk4Args<real> a = a0;
    a.nstrips = (a.Nc + kOutW - 1) / kOutW;
    a.nbands = (a.Nr + TY - 1) / TY;
    const int nblocks = a.nstrips * ((a.nbands + kWavesPerBlock - 1) / kWavesPerBlock);
    const void* fn = a.periodic_y ? (const void*)&rk4_scalar_flow<real, TY, ISO, PF, RES, PUMP, POT, true, CAMP>
                                  : (const void*)&rk4_scalar_flow<real, TY, ISO, PF, RES, PUMP, POT, false, CAMP>;
    /* every wave must be resident at once: the cooperative launch refuses a grid that is not */
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWave * kWavesPerBlock, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return hipErrorNotSupported;
    if ((long)per_cu * cus < nblocks) return hipErrorCooperativeLaunchTooLarge;
    if (!f.counters) return hipSuccess; /* dry run: the kernel takes this grid */
    Rk4Flow fl = f;
    void* args[] = {&a, &fl};
    return hipLaunchCooperativeKernel(fn, dim3(nblocks), dim3(kWave * kWavesPerBlock), args, 0, s);
}

template <typename real, bool ISO, int PF, bool RES, bool PUMP, bool POT, bool CAMP = false>
inline hipError_t launch_flow_tys(const Rk4Args<real>& a, const Rk4Tune& t, const Rk4Flow& f, hipStream_t s) {
    switch (t.ty) {
        case 8: return launch_flow_ty<real, 8, ISO, PF, RES, PUMP, POT, CAMP>(a, f, s);
        case 10: return launch_flow_ty<real, 10, ISO, PF, RES, PUMP, POT, CAMP>(a, f, s);
        case 12: return launch_flow_ty<real, 12, ISO, PF, RES, PUMP, POT, CAMP>(a, f, s);
        case 20: return launch_flow_ty<real, 20, ISO, PF, RES, PUMP, POT, CAMP>(a, f, s);
        case 16: return launch_flow_ty<real, 16, ISO, PF, RES, PUMP, POT, CAMP>(a, f, s);
        default: return hipErrorNotSupported;
    }
}

template <typename real, bool RES, bool PUMP, bool POT>
inline hipError_t launch_flow_flags(const Rk4Args<real>& a, const Rk4Tune& t, const Rk4Flow& f, hipStream_t s) {
    const bool iso = a.ph.kin_yx == real(1);
    if constexpr (RES && PUMP && !POT) {
        /* a pump of constant envelope (amplitude 1 at every stage) at the prefetch depths the
           multi-step autotuner picks (4 rows FP64, 6 rows FP32) */
        const bool camp = a.pamp[0] == real(1) && a.pamp[1] == real(1) && a.pamp[2] == real(1) && a.pamp[3] == real(1);
        if (iso && camp && (t.pf == 4 || t.pf == 5)) return launch_flow_tys<real, true, 4, RES, PUMP, POT, true>(a, t, f, s);
        if (iso && camp && (t.pf == 6 || t.pf == 7)) return launch_flow_tys<real, true, 6, RES, PUMP, POT, true>(a, t, f, s);
        if (iso && t.pf == 3) return launch_flow_tys<real, true, 3, RES, PUMP, POT>(a, t, f, s);
        if (iso && (t.pf == 4 || t.pf == 5)) return launch_flow_tys<real, true, 4, RES, PUMP, POT>(a, t, f, s);
        if (iso && (t.pf == 6 || t.pf == 7)) return launch_flow_tys<real, true, 6, RES, PUMP, POT>(a, t, f, s);
        if (iso && t.pf >= 8) return launch_flow_tys<real, true, 8, RES, PUMP, POT>(a, t, f, s);
    }
    return iso ? launch_flow_tys<real, true, 2, RES, PUMP, POT>(a, t, f, s) : launch_flow_tys<real, false, 2, RES, PUMP, POT>(a, t, f, s);
}

/* f.steps RK4 steps from a.psi_in / a.n_in; an even count leaves the result there, an odd
   one in a.psi_out / a.n_out.  Errors: hipErrorNotSupported for a shape or tile height the
   kernel does not take, hipErrorCooperativeLaunchTooLarge when its waves cannot all be
   resident; the caller then steps launch by launch.  f.counters == nullptr: only check. */
template <typename real>
hipError_t launch_rk4_flow(const Rk4Args<real>& a, const Rk4Tune& t, const Rk4Flow& f, hipStream_t s) {
    if ((size_t)a.Nc * (size_t)a.Nr * sizeof(typename Vec2<real>::type) > kMaxArrayBytes) return hipErrorInvalidValue;
    if (!rk4_flow_shape_ok(a.Nc, a.Nr, t.ty) || f.steps <= 0) return hipErrorNotSupported;
    const bool res = a.n_in != nullptr;
    const bool pump = res && a.pump != nullptr;
    const bool pot = a.pot != nullptr;
    if (!res) return pot ? launch_flow_flags<real, false, false, true>(a, t, f, s) : launch_flow_flags<real, false, false, false>(a, t, f, s);
    if (pump) return pot ? launch_flow_flags<real, true, true, true>(a, t, f, s) : launch_flow_flags<real, true, true, false>(a, t, f, s);
    return pot ? launch_flow_flags<real, true, false, true>(a, t, f, s) : launch_flow_flags<real, true, false, false>(a, t, f, s);
}

/* ---------------------------------------------------------------- ensembles */
template <typename real, int TY, bool ISO, bool RES, bool PUMP, bool POT>
inline hipError_t launch_batch_ty(const Rk4BatchArgs<real>& b, hipStream_t s) {
    const int nblocks = (b.ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
    if (b.common.periodic_y)
        hipLaunchKernelGGL((rk4_batched_fused<real, TY, ISO, 2, RES, PUMP, POT, true>), dim3(nblocks),
                           dim3(kWave * kWavesPerBlock), 0, s, b);
    else
        hipLaunchKernelGGL((rk4_batched_fused<real, TY, ISO, 2, RES, PUMP, POT, false>), dim3(nblocks),
                           dim3(kWave * kWavesPerBlock), 0, s, b);
    return hipGetLastError();
}

template <typename real, bool ISO, bool RES, bool PUMP, bool POT>
inline hipError_t launch_batch_tys(const Rk4BatchArgs<real>& b, int ty, hipStream_t s) {
    switch (ty) {
        case 8: return launch_batch_ty<real, 8, ISO, RES, PUMP, POT>(b, s);
        case 10: return launch_batch_ty<real, 10, ISO, RES, PUMP, POT>(b, s);
        case 16: return launch_batch_ty<real, 16, ISO, RES, PUMP, POT>(b, s);
        case 20: return launch_batch_ty<real, 20, ISO, RES, PUMP, POT>(b, s);
        default: return launch_batch_ty<real, 12, ISO, RES, PUMP, POT>(b, s);
    }
}

template <typename real, bool RES, bool PUMP, bool POT>
inline hipError_t launch_batch_flags(const Rk4BatchArgs<real>& b, int ty, hipStream_t s) {
    return b.iso ? launch_batch_tys<real, true, RES, PUMP, POT>(b, ty, s) : launch_batch_tys<real, false, RES, PUMP, POT>(b, ty, s);
}

template <typename real>
hipError_t launch_rk4_batched(const Rk4BatchArgs<real>& b, int ty, bool res, bool pump, bool pot, hipStream_t s) {
    /* (instance sizes are checked when the ensemble is created) */
    if (!b.entries && (size_t)b.common.Nc * (size_t)b.common.Nr * sizeof(typename Vec2<real>::type) > kMaxArrayBytes)
        return hipErrorInvalidValue;
    if (!res) return pot ? launch_batch_flags<real, false, false, true>(b, ty, s) : launch_batch_flags<real, false, false, false>(b, ty, s);
    if (pump) return pot ? launch_batch_flags<real, true, true, true>(b, ty, s) : launch_batch_flags<real, true, true, false>(b, ty, s);
    return pot ? launch_batch_flags<real, true, false, true>(b, ty, s) : launch_batch_flags<real, true, false, false>(b, ty, s);
}


} // namespace phx
