/*
 * phx_adjoint.hip -- discrete adjoint of the fused RK4 step (scalar GP +
 * reservoir) and the checkpointed forward/backward driver behind
 * phx_adjoint_* (gradients of an objective of the final state with respect
 * to the pump P(x,y), the potential V(x,y) and the initial state).
 *
 * The reference has no adjoint; the forward map differentiated here is its
 * RungeKutta4::step (runge_kutta.cu:74-94) with the RHS of gp_scalar
 * (kernel_gp_compute.cuh:50-110), as computed by rk4_scalar_fused.
 *
 * One step x' = x + dt/6 (k1 + 2 k2 + 2 k3 + k4), k_s = f(Y_s), Y_1 = x,
 * Y_2 = x + dt/2 k1, Y_3 = x + dt/2 k2, Y_4 = x + dt k3, is reversed as
 *   c4 = dt/6 l',            Yb4 = J(Y_4)^T c4
 *   c3 = dt/3 l' + dt Yb4,   Yb3 = J(Y_3)^T c3
 *   c2 = dt/3 l' + dt/2 Yb3, Yb2 = J(Y_2)^T c2
 *   c1 = dt/6 l' + dt/2 Yb2, Yb1 = J(Y_1)^T c1
 *   l  = l' + (Yb4 + Yb3 + Yb2) + Yb1,  dP += sum_s dk_s/dP^T c_s,  dV likewise
 * (J^T from phx_rhs_scalar_vjp).  The stage states are recomputed from x.
 *
 * Kernel mapping, as in the forward kernel: one wavefront streams a strip of
 * 64 extended columns (52 output columns, halo 6) top to bottom.  After row i
 * is loaded it evaluates the forward stages at rows i-1, i-2, i-3 and the
 * reverse stages at rows i-3 (c4 -> Yb4), i-4, i-5 and i-6, and writes row
 * i-6.  Every stage value lives in an 8-row register window indexed modulo 8,
 * so the row loop is unrolled 8 times whatever the tile height (a runtime
 * value): the code size does not grow with the tile.  Stages are evaluated on
 * every row of the loop, valid or not (only the store is guarded): a slot
 * that is written on every pass is dead between its last read and the next
 * write, which keeps the windows in about half the registers that
 * conditionally written slots would hold.  Rows where a stage is not valid
 * (before its inputs exist, or past the tile) produce values that no valid
 * row reads.
 * HBM traffic per cell and step: read x (psi, n), l' (psi, n), P, V; write l;
 * read+write dP, dV.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "custom_rhs.h"
#include "phx_config.hpp"
#include "phx_device.h"
#include "phx_kernels.h"

namespace phx {

int set_error(int code, const std::string& msg); /* phx_solver.hip */

namespace {
constexpr int kAdjOutW = kWave - 2 * kAdjHalo; /* 52 */
constexpr int kWin = 8;                        /* register window (rows) */
constexpr int kPF = 1;                         /* rows loaded ahead */

/* register slot of row it + d when `it` = base + j and base = -kPF (mod kWin) */
__host__ __device__ constexpr int slot(int j, int d) { return (j - kPF + d + 8 * kWin) & (kWin - 1); }
} // namespace

template <typename real, bool ISO, bool RES, bool PUMP, bool POT>
__global__ __launch_bounds__(256) void rk4_adjoint_fused(Rk4AdjArgs<real> a) {
    using v2 = typename Vec2<real>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int gw = xcd_contiguous(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
    if (gw >= a.ntiles) return;
    const int item = gw / a.tiles_per_item;
    const int local = gw - item * a.tiles_per_item;
    const int strip = local % a.nstrips, band = local / a.nstrips;
    const size_t so = (size_t)item * a.item_stride, po = (size_t)item * a.param_stride;
    const v2* psi = a.psi + so;
    const real* nres = RES ? a.n + so : nullptr;
    const v2* lpsi = a.lam_psi_in + so;
    const real* lres = RES ? a.lam_n_in + so : nullptr;
    const real* pump = PUMP ? a.pump + po : nullptr;
    const real* pot = POT ? a.pot + po : nullptr;

    const int TY = a.ty;
    const int E = TY + 2 * kAdjHalo;
    const int x0 = strip * kAdjOutW, y0 = band * TY;
    const int Nc = a.Nc, Nr = a.Nr;

    int gcol = x0 - kAdjHalo + lane;
    bool col_in = true;
    if (gcol < 0 || gcol >= Nc) {
        if (a.periodic_x)
            gcol = ((gcol % Nc) + Nc) % Nc;
        else
            col_in = false;
    }
    const uint32_t off_r = col_in ? (uint32_t)gcol * (uint32_t)sizeof(real) : kOutOfRange;
    const uint32_t off_c = col_in ? (uint32_t)gcol * (uint32_t)sizeof(v2) : kOutOfRange;
    const bool out_lane = lane >= kAdjHalo && lane < kWave - kAdjHalo && (x0 - kAdjHalo + lane) < Nc;
    const uint32_t rbytes = (uint32_t)Nc * (uint32_t)sizeof(real);
    const uint32_t cbytes = (uint32_t)Nc * (uint32_t)sizeof(v2);

    PhxPhys<real> ph = a.ph;
    ph.mhalf_gamma_c = in_vgpr(ph.mhalf_gamma_c);
    ph.gamma_r = in_vgpr(ph.gamma_r);
    const real kin_x = ph.kin_x;
    const real dth = a.dt * real(0.5), dt1 = a.dt, dt6 = a.dt / real(6), dt3 = a.dt / real(3);

    /* state, parameters and incoming cotangent */
    real xu[kWin], xv[kWin], xn[kWin], pp[kWin], vv[kWin], lu[kWin], lv[kWin], ln[kWin];
    /* forward stage states Y_2, Y_3 */
    real y2u[kWin], y2v[kWin], y2n[kWin], y3u[kWin], y3v[kWin], y3n[kWin];
    /* stage cotangents c3, c2, c1, running sum of Yb, parameter cotangents */
    real c3u[kWin], c3v[kWin], c3n[kWin], c2u[kWin], c2v[kWin], c2n[kWin], c1u[kWin], c1v[kWin], c1n[kWin];
    real au[kWin], av[kWin], an[kWin], gpa[kWin], gva[kWin];
    real d3u[kWin], d3v[kWin], d3n[kWin]; /* dt/3 l' */
#pragma unroll
    for (int k = 0; k < kWin; k++) {
        d3u[k] = d3v[k] = d3n[k] = real(0);
        xu[k] = xv[k] = xn[k] = pp[k] = vv[k] = lu[k] = lv[k] = ln[k] = real(0);
        y2u[k] = y2v[k] = y2n[k] = y3u[k] = y3v[k] = y3n[k] = real(0);
        c3u[k] = c3v[k] = c3n[k] = c2u[k] = c2v[k] = c2n[k] = c1u[k] = c1v[k] = c1n[k] = real(0);
        au[k] = av[k] = an[k] = gpa[k] = gva[k] = real(0);
    }

    auto load_row = [&](int r, int sl) {
        int grow = y0 - kAdjHalo + r;
        bool in = r < E; /* rows past the tile read as zero (no memory access) */
        if (in && (grow < 0 || grow >= Nr)) {
            if (a.periodic_y)
                grow = ((grow % Nr) + Nr) % Nr;
            else
                in = false;
        }
        const size_t row = (size_t)(in ? grow : 0) * (size_t)Nc;
        const v2 x = load_cplx<real>(row_rsrc(psi + row, in ? cbytes : 0u), off_c);
        xu[sl] = x.x;
        xv[sl] = x.y;
        const v2 l = load_cplx<real>(row_rsrc(lpsi + row, in ? cbytes : 0u), off_c);
        lu[sl] = l.x;
        lv[sl] = l.y;
        if (RES) {
            xn[sl] = load_real<real>(row_rsrc(nres + row, in ? rbytes : 0u), off_r);
            ln[sl] = load_real<real>(row_rsrc(lres + row, in ? rbytes : 0u), off_r);
        }
        if (PUMP) pp[sl] = load_real<real>(row_rsrc(pump + row, in ? rbytes : 0u), off_r);
        if (POT) vv[sl] = load_real<real>(row_rsrc(pot + row, in ? rbytes : 0u), off_r);
    };
    /* neighbour part kin_x (W + (E + r (N+S))) of the stencil, r = kin_y / kin_x (the RHS
       hooks add the centre); for dx == dy r is 1 and its multiply is skipped */
    const real kin_yx = ph.kin_yx;
    auto nbr = [&](real c, real ns) {
        if constexpr (ISO)
            return kin_x * (west(c) + (east(c) + ns));
        else
            return kin_x * (west(c) + (east(c) + kin_yx * ns));
    };
    auto rhs = [&](real u, real v, real nsu, real nsv, real n, real p, real vh, real& ku, real& kv, real& kn) {
        phx_rhs_scalar<real>(ph, u, v, nbr(u, nsu), nbr(v, nsv), RES ? n : real(0), PUMP ? p : real(0), POT ? vh : real(0), ku,
                             kv, kn);
    };

    for (int base = -kPF; base < E; base += kWin) {
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            const int it = base + j;
            load_row(it + kPF, slot(j, kPF));
            /* ---- F1 at e = it-1: Y_2 = x + dt/2 f(x) */
            {
                const int e = it - 1, s = slot(j, -1), sn = slot(j, -2), ss = slot(j, 0);
                {
                    real ku, kv, kn;
                    rhs(xu[s], xv[s], xu[sn] + xu[ss], xv[sn] + xv[ss], xn[s], pp[s] * a.pamp[0], vv[s] * a.vamp_h[0], ku, kv, kn);
                    y2u[s] = phx_fma(dth, ku, xu[s]);
                    y2v[s] = phx_fma(dth, kv, xv[s]);
                    if (RES) y2n[s] = phx_fma(dth, kn, xn[s]);
                }
            }
            /* ---- F2 at e = it-2: Y_3 = x + dt/2 f(Y_2) */
            {
                const int e = it - 2, s = slot(j, -2), sn = slot(j, -3), ss = slot(j, -1);
                {
                    real ku, kv, kn;
                    rhs(y2u[s], y2v[s], y2u[sn] + y2u[ss], y2v[sn] + y2v[ss], RES ? y2n[s] : real(0), pp[s] * a.pamp[1],
                        vv[s] * a.vamp_h[1], ku, kv, kn);
                    y3u[s] = phx_fma(dth, ku, xu[s]);
                    y3v[s] = phx_fma(dth, kv, xv[s]);
                    if (RES) y3n[s] = phx_fma(dth, kn, xn[s]);
                }
            }
            /* ---- F3 + B4 at e = it-3: Y_4 = x + dt f(Y_3); c4 = dt/6 l', Yb4 = J(Y_4)^T c4, c3 */
            {
                const int e = it - 3, s = slot(j, -3), sn = slot(j, -4), ss = slot(j, -2);
                {
                    real ku, kv, kn;
                    rhs(y3u[s], y3v[s], y3u[sn] + y3u[ss], y3v[sn] + y3v[ss], RES ? y3n[s] : real(0), pp[s] * a.pamp[2],
                        vv[s] * a.vamp_h[2], ku, kv, kn);
                    const real y4u = phx_fma(dt1, ku, xu[s]);
                    const real y4v = phx_fma(dt1, kv, xv[s]);
                    const real y4n = RES ? phx_fma(dt1, kn, xn[s]) : real(0);
                    const real c4u = dt6 * lu[s], c4v = dt6 * lv[s], c4n = RES ? dt6 * ln[s] : real(0);
                    const real K4u = dt6 * nbr(lu[s], lu[sn] + lu[ss]);
                    const real K4v = dt6 * nbr(lv[s], lv[sn] + lv[ss]);
                    real bu, bv, bn;
                    phx_rhs_scalar_vjp<real>(ph, y4u, y4v, y4n, POT ? vv[s] * a.vamp_h[3] : real(0), c4u, c4v, c4n, K4u, K4v, bu,
                                             bv, bn);
                    au[s] = bu;
                    av[s] = bv;
                    d3u[s] = dt3 * lu[s];
                    d3v[s] = dt3 * lv[s];
                    c3u[s] = phx_fma(dt1, bu, d3u[s]);
                    c3v[s] = phx_fma(dt1, bv, d3v[s]);
                    real gp4, gv4, gp3, gv3;
                    phx_rhs_scalar_param_vjp<real>(y4u, y4v, c4u, c4v, c4n, a.pamp[3], a.vamp_h[3], gp4, gv4);
                    if (RES) {
                        an[s] = bn;
                        d3n[s] = dt3 * ln[s];
                        c3n[s] = phx_fma(dt1, bn, d3n[s]);
                    }
                    phx_rhs_scalar_param_vjp<real>(y3u[s], y3v[s], c3u[s], c3v[s], RES ? c3n[s] : real(0), a.pamp[2], a.vamp_h[2],
                                                   gp3, gv3);
                    gpa[s] = gp4 + gp3;
                    gva[s] = gv4 + gv3;
                }
            }
            /* ---- B3 at e = it-4: Yb3 = J(Y_3)^T c3, c2 */
            {
                const int e = it - 4, s = slot(j, -4), sn = slot(j, -5), ss = slot(j, -3);
                {
                    const real Ku = nbr(c3u[s], c3u[sn] + c3u[ss]);
                    const real Kv = nbr(c3v[s], c3v[sn] + c3v[ss]);
                    real bu, bv, bn;
                    phx_rhs_scalar_vjp<real>(ph, y3u[s], y3v[s], RES ? y3n[s] : real(0), POT ? vv[s] * a.vamp_h[2] : real(0), c3u[s],
                                             c3v[s], RES ? c3n[s] : real(0), Ku, Kv, bu, bv, bn);
                    au[s] += bu;
                    av[s] += bv;
                    c2u[s] = phx_fma(dth, bu, d3u[s]);
                    c2v[s] = phx_fma(dth, bv, d3v[s]);
                    if (RES) {
                        an[s] += bn;
                        c2n[s] = phx_fma(dth, bn, d3n[s]);
                    }
                    real gp2, gv2;
                    phx_rhs_scalar_param_vjp<real>(y2u[s], y2v[s], c2u[s], c2v[s], RES ? c2n[s] : real(0), a.pamp[1], a.vamp_h[1],
                                                   gp2, gv2);
                    gpa[s] += gp2;
                    gva[s] += gv2;
                }
            }
            /* ---- B2 at e = it-5: Yb2 = J(Y_2)^T c2, c1 */
            {
                const int e = it - 5, s = slot(j, -5), sn = slot(j, -6), ss = slot(j, -4);
                {
                    const real Ku = nbr(c2u[s], c2u[sn] + c2u[ss]);
                    const real Kv = nbr(c2v[s], c2v[sn] + c2v[ss]);
                    real bu, bv, bn;
                    phx_rhs_scalar_vjp<real>(ph, y2u[s], y2v[s], RES ? y2n[s] : real(0), POT ? vv[s] * a.vamp_h[1] : real(0), c2u[s],
                                             c2v[s], RES ? c2n[s] : real(0), Ku, Kv, bu, bv, bn);
                    au[s] += bu;
                    av[s] += bv;
                    c1u[s] = phx_fma(dth, bu, dt6 * lu[s]);
                    c1v[s] = phx_fma(dth, bv, dt6 * lv[s]);
                    if (RES) {
                        an[s] += bn;
                        c1n[s] = phx_fma(dth, bn, dt6 * ln[s]);
                    }
                    real gp1, gv1;
                    phx_rhs_scalar_param_vjp<real>(xu[s], xv[s], c1u[s], c1v[s], RES ? c1n[s] : real(0), a.pamp[0], a.vamp_h[0], gp1,
                                                   gv1);
                    gpa[s] += gp1;
                    gva[s] += gv1;
                }
            }
            /* ---- B1 at e = it-6: Yb1 = J(x)^T c1, l = l' + sum Yb; write row e */
            {
                const int e = it - 6, s = slot(j, -6), sn = slot(j, -7), ss = slot(j, -5);
                if (e >= 6 && e <= E - 7) {
                    const real Ku = nbr(c1u[s], c1u[sn] + c1u[ss]);
                    const real Kv = nbr(c1v[s], c1v[sn] + c1v[ss]);
                    real bu, bv, bn;
                    phx_rhs_scalar_vjp<real>(ph, xu[s], xv[s], RES ? xn[s] : real(0), POT ? vv[s] * a.vamp_h[0] : real(0), c1u[s],
                                             c1v[s], RES ? c1n[s] : real(0), Ku, Kv, bu, bv, bn);
                    const int grow = y0 + e - kAdjHalo;
                    if (out_lane && grow < Nr) {
                        const size_t idx = (size_t)grow * (size_t)Nc + (size_t)(x0 - kAdjHalo + lane);
                        v2 o;
                        o.x = lu[s] + (au[s] + bu);
                        o.y = lv[s] + (av[s] + bv);
                        a.lam_psi_out[so + idx] = o;
                        if (RES) a.lam_n_out[so + idx] = ln[s] + (an[s] + bn);
                        if (RES && a.g_pump) a.g_pump[so + idx] += gpa[s];
                        if (a.g_pot) a.g_pot[so + idx] += gva[s];
                    }
                }
            }
        }
    }
}

template <typename real, bool RES, bool PUMP, bool POT>
static hipError_t launch_adj(const Rk4AdjArgs<real>& a, hipStream_t s) {
    const int nblocks = (a.ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
    if (a.ph.kin_yx == real(1))
        hipLaunchKernelGGL((rk4_adjoint_fused<real, true, RES, PUMP, POT>), dim3(nblocks), dim3(kWave * kWavesPerBlock), 0, s, a);
    else
        hipLaunchKernelGGL((rk4_adjoint_fused<real, false, RES, PUMP, POT>), dim3(nblocks), dim3(kWave * kWavesPerBlock), 0, s, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_rk4_adjoint(const Rk4AdjArgs<real>& a, bool res, bool pump, bool pot, hipStream_t s) {
    if (!res) return pot ? launch_adj<real, false, false, true>(a, s) : launch_adj<real, false, false, false>(a, s);
    if (pump) return pot ? launch_adj<real, true, true, true>(a, s) : launch_adj<real, true, true, false>(a, s);
    return pot ? launch_adj<real, true, false, true>(a, s) : launch_adj<real, true, false, false>(a, s);
}
template hipError_t launch_rk4_adjoint<float>(const Rk4AdjArgs<float>&, bool, bool, bool, hipStream_t);
template hipError_t launch_rk4_adjoint<double>(const Rk4AdjArgs<double>&, bool, bool, bool, hipStream_t);

template <typename real>
__global__ __launch_bounds__(256) void sum_items(const real* __restrict__ in, real* __restrict__ out, size_t N, int items) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    real acc = in[i];
    for (int b = 1; b < items; b++) acc += in[(size_t)b * N + i];
    out[i] = acc;
}

template <typename real>
hipError_t launch_sum_items(const real* in, real* out, size_t N, int items, hipStream_t s) {
    const unsigned blocks = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL((sum_items<real>), dim3(blocks), dim3(256), 0, s, in, out, N, items);
    return hipGetLastError();
}
template hipError_t launch_sum_items<float>(const float*, float*, size_t, int, hipStream_t);
template hipError_t launch_sum_items<double>(const double*, double*, size_t, int, hipStream_t);

} // namespace phx

/* ======================================================================== host */

namespace {
int afail(int code, const std::string& m) { return phx::set_error(code, m); }
#define ADJ_HIP(call)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return afail(PHX_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)
} // namespace

struct phx_adjoint {
    int precision = PHX_FP32;
    size_t rsize = 4;
    phx_physics ph{};
    int Nc = 0, Nr = 0, B = 1, px = 1, py = 1;
    bool res = true, pump = true, pot = false, shared = true;
    double dt = 0;
    int64_t steps = 0;
    int K = 50;
    int64_t nseg = 0;
    size_t N = 0;      /* cells per grid */
    size_t sbytes = 0; /* bytes of one batched state: B grids of psi, then B grids of n */
    void* ckpt = nullptr;              /* nseg states: x at steps 0, K, 2K, ... */
    void* seg[2] = {nullptr, nullptr}; /* K-1 states each: x at steps jK+1 .. jK+K-1 of segment j in seg[j % 2] */
    int64_t seg_content[2] = {-1, -1}; /* segment held by each buffer (valid in stream order) */
    std::vector<hipEvent_t> ev_rec, ev_adj; /* per segment: recompute done (aux), adjoint done (main) */
    hipEvent_t ev_start = nullptr;
    hipStream_t aux = nullptr; /* recomputes the next segment while the main stream runs the adjoint */
    void* lam[2] = {nullptr, nullptr};
    void* gp = nullptr; /* B grids each */
    void* gv = nullptr;
    const void* P = nullptr; /* parameters of the last forward call (caller-owned) */
    const void* V = nullptr;
    bool have_forward = false;
    int ty_fwd = 12, ty_adj = 32;
    bool tuned = false;
    hipStream_t own = nullptr, stream = nullptr;

    char* state_psi(void* st) const { return (char*)st; }
    char* state_n(void* st) const { return (char*)st + B * N * 2 * rsize; }
    void* ckpt_at(int64_t j) const { return (char*)ckpt + j * sbytes; }
    void* seg_at(int64_t j, int64_t i) const { return (char*)seg[j & 1] + (i - 1) * sbytes; } /* segment j, i = 1 .. K-1 */
};

namespace {

template <typename real>
PhxPhys<real> adj_phys(const phx_adjoint* a) {
    phx::SystemConfig c;
    c.dx = a->ph.dx;
    c.dy = a->ph.dy;
    c.dV = a->ph.dx * a->ph.dy;
    c.one_over_dx2 = 1.0 / (c.dx * c.dx);
    c.one_over_dy2 = 1.0 / (c.dy * c.dy);
    c.m2_over_dx2_p_dy2 = -2.0 * (c.one_over_dx2 + c.one_over_dy2);
    c.m_eff_scaled = a->ph.m_eff_scaled;
    c.one_over_h_bar_s = a->ph.one_over_h_bar_s;
    c.gamma_c = a->ph.gamma_c;
    c.gamma_r = a->ph.gamma_r;
    c.g_c = a->ph.g_c;
    c.g_r = a->ph.g_r;
    c.R = a->ph.R;
    c.g_pm = a->ph.g_pm;
    c.delta_LT = a->ph.delta_LT;
    return phx::make_phys<real>(c);
}

/* one forward RK4 step of all items: state (psi_in, n_in) -> (psi_out, n_out) */
template <typename real>
int fwd_step(phx_adjoint* a, const void* psi_in, const void* n_in, void* psi_out, void* n_out, int ty, hipStream_t st = nullptr) {
    using v2 = typename phx::Vec2<real>::type;
    phx::Rk4BatchArgs<real> b{};
    b.entries = nullptr;
    b.tiles_per_item = phx::rk4_tiles(a->Nc, a->Nr, ty);
    b.ntiles = b.tiles_per_item * a->B;
    b.item_stride = a->N;
    b.param_stride = a->shared ? 0 : a->N;
    auto& c = b.common;
    c.psi_in = (const v2*)psi_in;
    c.psi_out = (v2*)psi_out;
    c.n_in = a->res ? (const real*)n_in : nullptr;
    c.n_out = a->res ? (real*)n_out : nullptr;
    c.pump = a->pump ? (const real*)a->P : nullptr;
    c.pot = a->pot ? (const real*)a->V : nullptr;
    c.Nc = a->Nc;
    c.Nr = a->Nr;
    c.periodic_x = a->px;
    c.periodic_y = a->py;
    c.dt = (real)a->dt;
    for (int k = 0; k < 4; k++) {
        c.pamp[k] = 1;
        c.vamp_h[k] = (real)a->ph.one_over_h_bar_s;
    }
    c.ph = adj_phys<real>(a);
    b.iso = c.ph.kin_yx == real(1);
    ADJ_HIP(phx::launch_rk4_batched<real>(b, ty, a->res, a->pump, a->pot, st ? st : a->stream));
    return PHX_OK;
}

/* state location of x_m (m < steps): a checkpoint, or a slot of its segment's buffer */
void* fwd_state(phx_adjoint* a, int64_t m) { return m % a->K == 0 ? a->ckpt_at(m / a->K) : a->seg_at(m / a->K, m % a->K); }

/* states jK+1 .. jK+len-1 of segment j from its checkpoint, on the aux stream */
template <typename real>
int recompute(phx_adjoint* a, int64_t j) {
    const int len = (int)std::min<int64_t>(a->K, a->steps - j * a->K);
    for (int i = 1; i < len; i++) {
        void* in = i == 1 ? a->ckpt_at(j) : a->seg_at(j, i - 1);
        void* out = a->seg_at(j, i);
        if (int r = fwd_step<real>(a, a->state_psi(in), a->state_n(in), a->state_psi(out), a->state_n(out), a->ty_fwd, a->aux)) return r;
    }
    ADJ_HIP(hipEventRecord(a->ev_rec[j], a->aux));
    a->seg_content[j & 1] = j;
    return PHX_OK;
}

template <typename real>
int adj_step(phx_adjoint* a, void* state, int lin, int ty) {
    using v2 = typename phx::Vec2<real>::type;
    phx::Rk4AdjArgs<real> g{};
    g.psi = (const v2*)a->state_psi(state);
    g.n = a->res ? (const real*)a->state_n(state) : nullptr;
    g.lam_psi_in = (const v2*)a->state_psi(a->lam[lin]);
    g.lam_n_in = a->res ? (const real*)a->state_n(a->lam[lin]) : nullptr;
    g.lam_psi_out = (v2*)a->state_psi(a->lam[lin ^ 1]);
    g.lam_n_out = a->res ? (real*)a->state_n(a->lam[lin ^ 1]) : nullptr;
    g.pump = a->pump ? (const real*)a->P : nullptr;
    g.pot = a->pot ? (const real*)a->V : nullptr;
    g.g_pump = a->res ? (real*)a->gp : nullptr;
    g.g_pot = (real*)a->gv;
    g.Nc = a->Nc;
    g.Nr = a->Nr;
    g.periodic_x = a->px;
    g.periodic_y = a->py;
    g.ty = ty;
    g.nstrips = phx::adj_strips(a->Nc);
    g.tiles_per_item = phx::adj_tiles(a->Nc, a->Nr, ty);
    g.ntiles = g.tiles_per_item * a->B;
    g.item_stride = a->N;
    g.param_stride = a->shared ? 0 : a->N;
    g.dt = (real)a->dt;
    for (int k = 0; k < 4; k++) {
        g.pamp[k] = 1;
        g.vamp_h[k] = (real)a->ph.one_over_h_bar_s;
    }
    g.ph = adj_phys<real>(a);
    ADJ_HIP(phx::launch_rk4_adjoint<real>(g, a->res, a->pump, a->pot, a->stream));
    return PHX_OK;
}

/* Tile heights of both kernels, timed on the first backward call.  The timed
   launches read the current state and cotangent and write the other buffers,
   so nothing advances; the gradient accumulators are cleared afterwards.
   Three round-robin passes over the candidates, the fastest pass of each
   counting: one pass in a fixed order timed the first candidates while the
   clocks were still ramping (after a long idle or a different workload) and
   could keep a tile height 40 % slower than the best. */
template <typename real>
int autotune(phx_adjoint* a) {
    hipEvent_t e0, e1;
    ADJ_HIP(hipEventCreate(&e0));
    ADJ_HIP(hipEventCreate(&e1));
    auto timeit = [&](auto&& launch, float& ms) -> int {
        for (int w = 0; w < 2; w++)
            if (int r = launch()) return r;
        ADJ_HIP(hipEventRecord(e0, a->stream));
        for (int w = 0; w < 4; w++)
            if (int r = launch()) return r;
        ADJ_HIP(hipEventRecord(e1, a->stream));
        ADJ_HIP(hipEventSynchronize(e1));
        ADJ_HIP(hipEventElapsedTime(&ms, e0, e1));
        return PHX_OK;
    };
    const int fwd_tys[] = {8, 10, 12, 16, 20};
    const int adj_tys[] = {16, 20, 24, 28, 32, 40, 48, 56, 64, 80};
    float fwd_ms[5], adj_ms[10];
    for (float& m : fwd_ms) m = 1e30f;
    for (float& m : adj_ms) m = 1e30f;
    int rc = PHX_OK;
    for (int pass = 0; pass < 3 && !rc; pass++) {
        for (int i = 0; i < 5 && !rc; i++) {
            float ms = 0;
            rc = timeit([&] { return fwd_step<real>(a, a->state_psi(a->ckpt_at(0)), a->state_n(a->ckpt_at(0)), a->state_psi(a->lam[1]),
                                                    a->state_n(a->lam[1]), fwd_tys[i]); },
                        ms);
            if (!rc) fwd_ms[i] = std::min(fwd_ms[i], ms);
        }
        for (int i = 0; i < 10 && !rc; i++) {
            float ms = 0;
            rc = timeit([&] { return adj_step<real>(a, a->ckpt_at(0), 0, adj_tys[i]); }, ms);
            if (!rc) adj_ms[i] = std::min(adj_ms[i], ms);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    a->ty_fwd = fwd_tys[std::min_element(fwd_ms, fwd_ms + 5) - fwd_ms];
    a->ty_adj = adj_tys[std::min_element(adj_ms, adj_ms + 10) - adj_ms];
    a->tuned = true;
    return PHX_OK;
}

template <typename real>
int forward_t(phx_adjoint* a, const void* psi0, const void* n0, const void* P, const void* V, void* psiT, void* nT) {
    const size_t cb = a->B * a->N * 2 * a->rsize, rb = a->B * a->N * a->rsize;
    a->P = P;
    a->V = V;
    void* x0 = a->ckpt_at(0);
    ADJ_HIP(hipMemcpyAsync(a->state_psi(x0), psi0, cb, hipMemcpyDeviceToDevice, a->stream));
    if (a->res) ADJ_HIP(hipMemcpyAsync(a->state_n(x0), n0, rb, hipMemcpyDeviceToDevice, a->stream));
    for (int64_t k = 0; k < a->steps; k++) {
        void* in = fwd_state(a, k);
        const bool last = k + 1 == a->steps;
        void* out = last ? nullptr : fwd_state(a, k + 1);
        if (int r = fwd_step<real>(a, a->state_psi(in), a->state_n(in), last ? psiT : a->state_psi(out), last ? nT : a->state_n(out),
                                   a->ty_fwd))
            return r;
    }
    if (a->steps == 0) {
        ADJ_HIP(hipMemcpyAsync(psiT, psi0, cb, hipMemcpyDeviceToDevice, a->stream));
        if (a->res) ADJ_HIP(hipMemcpyAsync(nT, n0, rb, hipMemcpyDeviceToDevice, a->stream));
    }
    /* the last two segments' states are still in the two buffers */
    a->seg_content[(a->nseg - 1) & 1] = a->nseg - 1;
    a->seg_content[a->nseg & 1] = a->nseg >= 2 ? a->nseg - 2 : -1;
    a->have_forward = true;
    return PHX_OK;
}

template <typename real>
int backward_t(phx_adjoint* a, const void* gpsiT, const void* gnT, void* gP, void* gV, void* gpsi0, void* gn0) {
    const size_t cb = a->B * a->N * 2 * a->rsize, rb = a->B * a->N * a->rsize;
    ADJ_HIP(hipMemcpyAsync(a->state_psi(a->lam[0]), gpsiT, cb, hipMemcpyDeviceToDevice, a->stream));
    if (a->res) {
        if (gnT)
            ADJ_HIP(hipMemcpyAsync(a->state_n(a->lam[0]), gnT, rb, hipMemcpyDeviceToDevice, a->stream));
        else
            ADJ_HIP(hipMemsetAsync(a->state_n(a->lam[0]), 0, rb, a->stream));
    }
    if (!a->tuned)
        if (int r = autotune<real>(a)) return r;
    if (const char* e = std::getenv("PHX_ADJ_TY")) a->ty_adj = std::max(1, std::atoi(e)); /* tuning experiments */
    if (const char* e = std::getenv("PHX_ADJ_FWD_TY")) a->ty_fwd = std::max(8, std::atoi(e));
    ADJ_HIP(hipMemsetAsync(a->gp, 0, rb, a->stream));
    ADJ_HIP(hipMemsetAsync(a->gv, 0, rb, a->stream));
    /* Segment j's states live in seg[j % 2].  While the main stream runs the adjoint
       of segment j, the aux stream recomputes segment j-1 into the other buffer, after
       the adjoint of segment j+1 (the buffer's previous reader) has finished. */
    ADJ_HIP(hipEventRecord(a->ev_start, a->stream));
    ADJ_HIP(hipStreamWaitEvent(a->aux, a->ev_start, 0));
    std::vector<char> pending(a->nseg, 0); /* recomputed on aux in this call: main must wait */
    const int64_t last = a->nseg - 1;
    if (a->K >= 2 && a->seg_content[last & 1] != last) {
        if (int r = recompute<real>(a, last)) return r;
        pending[last] = 1;
    }
    int lin = 0;
    for (int64_t j = last; j >= 0; j--) {
        if (a->K >= 2 && j >= 1 && a->seg_content[(j - 1) & 1] != j - 1) {
            if (j + 1 <= last) ADJ_HIP(hipStreamWaitEvent(a->aux, a->ev_adj[j + 1], 0));
            if (int r = recompute<real>(a, j - 1)) return r;
            pending[j - 1] = 1;
        }
        if (pending[j]) ADJ_HIP(hipStreamWaitEvent(a->stream, a->ev_rec[j], 0));
        const int len = (int)std::min<int64_t>(a->K, a->steps - j * a->K);
        for (int i = len - 1; i >= 0; i--) {
            void* x = i == 0 ? a->ckpt_at(j) : a->seg_at(j, i);
            if (int r = adj_step<real>(a, x, lin, a->ty_adj)) return r;
            lin ^= 1;
        }
        ADJ_HIP(hipEventRecord(a->ev_adj[j], a->stream));
    }
    if (gpsi0) ADJ_HIP(hipMemcpyAsync(gpsi0, a->state_psi(a->lam[lin]), cb, hipMemcpyDeviceToDevice, a->stream));
    if (gn0 && a->res) ADJ_HIP(hipMemcpyAsync(gn0, a->state_n(a->lam[lin]), rb, hipMemcpyDeviceToDevice, a->stream));
    if (a->shared) {
        if (gP && a->res) ADJ_HIP(phx::launch_sum_items<real>((const real*)a->gp, (real*)gP, a->N, a->B, a->stream));
        if (gV) ADJ_HIP(phx::launch_sum_items<real>((const real*)a->gv, (real*)gV, a->N, a->B, a->stream));
    } else {
        if (gP && a->res) ADJ_HIP(hipMemcpyAsync(gP, a->gp, rb, hipMemcpyDeviceToDevice, a->stream));
        if (gV) ADJ_HIP(hipMemcpyAsync(gV, a->gv, rb, hipMemcpyDeviceToDevice, a->stream));
    }
    if (gP && !a->res) ADJ_HIP(hipMemsetAsync(gP, 0, (a->shared ? 1 : a->B) * a->N * a->rsize, a->stream));
    return PHX_OK;
}

void release(phx_adjoint* a) {
    for (void* p : {a->ckpt, a->seg[0], a->seg[1], a->lam[0], a->lam[1], a->gp, a->gv})
        if (p) (void)hipFree(p);
    for (auto* v : {&a->ev_rec, &a->ev_adj})
        for (hipEvent_t e : *v)
            if (e) (void)hipEventDestroy(e);
    if (a->ev_start) (void)hipEventDestroy(a->ev_start);
    if (a->aux) (void)hipStreamDestroy(a->aux);
    if (a->own) (void)hipStreamDestroy(a->own);
    delete a;
}

} // namespace

extern "C" {

int phx_adjoint_create(phx_adjoint** out, int precision, const phx_physics* ph, int N_c, int N_r, int batch, int periodic_x,
                       int periodic_y, int flags, int shared_params, double dt, int64_t n_steps, int checkpoint_every) {
    if (!out || !ph || N_c <= 0 || N_r <= 0 || batch <= 0 || dt <= 0 || n_steps < 0 || checkpoint_every < 1)
        return afail(PHX_ERR_INVALID, "invalid argument");
    if (precision != PHX_FP32 && precision != PHX_FP64) return afail(PHX_ERR_INVALID, "precision");
    if ((flags & PHX_ENS_PUMP) && !(flags & PHX_ENS_RESERVOIR)) return afail(PHX_ERR_INVALID, "a pump needs a reservoir");
    auto* a = new phx_adjoint();
    a->precision = precision;
    a->rsize = precision == PHX_FP32 ? 4 : 8;
    a->ph = *ph;
    a->Nc = N_c;
    a->Nr = N_r;
    a->B = batch;
    a->px = periodic_x;
    a->py = periodic_y;
    a->res = flags & PHX_ENS_RESERVOIR;
    a->pump = flags & PHX_ENS_PUMP;
    a->pot = flags & PHX_ENS_POTENTIAL;
    a->shared = shared_params != 0;
    a->dt = dt;
    a->steps = n_steps;
    a->K = (int)std::min<int64_t>(checkpoint_every, std::max<int64_t>(n_steps, 1));
    a->nseg = std::max<int64_t>((n_steps + a->K - 1) / a->K, 1);
    a->N = (size_t)N_c * N_r;
    a->sbytes = (size_t)batch * a->N * 3 * a->rsize;
    const size_t rb = (size_t)batch * a->N * a->rsize;
    bool ok = hipMalloc(&a->ckpt, a->nseg * a->sbytes) == hipSuccess;
    ok = ok && (a->K < 2 || (hipMalloc(&a->seg[0], (a->K - 1) * a->sbytes) == hipSuccess &&
                             hipMalloc(&a->seg[1], (a->K - 1) * a->sbytes) == hipSuccess));
    ok = ok && hipMalloc(&a->lam[0], a->sbytes) == hipSuccess && hipMalloc(&a->lam[1], a->sbytes) == hipSuccess;
    ok = ok && hipMalloc(&a->gp, rb) == hipSuccess && hipMalloc(&a->gv, rb) == hipSuccess;
    ok = ok && hipStreamCreate(&a->own) == hipSuccess && hipStreamCreate(&a->aux) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&a->ev_start, hipEventDisableTiming) == hipSuccess;
    a->ev_rec.assign(a->nseg, nullptr);
    a->ev_adj.assign(a->nseg, nullptr);
    for (int64_t j = 0; ok && j < a->nseg; j++)
        ok = hipEventCreateWithFlags(&a->ev_rec[j], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&a->ev_adj[j], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        const size_t total = (a->nseg + 2 * a->K) * a->sbytes + 2 * rb;
        release(a);
        return afail(PHX_ERR_HIP, "device allocation of about " + std::to_string(total >> 20) + " MiB failed");
    }
    a->stream = a->own;
    *out = a;
    return PHX_OK;
}

int phx_adjoint_forward(phx_adjoint* a, const void* psi0, const void* n0, const void* pump, const void* pot, void* psi_T, void* n_T) {
    if (!a || !psi0 || !psi_T) return afail(PHX_ERR_INVALID, "null argument");
    if (a->res && (!n0 || !n_T)) return afail(PHX_ERR_INVALID, "the reservoir needs n0 and n_T");
    if (a->pump && !pump) return afail(PHX_ERR_INVALID, "pump expected");
    if (a->pot && !pot) return afail(PHX_ERR_INVALID, "potential expected");
    return a->precision == PHX_FP32 ? forward_t<float>(a, psi0, n0, pump, pot, psi_T, n_T)
                                    : forward_t<double>(a, psi0, n0, pump, pot, psi_T, n_T);
}

int phx_adjoint_backward(phx_adjoint* a, const void* g_psi_T, const void* g_n_T, void* g_pump, void* g_pot, void* g_psi0,
                         void* g_n0) {
    if (!a || !g_psi_T) return afail(PHX_ERR_INVALID, "null argument");
    if (!a->have_forward) return afail(PHX_ERR_STATE, "backward before forward");
    return a->precision == PHX_FP32 ? backward_t<float>(a, g_psi_T, g_n_T, g_pump, g_pot, g_psi0, g_n0)
                                    : backward_t<double>(a, g_psi_T, g_n_T, g_pump, g_pot, g_psi0, g_n0);
}

int phx_adjoint_set_stream(phx_adjoint* a, void* stream) {
    if (!a) return afail(PHX_ERR_INVALID, "null handle");
    a->stream = stream ? (hipStream_t)stream : a->own;
    return PHX_OK;
}

int phx_adjoint_sync(phx_adjoint* a) {
    if (!a) return afail(PHX_ERR_INVALID, "null handle");
    ADJ_HIP(hipStreamSynchronize(a->stream));
    return PHX_OK;
}

void phx_adjoint_destroy(phx_adjoint* a) {
    if (!a) return;
    (void)hipStreamSynchronize(a->stream);
    (void)hipStreamSynchronize(a->aux);
    release(a);
}

} /* extern "C" */
