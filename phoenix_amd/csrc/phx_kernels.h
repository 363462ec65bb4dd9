/*
 * phx_kernels.h -- host-side launch interface of the gfx950 kernels.
 * Device layout: full grid, row-major, psi interleaved complex (float2/double2),
 * n, P, V real.  No padding: boundaries are resolved inside the kernels.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>

#include "custom_rhs.h"
#include "phx_tableaux.h"

namespace phx {

template <typename real>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

template <typename real>
struct Rk4Args {
    const typename Vec2<real>::type* psi_in;
    const real* n_in; /* nullptr: no reservoir */
    typename Vec2<real>::type* psi_out;
    real* n_out;
    const real* pump; /* nullptr: no pump */
    const real* pot;  /* nullptr: no potential */
    int Nc, Nr;
    int periodic_x, periodic_y;
    int nstrips, nbands; /* set by the launcher */
    int alt_rows;        /* 1: odd bands of tiles stream their rows bottom to top (same results) */
    real dt;
    real pamp[4];   /* pump temporal amplitude per stage: a(t)(1-w)+w a(t+dt) */
    real vamp_h[4]; /* potential temporal amplitude per stage, divided by hbar */
    PhxPhys<real> ph;
};

/* Row order of the fused RK4 tiles (Rk4Args::alt_rows): alternating by band unless PHX_RK4_ALT=0. */
inline int rk4_alt_rows() {
    const char* e = std::getenv("PHX_RK4_ALT");
    return e && std::atoi(e) == 0 ? 0 : 1;
}

/* Launch shape of the fused RK4 kernel.
   ty: output rows per wavefront tile (8, 10, 12, 16, 20, 24);
   ew, pf: kept for the tuning interface; the kernel uses DPP wave shifts for
   east/west and loads rows 2 ahead of first use. */
struct Rk4Tune {
    int ty = 16, ew = 1, pf = 2;
};

template <typename real>
hipError_t launch_rk4_scalar(const Rk4Args<real>& a, const Rk4Tune& t, hipStream_t s);

/* Many RK4 steps in one launch (rk4_scalar_flow): one step counter per wavefront tile in
   device memory (kFlowCounterStride ints apart), all equal to `base` when the launch starts
   and base + steps after it;
   *error is set (and the state undefined) if a tile waited more than spin_limit polls for
   a neighbour.  Tiles: nstrips x nbands = ceil(Nc / 56) x ceil(Nr / ty). */
constexpr int kFlowCounterStride = 32; /* ints: one 128-byte line per tile counter */
struct Rk4Flow {
    int* counters; /* tile t's counter at counters[t * kFlowCounterStride] */
    int* error;
    int base;
    int steps;
    int spin_limit;
};
template <typename real>
hipError_t launch_rk4_flow(const Rk4Args<real>& a, const Rk4Tune& t, const Rk4Flow& f, hipStream_t s);

/* One instance of an ensemble: its buffers, its size, and its first wavefront tile
   in the flattened tile list.  The kinetic weights carry the instance's dx, dy. */
template <typename real>
struct Rk4BatchEntry {
    const typename Vec2<real>::type* psi_in;
    const real* n_in;
    typename Vec2<real>::type* psi_out;
    real* n_out;
    const real* pump;
    const real* pot;
    int Nc, Nr, nstrips, tile0;
    real kin_c, kin_x, kin_y, kin_yx;
};

template <typename real>
struct Rk4BatchArgs {
    const Rk4BatchEntry<real>* entries; /* device array, one per instance; nullptr: uniform batch */
    const int* tile_inst;               /* instance of each wavefront tile */
    int ntiles;
    Rk4Args<real> common;               /* dt, amplitudes, physics, boundary flags */
    /* uniform batch (entries == nullptr): grids of size common.Nc x common.Nr whose
       fields start at common.<field> + item * item_stride (pump / pot: param_stride,
       0 = shared by all items); ntiles = items * tiles_per_item */
    int tiles_per_item;
    size_t item_stride, param_stride;
    int iso; /* every instance has dx == dy (kin_yx == 1): stencil without the y/x ratio */
};

/* All instances share the feature set (reservoir / pump / potential) and dt. */
template <typename real>
hipError_t launch_rk4_batched(const Rk4BatchArgs<real>& b, int ty, bool res, bool pump, bool pot, hipStream_t s);

/* wavefront tiles of one grid for the fused RK4 kernel with tile height ty */
__host__ __device__ inline int rk4_tiles(int Nc, int Nr, int ty) { return ((Nc + 55) / 56) * ((Nr + ty - 1) / ty); }
__host__ __device__ inline int rk4_strips(int Nc) { return (Nc + 55) / 56; }

/* Discrete adjoint of one RK4 step of the scalar model (phx_adjoint.hip), for
   `ntiles / tiles_per_item` grids of one size at a fixed element stride.
   lam_*_in is the cotangent of the step's output state, lam_*_out receives the
   cotangent of its input state; g_pump / g_pot (one grid per item) accumulate
   the cotangents of P and V. */
template <typename real>
struct Rk4AdjArgs {
    const typename Vec2<real>::type* psi; /* input state of the step */
    const real* n;
    const typename Vec2<real>::type* lam_psi_in;
    const real* lam_n_in;
    typename Vec2<real>::type* lam_psi_out;
    real* lam_n_out;
    const real* pump; /* nullptr: no pump */
    const real* pot;  /* nullptr: V = 0 (its cotangent is still accumulated) */
    real* g_pump;     /* nullptr: not accumulated */
    real* g_pot;
    int Nc, Nr;
    int periodic_x, periodic_y;
    int ty, nstrips, tiles_per_item, ntiles; /* ty: output rows per wavefront tile */
    size_t item_stride, param_stride;       /* elements; param_stride 0: P, V shared */
    real dt;
    real pamp[4], vamp_h[4];
    PhxPhys<real> ph;
};

/* halo of the adjoint tile: the state is needed 6 cells beyond the output */
constexpr int kAdjHalo = 6;
__host__ __device__ inline int adj_strips(int Nc) { return (Nc + 51) / 52; }
__host__ __device__ inline int adj_tiles(int Nc, int Nr, int ty) { return adj_strips(Nc) * ((Nr + ty - 1) / ty); }

template <typename real>
hipError_t launch_rk4_adjoint(const Rk4AdjArgs<real>& a, bool res, bool pump, bool pot, hipStream_t s);

/* out[i] = sum_b in[b * N + i], b = 0 .. items-1 in order */
template <typename real>
hipError_t launch_sum_items(const real* in, real* out, size_t N, int items, hipStream_t s);

/* Arguments of the generic fused Runge-Kutta kernel (phx_rk_generic.hip).
   Index 0 = plus component, 1 = minus component (TE-TM only). */
template <typename real>
struct RkGenArgs {
    const typename Vec2<real>::type* psi_in[2];
    const real* n_in[2]; /* nullptr: no reservoir */
    typename Vec2<real>::type* psi_out[2];
    real* n_out[2];
    const real* pump[2]; /* nullptr: no pump */
    const real* pot[2];  /* nullptr: no potential */
    int Nc, Nr;
    int periodic_x, periodic_y;
    int nstrips, nbands; /* set by the launcher */
    int tetm, stochastic;
    int imag; /* imaginary time: psi advances by -i dt */
    real dt;
    /* temporal amplitudes at t and t+dt; a(w) = a(t) + w (a(t+dt) - a(t)) */
    real amp_pump, amp_next_pump, amp_pot_h, amp_next_pot_h;
    /* noise field of this step (launch_noise_field), read with the state; nullptr without noise */
    const typename Vec2<real>::type* noise;
    double* partials; /* 2 doubles per wavefront (error, norm), adaptive tableaux only */
    double* sums;     /* 2 doubles: reduced error and norm */
    PhxPhys<real> ph;
};

template <typename real>
hipError_t launch_rk_generic(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s);
/* groups of tableaux instantiated in their own translation units (1-3, 3-4 and 6 stages) */
template <typename real>
hipError_t launch_rk_tab3(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s);
template <typename real>
hipError_t launch_rk_tab4(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s);
template <typename real>
hipError_t launch_rk_tab6(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s);
template <typename real>
hipError_t launch_rk_tab6f(const RkGenArgs<real>& a, RkTableau tab, hipStream_t s);
/* adaptive DP45, TE-TM, FP32, tile height 32 (phx_rk_dp45t.hip) */
hipError_t launch_dp45_tetm_tall(const RkGenArgs<float>& a, int ty, hipStream_t s);
/* tile height of the adaptive DP45 kernel for a grid: the height among the compiled ones that
   minimises (wavefront tiles per SIMD, rounded up) x (row evaluations per tile), a VALU-bound
   model; PHX_GEN_TY fixes it (16 or 32) */
int dp45_adaptive_tile_rows(bool tetm, bool fp32, int Nc, int Nr);

/* Staged Runge-Kutta step (phx_staged.hip): one launch per CALCULATE_K, INTERMEDIATE_SUM_K
   and FINAL_SUM_K / ERROR_K of the reference's sequence, K's in HBM on a grid padded by H
   cells.  For tableaux with more stages than the fused kernel holds, and for sources the
   fused kernels do not take: pump, potential and pulse chains of several temporal groups,
   complex temporal envelopes of the potential and the pulse. */
constexpr int kStagedMaxGroups = 8;
enum class StagedOp { K, SUM, FINAL };
template <typename real>
struct StagedArgs {
    using v2 = typename Vec2<real>::type;
    /* step-start state and output (full grid, row-major); x_n[0] == nullptr: no reservoir.  With
       cres the reservoir arrays (x_n, out_n, y_n, k_n) hold complex {re, im} cells. */
    const v2* x_psi[2];
    const real* x_n[2];
    v2* out_psi[2];
    real* out_n[2];
    /* stage input and the K slots (padded grid; slot s at s * padN) */
    v2* y_psi[2];
    real* y_n[2];
    v2* k_psi[2];
    real* k_n[2];
    /* sources: n groups of full-grid fields per component, group g at g * N */
    const real* pump[2];
    const real* pot[2];
    const v2* pulse[2];
    int npump, npot, npulse;
    /* this stage's temporal amplitudes a_g(w) = a_g(t)(1 - w) + w a_g(t + dt) */
    real pump_amp[kStagedMaxGroups];
    real pump_amp_im[kStagedMaxGroups]; /* imaginary parts (complex reservoir only) */
    v2 pot_amp[kStagedMaxGroups];
    v2 pulse_amp[kStagedMaxGroups];
    const v2* noise; /* this step's noise field (full grid) or nullptr */
    int stochastic, tetm;
    int cres; /* complex reservoir */
    int Nc, Nr, H, P; /* P = Nc + 2H */
    size_t N, padN;   /* Nc Nr, (Nr + 2H) P */
    int periodic_x, periodic_y;
    int stage; /* K: slot written (0-based) */
    int ext;   /* K, SUM: padded cells within ext of the grid */
    /* SUM / FINAL: weights w[j] of slots slot[j]; wsum = their sum (noise increment) */
    int nw;
    int slot[kStagedMaxStages];
    real w[kStagedMaxStages];
    real wsum;
    /* FINAL: error weights e[j] of slots eslot[j] (ne = 0: no error) */
    int ne;
    int eslot[kStagedMaxStages];
    real e[kStagedMaxStages];
    real dt, dtr, dti; /* psi advances by (dtr + i dti) sum; a real reservoir by dt, a complex one like psi */
    double* partials;  /* FINAL with ne > 0: (error, norm) per block */
    PhxPhys<real> ph;
};
template <typename real>
hipError_t launch_staged(const StagedArgs<real>& a, StagedOp op, hipStream_t s);
/* deterministic reduction of (error, norm) pairs into sums[0..1] (one workgroup, fixed order) */
hipError_t launch_reduce_partials(const double* partials, int nwaves, double* sums, hipStream_t s);
/* blocks of the FINAL launch (partials buffer size = 2 * this) */
inline size_t staged_final_blocks(int Nc, int Nr) { return (size_t)((Nc + 63) / 64) * (size_t)((Nr + 3) / 4); }

/* Imaginary-time normalisation (Solver::normalizeImaginaryTimePropagation): psi_k *= sqrt(amp /
   (S_k dV)) with S_k = sum over the grid of |psi_k|^2 (FP64, fixed order; S_k < 1e-10 -> 1), for
   psi_plus and, if not null, psi_minus.  partials: 2 doubles per 256-cell block, sums: 2 doubles;
   no host synchronisation. */
template <typename real>
hipError_t launch_imag_normalize(typename Vec2<real>::type* psi_plus, typename Vec2<real>::type* psi_minus, size_t cells,
                                 double amp, double dV, double* partials, double* sums, hipStream_t s);
inline size_t imag_norm_blocks(size_t cells) { return (cells + 255) / 256; }

/* noise field of one step: xi[cell] = amp * N2(seed, step, cell) over the whole grid (row-major
   cell index), N2 the Philox4x32-10 + Box-Muller pair of phx_noise.h; amp = stochastic amplitude
   * sqrt(dt).  The fused kernels read it like a field (0 outside a zero-BC domain). */
template <typename real>
hipError_t launch_noise_field(typename Vec2<real>::type* xi, int Nc, int Nr, uint64_t seed, uint64_t step, real amp,
                              hipStream_t s);

/* wavefront tiles of the generic kernel for a grid at tile height ty; the partials buffer of an
   adaptive tableau holds 2 doubles per tile at the smallest height the launcher uses (16) */
constexpr int kRkGenMinTileRows = 16;
inline int rk_generic_waves(RkTableau tab, int Nc, int Nr, int ty = kRkGenMinTileRows) {
    const int S = rk_tableau_stages(tab);
    const int outw = 64 - 2 * S;
    return ((Nc + outw - 1) / outw) * ((Nr + ty - 1) / ty);
}

/* k-space filter psi <- IDFT(DFT(psi) / (N_c N_r) * fftshift(mask)), in place.
   tw_c / tw_r: forward twiddle tables exp(-2 pi i k / N) for k < N (complex, real
   precision) of the row length N_c and column length N_r. */
template <typename real>
hipError_t launch_fft_filter(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const real* mask, int Nc, int Nr,
                             hipStream_t s, typename Vec2<real>::type* spec = nullptr);
/* The filter above also stores the filtered spectrum DFT(psi) / (N_c N_r) * fftshift(mask) in spec
   (row-major, k = 0 first) when spec != nullptr.  launch_fft_forward: DFT(data) in place. */
template <typename real>
hipError_t launch_fft_forward(typename Vec2<real>::type* data, const void* tw_c, const void* tw_r, int Nc, int Nr, hipStream_t s);

/* Split-step linear propagator in place: psi <- IDFT(DFT(psi) * colf[c] * rowf[r]), with the
   separable complex factors rowf (N_r), colf (N_c) computed on the host (they carry 1/(N_c N_r)). */
template <typename real>
hipError_t launch_fft_propagate(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const void* rowf, const void* colf,
                                int Nc, int Nr, hipStream_t s);

/* Pointwise parts of the split-step Fourier iterator (phx_ssfm.hip).  Index 0 = plus, 1 = minus. */
template <typename real>
struct SsfmArgs {
    typename Vec2<real>::type* psi[2];
    real* n_cur[2]; /* reservoir of the step start (nullptr without reservoir) */
    real* n_nxt[2]; /* reservoir after the nonlinear step */
    /* sources as chains of temporal groups (a single field is one group), each group at its
       temporal amplitude of the step start: pump P_g Re a_g, potential V_g a_g (complex a_g),
       pulse F_g a_g (complex field and amplitude) */
    int npump, npot, npulse;
    const real* pump[2][kStagedMaxGroups];
    const real* pot[2][kStagedMaxGroups];
    const typename Vec2<real>::type* pulse[2][kStagedMaxGroups];
    real pamp[kStagedMaxGroups], pamp_im[kStagedMaxGroups]; /* pump amplitudes; Im only with cres */
    typename Vec2<real>::type vamp[kStagedMaxGroups], lamp[kStagedMaxGroups];
    const typename Vec2<real>::type* noise; /* this step's noise field (stochastic only) */
    size_t cells;
    int tetm, stochastic;
    int cres; /* complex reservoir: n_cur / n_nxt hold {re, im} cells */
    real dt;
    PhxPhys<real> ph;
};
template <typename real>
hipError_t launch_ssfm_nonlinear(const SsfmArgs<real>& a, hipStream_t s);
template <typename real>
hipError_t launch_ssfm_independent(const SsfmArgs<real>& a, hipStream_t s);

/* Power of two in [4, 4096] whose two LDS row buffers fit in 64 KiB. */
bool fft_supported(int N, size_t real_size);

} // namespace phx
