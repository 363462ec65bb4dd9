/*
 * phx_fft.hip -- the k-space low-pass filter of PHOENIX as three LDS FFT kernels.
 *
 * Replaces Solver::applyFFTFilter (solver_fft.cu:67-100): toFull, cuFFT forward,
 * kernel_mask_fft (F <- F / (N_c N_r) * mask, kernel_fft.cuh:34-39), cuFFT
 * inverse, toSubgrids.  Here, in place on the full-grid psi:
 *   1. fft_rows<forward>  : one row per workgroup, length-N_c FFT in LDS;
 *   2. fft_cols_filter    : CB adjacent columns per workgroup, length-N_r forward
 *                           FFT, multiply by the fftshift-ed mask / (N_c N_r),
 *                           inverse FFT -- all without leaving LDS;
 *   3. fft_rows<inverse>  : inverse row FFTs.
 * Each kernel reads and writes the grid once (6 passes over N_c N_r complex).
 * Lengths 256, 512 and 1024 on both axes take the four-step kernels instead (fft4_rows,
 * fft4_cols: N = N1 N2, sub-transforms in registers, one LDS transpose per transform, two
 * lanes per sub-sequence); other powers of two the Stockham kernels below, other even
 * lengths direct DFTs.
 *
 * The transforms are unnormalised DFTs with the forward sign e^{-2 pi i jk/N}
 * (cuFFT/FFTW convention).  Algorithm: Stockham autosort, radix 4 with a final
 * radix-2 pass when log2 N is odd (textbook form, e.g. Van Loan, "Computational
 * Frameworks for the FFT", 1992).  Twiddles come from a table computed on the
 * host in double precision.  The passes run in place in one LDS buffer (inputs of
 * a pass staged in registers), so a row of up to 8192 (FP32) / 4096 (FP64) points
 * fits in 64 KiB; column blocks hold 8 columns where they fit.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>

#include "custom_rhs.h"
#include "phx_device.h"
#include "phx_kernels.h"

namespace phx {

constexpr int kFftThreads = 256;

/* aligned to its full size, so an element moves with one 64-bit (FP32) / 128-bit (FP64) LDS access */
template <typename real>
struct alignas(2 * sizeof(real)) Cx {
    real r, i;
};

template <typename real>
__device__ __forceinline__ Cx<real> cmul(Cx<real> a, Cx<real> w) {
    return {phx_fma(a.r, w.r, -(a.i * w.i)), phx_fma(a.r, w.i, a.i * w.r)};
}

/* LDS slot of element i: one spare slot after every 32 elements (PAD), so that the strided
   Stockham writes of the early passes spread over all banks instead of a few. */
template <bool PAD>
__device__ __forceinline__ int lslot(int i) { return PAD ? i + (i >> 5) : i; }
/* LDS elements that hold n elements */
template <bool PAD>
__host__ __device__ constexpr size_t lds_elems(size_t n) { return PAD ? n + n / 32 + 1 : n; }

/* Butterflies a thread may own in one pass: the launchers pick KB, the smallest
   power of two with count * N / 4 <= KB * threads, up to 8 in FP32 and 4 in FP64,
   so the staged inputs stay in registers. */
template <typename real>
__host__ __device__ constexpr int max_bfly() { return sizeof(real) == 4 ? 8 : 4; }

/* Stockham autosort FFT of `count` independent length-N sequences stored back to
   back in one LDS buffer (sequence j at buf + j*N), in place: in every radix-4 pass
   each thread first reads the inputs of its butterflies into registers, the block
   synchronises, and then the outputs are written to their Stockham positions in the
   same buffer.  One buffer instead of two doubles the sequences a workgroup holds.
   sign = -1 forward, +1 inverse (conjugated twiddles). */
template <typename real, int SIGN, int KB, bool PAD>
__device__ void stockham(Cx<real>* buf, const Cx<real>* __restrict__ tw, int N, int log2N, int count) {
    const int nb = N >> 2; /* radix-4 butterflies per sequence */
    const int total = nb * count;
    int n = N, log2n = log2N, s = 1, log2s = 0;
    while (n >= 4) {
        const int m = n >> 2;
        const int stepshift = log2N - log2n; /* twiddle index p * N / n */
        Cx<real> in[KB][4], wt[KB][3];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int b = threadIdx.x + k * kFftThreads;
            if (b < total) {
                const int seq = b / nb, bb = b - seq * nb;
                const int p = bb >> log2s, q = bb & (s - 1);
                const int x = seq * N;
                in[k][0] = buf[lslot<PAD>(x + q + s * p)];
                in[k][1] = buf[lslot<PAD>(x + q + s * (p + m))];
                in[k][2] = buf[lslot<PAD>(x + q + s * (p + 2 * m))];
                in[k][3] = buf[lslot<PAD>(x + q + s * (p + 3 * m))];
                /* the pass's twiddles are read before the barrier, so their latency overlaps it */
                wt[k][0] = tw[p << stepshift];
                wt[k][1] = tw[(2 * p) << stepshift];
                wt[k][2] = tw[(3 * p) << stepshift];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int b = threadIdx.x + k * kFftThreads;
            if (b < total) {
                const int seq = b / nb, bb = b - seq * nb;
                const int p = bb >> log2s, q = bb & (s - 1);
                const int y = seq * N;
                const Cx<real> a = in[k][0], bv = in[k][1], c = in[k][2], d = in[k][3];
                Cx<real> w1 = wt[k][0], w2 = wt[k][1], w3 = wt[k][2];
                if (SIGN > 0) {
                    w1.i = -w1.i;
                    w2.i = -w2.i;
                    w3.i = -w3.i;
                }
                const Cx<real> apc = {a.r + c.r, a.i + c.i}, amc = {a.r - c.r, a.i - c.i};
                const Cx<real> bpd = {bv.r + d.r, bv.i + d.i};
                /* j(b - d) with j = -i (forward) or +i (inverse) */
                const Cx<real> jbmd = SIGN < 0 ? Cx<real>{bv.i - d.i, -(bv.r - d.r)} : Cx<real>{-(bv.i - d.i), bv.r - d.r};
                buf[lslot<PAD>(y + q + s * (4 * p))] = {apc.r + bpd.r, apc.i + bpd.i};
                buf[lslot<PAD>(y + q + s * (4 * p + 1))] = cmul(Cx<real>{amc.r + jbmd.r, amc.i + jbmd.i}, w1);
                buf[lslot<PAD>(y + q + s * (4 * p + 2))] = cmul(Cx<real>{apc.r - bpd.r, apc.i - bpd.i}, w2);
                buf[lslot<PAD>(y + q + s * (4 * p + 3))] = cmul(Cx<real>{amc.r - jbmd.r, amc.i - jbmd.i}, w3);
            }
        }
        __syncthreads();
        n = m;
        log2n -= 2;
        s <<= 2;
        log2s += 2;
    }
    if (n == 2) {
        /* final radix-2 pass: the butterfly (q, q + s) reads and writes the same two slots */
        for (int b = threadIdx.x; b < s * count; b += blockDim.x) {
            const int seq = b >> log2s;
            const int q = b & (s - 1);
            const int x = seq * N;
            const Cx<real> a = buf[lslot<PAD>(x + q)], bv = buf[lslot<PAD>(x + q + s)];
            buf[lslot<PAD>(x + q)] = {a.r + bv.r, a.i + bv.i};
            buf[lslot<PAD>(x + q + s)] = {a.r - bv.r, a.i - bv.i};
        }
        __syncthreads();
    }
}

template <typename real, int SIGN, int KB, bool PAD>
__global__ __launch_bounds__(kFftThreads) void fft_rows(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw, int N, int log2N) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<real>* A = reinterpret_cast<Cx<real>*>(lds_raw);
    typename Vec2<real>::type* row = data + (size_t)blockIdx.x * (size_t)N;
    /* N <= PT * threads (the launcher's KB): every load of the row is issued before the first
       LDS write; indices past the end read the last element and are not written */
    constexpr int PT = 4 * KB;
    typename Vec2<real>::type v[PT];
#pragma unroll
    for (int k = 0; k < PT; k++) v[k] = row[min((int)threadIdx.x + k * kFftThreads, N - 1)];
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int i = threadIdx.x + k * kFftThreads;
        if (i < N) A[lslot<PAD>(i)] = {v[k].x, v[k].y};
    }
    __syncthreads();
    stockham<real, SIGN, KB, PAD>(A, tw, N, log2N, 1);
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int i = threadIdx.x + k * kFftThreads;
        if (i < N) {
            typename Vec2<real>::type o;
            const Cx<real> f = A[lslot<PAD>(i)];
            o.x = f.r;
            o.y = f.i;
            row[i] = o;
        }
    }
}

/* Column groups are handed out XCD by XCD (block b runs on XCD b % 8): consecutive
   groups share cache lines, so they go to the same L2. */
__device__ __forceinline__ int fft_xcd_group(int b, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

/* What the column pass does in k-space. */
enum { kColFilter = 0, kColPropagate = 1, kColForward = 2 };

/* CB adjacent columns: forward FFT along the column, a pointwise factor in k-space, inverse FFT.
   MODE kColFilter:    F <- F / (N_c N_r) * mask[shifted] (the k-space filter); with spec != nullptr
                       the filtered spectrum is also stored there (row-major, k = 0 first), as the
                       reference keeps it in fft_plus / fft_minus (solver_fft.cu:67-100);
   MODE kColPropagate: F <- F * colf[c] * rowf[r] (the split-step propagator; the tables carry 1 / (N_c N_r));
   MODE kColForward:   the forward transform only (data becomes the unnormalised spectrum). */
template <typename real, int CB, int KB, int MODE, bool PAD>
__global__ __launch_bounds__(kFftThreads) void fft_cols_filter(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw,
                                                              const real* __restrict__ mask, const Cx<real>* __restrict__ rowf,
                                                              const Cx<real>* __restrict__ colf, int Nc, int Nr, int log2Nr,
                                                              real Ntot, typename Vec2<real>::type* spec) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<real>* A = reinterpret_cast<Cx<real>*>(lds_raw);
    const int c0 = fft_xcd_group(blockIdx.x, gridDim.x) * CB;
    /* CB * Nr <= PT * threads (the launcher's KB): all loads issued before the first LDS write;
       consecutive threads read consecutive columns; indices past the end are clamped, not written */
    constexpr int PT = 4 * KB;
    const int cnt = CB * Nr;
    {
        typename Vec2<real>::type v[PT];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = min((int)threadIdx.x + k * kFftThreads, cnt - 1);
            const int r = i / CB, j = i - r * CB;
            v[k] = data[(size_t)r * Nc + c0 + j];
        }
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = threadIdx.x + k * kFftThreads;
            const int r = i / CB, j = i - r * CB;
            if (i < cnt) A[lslot<PAD>(j * Nr + r)] = {v[k].x, v[k].y};
        }
    }
    __syncthreads();
    stockham<real, -1, KB, PAD>(A, tw, Nr, log2Nr, CB);
    if constexpr (MODE == kColForward) {
    } else if constexpr (MODE == kColPropagate) {
        /* factor tables read before they are used (all loads in flight together) */
        Cx<real> rf[PT];
#pragma unroll
        for (int k = 0; k < PT; k++) rf[k] = rowf[min((int)threadIdx.x + k * kFftThreads, cnt - 1) / CB];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = threadIdx.x + k * kFftThreads;
            const int r = i / CB, j = i - r * CB;
            if (i < cnt) {
                Cx<real>& f = A[lslot<PAD>(j * Nr + r)];
                f = cmul(cmul(f, colf[c0 + j]), rf[k]);
            }
        }
    } else {
        /* mask of the quadrant-swapped grid (fft_shift_2D, kernel_fft.cuh:15-32), applied as data / N * mask;
           row-major walk: consecutive threads read consecutive mask cells, all loads issued first */
        real mk[PT];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = min((int)threadIdx.x + k * kFftThreads, cnt - 1);
            const int r = i / CB, j = i - r * CB;
            const int rs = (r + (Nr >> 1)) & (Nr - 1);
            const int cs = (c0 + j + (Nc >> 1)) & (Nc - 1);
            mk[k] = mask[(size_t)rs * Nc + cs];
        }
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = threadIdx.x + k * kFftThreads;
            const int r = i / CB, j = i - r * CB;
            if (i < cnt) {
                Cx<real>& f = A[lslot<PAD>(j * Nr + r)];
                f.r = f.r / Ntot * mk[k];
                f.i = f.i / Ntot * mk[k];
                if (spec) {
                    typename Vec2<real>::type o;
                    o.x = f.r;
                    o.y = f.i;
                    spec[(size_t)r * Nc + c0 + j] = o;
                }
            }
        }
    }
    __syncthreads();
    if constexpr (MODE != kColForward) stockham<real, +1, KB, PAD>(A, tw, Nr, log2Nr, CB);
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int i = threadIdx.x + k * kFftThreads;
        const int r = i / CB, j = i - r * CB;
        if (i < cnt) {
            typename Vec2<real>::type o;
            const Cx<real> f = A[lslot<PAD>(j * Nr + r)];
            o.x = f.r;
            o.y = f.i;
            data[(size_t)r * Nc + c0 + j] = o;
        }
    }
}

/* ------------------------------------------------------------------------------------------
   Four-step transforms for N = N1 N2 (N1, N2 in {16, 32}).  With n = N2 n1 + n2 and k = k1 + N1 k2,
       X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_N^(n2 k1) sum_n1 x[N2 n1 + n2] W_N1^(n1 k1),
   so a thread loads one strided sub-sequence (n2 fixed, n1 = 0..N1-1) into registers, transforms
   it there, applies the twiddle W_N^(n2 k1), and writes it to LDS transposed; after one barrier a
   thread reads one sub-sequence of the other kind (k1 fixed) and transforms it in registers.  One
   LDS round trip and one barrier per length-N transform instead of log4 N Stockham passes with two
   barriers each.  Global accesses: in a row, the threads of a sequence read consecutive elements;
   in the column pass, consecutive threads read consecutive columns. */

/* in-register radix-2 FFT of R elements: X[k] = sum_j x[j] e^(SIGN 2 pi i j k / R); tw holds
   W_N^m = e^(-2 pi i m / N) for m < N and R divides N (uniform indices: scalar loads) */
template <typename real, int R, int SIGN>
__device__ __forceinline__ void fft_reg(Cx<real> (&x)[R], const Cx<real>* __restrict__ tw, int N) {
    constexpr int bits = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : R == 32 ? 5 : 6;
    Cx<real> y[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        int r = 0;
#pragma unroll
        for (int b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
        y[r] = x[i];
    }
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int j = 0; j < half; j++) {
            Cx<real> w = tw[j * (N / len)];
            if (SIGN > 0) w.i = -w.i;
#pragma unroll
            for (int i = 0; i < R; i += len) {
                const Cx<real> u = y[i + j];
                const Cx<real> v = j == 0 ? y[i + j + half] : cmul(y[i + j + half], w);
                y[i + j] = {u.r + v.r, u.i + v.i};
                y[i + j + half] = {u.r - v.r, u.i - v.i};
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; i++) x[i] = y[i];
}

/* LDS slot of (row, col) in a [rows][W] tile, W a power of two: the column is XORed with the row,
   so that both the row-wise writes and the column-wise reads of the transpose spread over the banks
   without padding (a tile stays within 64 KiB) */
template <int W>
__device__ __forceinline__ int swz(int row, int col) { return row * W + (col ^ (row & (W - 1))); }

template <typename real, int SIGN>
__device__ __forceinline__ Cx<real> twiddle(const Cx<real>* __restrict__ tw, int m) {
    Cx<real> w = tw[m];
    if (SIGN > 0) w.i = -w.i;
    return w;
}

/* Lane pairs: the four-step kernels give each sub-sequence of R = 2H points to two adjacent lanes (p = 0, 1),
   one radix-2 stage between them (one DPP lane swap per element) and an H-point transform in each lane, so
   that a row or column has 2 max(N1, N2) threads.  pair_dit: lane p holds x[2j + p] on entry and X[m + p H]
   on return (X[m] = E[m] + w^m O[m], X[m + H] = E[m] - w^m O[m], E / O the transforms of the even /
   odd points); pair_dif: lane p holds x[m + p H] on entry and X[2k + p] on return (the halves
   x[m] + x[m + H] and (x[m] - x[m + H]) w^m transformed); w = e^(SIGN 2 pi i / R). */
template <typename real>
__device__ __forceinline__ Cx<real> shfl_pair(Cx<real> v) {
    return {dpp_move<kDppPairSwap>(v.r), dpp_move<kDppPairSwap>(v.i)};
}

template <typename real, int H, int SIGN>
__device__ __forceinline__ void pair_dit(Cx<real> (&x)[H], const Cx<real>* __restrict__ tw, int N, int p) {
    fft_reg<real, H, SIGN>(x, tw, N);
    const int step = N / (2 * H);
#pragma unroll
    for (int m = 0; m < H; m++) {
        const Cx<real> b = shfl_pair(x[m]);
        const Cx<real> e = p ? b : x[m], o = p ? x[m] : b;
        const Cx<real> q = m == 0 ? o : cmul(o, twiddle<real, SIGN>(tw, m * step));
        const Cx<real> sum = {e.r + q.r, e.i + q.i}, dif = {e.r - q.r, e.i - q.i};
        x[m] = p ? dif : sum;
    }
}

template <typename real, int H, int SIGN>
__device__ __forceinline__ void pair_dif(Cx<real> (&x)[H], const Cx<real>* __restrict__ tw, int N, int p) {
    const int step = N / (2 * H);
#pragma unroll
    for (int m = 0; m < H; m++) {
        const Cx<real> b = shfl_pair(x[m]);
        const Cx<real> sum = {x[m].r + b.r, x[m].i + b.i}, d = {b.r - x[m].r, b.i - x[m].i};
        const Cx<real> dw = m == 0 ? d : cmul(d, twiddle<real, SIGN>(tw, m * step));
        x[m] = p ? dw : sum;
    }
    fft_reg<real, H, SIGN>(x, tw, N);
}

/* LDS elements between the column slabs of fft4_cols beyond N: the slabs of the CB columns a wave
   touches then start on different banks (N apart, the same (k1, t) of every column fell on one
   bank).  SQ_LDS_BANK_CONFLICT per LDS instruction at 1024^2, FP32 / FP64: skew 0 21 / -, 2 1 / 2,
   4 2 / 4, 8 7 / 14; the pass time moved by under 1 % between skews 2 and 8 (parallelism-bound). */
template <typename real>
__host__ __device__ constexpr int fft4_col_skew() { return 2; }

/* threads per workgroup of the four-step kernels (two lanes per sub-sequence) */
constexpr int kFft4Threads = 256;

/* rows: kFft4Threads / (2 max(N1, N2)) rows per workgroup, two lanes per sub-sequence (lane pairs
   above), LDS [rows][N1][N2] (swizzled).  A wave holds one 1024-point row: each load and store
   instruction moves two contiguous 32-element runs of it. */
template <typename real, int N1, int N2, int SIGN>
__global__ __launch_bounds__(kFft4Threads) void fft4_rows(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw, int nrows) {
    constexpr int N = N1 * N2, T = N1 > N2 ? N1 : N2, H1 = N1 / 2, H2 = N2 / 2, RPB = kFft4Threads / (2 * T);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<real>* S = reinterpret_cast<Cx<real>*>(lds_raw);
    const int p = threadIdx.x & 1, r = threadIdx.x >> 1;
    const int lr = r / T, t = r - lr * T;
    const int row = blockIdx.x * RPB + lr;
    const bool live = row < nrows;
    typename Vec2<real>::type* rp = data + (size_t)(live ? row : 0) * N;
    Cx<real>* sr = S + lr * N;
    /* n2 = t; lane p holds n1 = 2j + p, then k1 = m + p H1 */
    if (t < N2) {
        Cx<real> x[H1];
#pragma unroll
        for (int j = 0; j < H1; j++) {
            const auto v = rp[N2 * (2 * j + p) + t];
            x[j] = {v.x, v.y};
        }
        pair_dit<real, H1, SIGN>(x, tw, N, p);
#pragma unroll
        for (int m = 0; m < H1; m++) {
            const int k1 = m + p * H1;
            sr[swz<N2>(k1, t)] = k1 == 0 ? x[m] : cmul(x[m], twiddle<real, SIGN>(tw, t * k1));
        }
    }
    __syncthreads();
    /* k1 = t; lane p reads n2 = 2j + p, then holds X[t + N1 k2], k2 = m + p H2 */
    if (t < N1 && live) {
        Cx<real> y[H2];
#pragma unroll
        for (int j = 0; j < H2; j++) y[j] = sr[swz<N2>(t, 2 * j + p)];
        pair_dit<real, H2, SIGN>(y, tw, N, p);
#pragma unroll
        for (int m = 0; m < H2; m++) {
            typename Vec2<real>::type o;
            o.x = y[m].r;
            o.y = y[m].i;
            rp[t + N1 * (m + p * H2)] = o;
        }
    }
}

/* columns: CB adjacent columns per workgroup, 2 max(N1, N2) threads per column (lane pairs above):
   forward transform, the k-space factor of MODE (as fft_cols_filter), inverse transform; LDS
   [CB][N1][N2] for the forward transpose and [CB][N2][N1] for the inverse (swizzled).
   thread = p + 2 (column + CB sub-sequence).  At 1024^2: 4 columns per workgroup, 256 workgroups,
   1024 waves (one lane per sub-sequence gave 512 waves, half the SIMDs, and 16.9 / 19.5 us FP32 /
   FP64).  Whole 128-byte row segments per workgroup (16 FP32 columns) were slower: the pass is bound
   by its parallelism, not by the segment width. */
template <typename real, int N1, int N2, int CB, int MODE>
__global__ __launch_bounds__(2 * CB *(N1 > N2 ? N1 : N2)) void fft4_cols(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw,
                                                     const real* __restrict__ mask, const Cx<real>* __restrict__ rowf,
                                                     const Cx<real>* __restrict__ colf, int Nc, real Ntot, typename Vec2<real>::type* spec) {
    constexpr int N = N1 * N2, H1 = N1 / 2, H2 = N2 / 2;
    constexpr int SLAB = N + fft4_col_skew<real>();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<real>* S = reinterpret_cast<Cx<real>*>(lds_raw);
    const int p = threadIdx.x & 1, r = threadIdx.x >> 1;
    const int c = r % CB, t = r / CB;
    const int col = fft_xcd_group(blockIdx.x, gridDim.x) * CB + c;
    Cx<real>* sc = S + c * SLAB;
    /* forward, first half: n2 = t; lane p holds n1 = 2j + p, then k1 = m + p H1 */
    if (t < N2) {
        Cx<real> x[H1];
#pragma unroll
        for (int j = 0; j < H1; j++) {
            const auto v = data[(size_t)(N2 * (2 * j + p) + t) * Nc + col];
            x[j] = {v.x, v.y};
        }
        pair_dit<real, H1, -1>(x, tw, N, p);
#pragma unroll
        for (int m = 0; m < H1; m++) {
            const int k1 = m + p * H1;
            sc[swz<N2>(k1, t)] = k1 == 0 ? x[m] : cmul(x[m], twiddle<real, -1>(tw, t * k1));
        }
    }
    __syncthreads();
    /* forward, second half: k1 = t; lane p reads n2 = 2j + p, then holds X[t + N1 k2], k2 = m + p H2 */
    Cx<real> y[H2];
    if (t < N1) {
#pragma unroll
        for (int j = 0; j < H2; j++) y[j] = sc[swz<N2>(t, 2 * j + p)];
        pair_dit<real, H2, -1>(y, tw, N, p);
        if constexpr (MODE == kColForward) {
#pragma unroll
            for (int m = 0; m < H2; m++) {
                typename Vec2<real>::type o;
                o.x = y[m].r;
                o.y = y[m].i;
                data[(size_t)(t + N1 * (m + p * H2)) * Nc + col] = o;
            }
        } else if constexpr (MODE == kColPropagate) {
            const Cx<real> cf = colf[col];
#pragma unroll
            for (int m = 0; m < H2; m++) y[m] = cmul(cmul(y[m], cf), rowf[t + N1 * (m + p * H2)]);
        } else {
            const int cs = (col + (Nc >> 1)) & (Nc - 1);
            real mk[H2];
#pragma unroll
            for (int m = 0; m < H2; m++) mk[m] = mask[(size_t)(((t + N1 * (m + p * H2)) + (N >> 1)) & (N - 1)) * Nc + cs];
#pragma unroll
            for (int m = 0; m < H2; m++) {
                y[m].r = y[m].r / Ntot * mk[m];
                y[m].i = y[m].i / Ntot * mk[m];
                if (spec) {
                    typename Vec2<real>::type o;
                    o.x = y[m].r;
                    o.y = y[m].i;
                    spec[(size_t)(t + N1 * (m + p * H2)) * Nc + col] = o;
                }
            }
        }
    }
    if constexpr (MODE == kColForward) return;
    __syncthreads(); /* the forward transpose slab is reused */
    /* inverse, first half: k1 = t; the pair holds k2 in halves and leaves Z[p'], p' = 2j + p,
       Z[p'] = sum_k2 X[t + N1 k2] W_N2^(-p' k2), then times W_N^(-p' t) */
    if (t < N1) {
        pair_dif<real, H2, +1>(y, tw, N, p);
#pragma unroll
        for (int j = 0; j < H2; j++) {
            const int q = 2 * j + p;
            sc[swz<N1>(q, t)] = q == 0 ? y[j] : cmul(y[j], twiddle<real, +1>(tw, q * t));
        }
    }
    __syncthreads();
    /* inverse, second half: p' = t; lane p reads k1 = 2j + p, then holds x[t + N2 q], q = m + p H1 */
    if (t < N2) {
        Cx<real> w[H1];
#pragma unroll
        for (int j = 0; j < H1; j++) w[j] = sc[swz<N1>(t, 2 * j + p)];
        pair_dit<real, H1, +1>(w, tw, N, p);
#pragma unroll
        for (int m = 0; m < H1; m++) {
            typename Vec2<real>::type o;
            o.x = w[m].r;
            o.y = w[m].i;
            data[(size_t)(t + N2 * (m + p * H1)) * Nc + col] = o;
        }
    }
}

/* Lengths of the four-step path: (N1, N2) with N1 N2 = N */
inline bool fft4_split(int N, int& N1, int& N2) {
    switch (N) {
        case 256: N1 = 16, N2 = 16; return true;
        case 512: N1 = 32, N2 = 16; return true;
        case 1024: N1 = 32, N2 = 32; return true;
        default: return false;
    }
}

static bool fft4_enabled() {
    const char* e = std::getenv("PHX_FFT4");
    return !(e && std::atoi(e) == 0);
}

/* rows of the four-step path by the four-step row kernel (PHX_FFT4_ROWS=0: the Stockham one).  With
   lane pairs (1024 waves at 1024^2) it is the faster: FP32 7.38 / 9.93 us forward / inverse against
   7.41 / 10.02 for the Stockham rows, FP64 9.3 / 11.5 against 10.0 / 15.7; with one lane per
   sub-sequence (512 waves) it had been slower in FP32 (7.88 / 10.14). */
static bool fft4_rows_enabled() {
    static const bool on = [] { const char* e = std::getenv("PHX_FFT4_ROWS"); return !(e && std::atoi(e) == 0); }();
    return on;
}

template <typename real, int N1, int N2, int SIGN>
static hipError_t launch_fft4_rows_t(typename Vec2<real>::type* psi, const void* tw, int Nr, hipStream_t s) {
    constexpr int T = N1 > N2 ? N1 : N2, RPB = kFft4Threads / (2 * T);
    const size_t lds = (size_t)RPB * N1 * N2 * sizeof(Cx<real>);
    hipLaunchKernelGGL((fft4_rows<real, N1, N2, SIGN>), dim3((Nr + RPB - 1) / RPB), dim3(kFft4Threads), lds, s, psi,
                       (const Cx<real>*)tw, Nr);
    return hipGetLastError();
}

template <typename real, int SIGN>
static hipError_t launch_fft4_rows(typename Vec2<real>::type* psi, const void* tw, int Nc, int Nr, hipStream_t s) {
    switch (Nc) {
        case 256: return launch_fft4_rows_t<real, 16, 16, SIGN>(psi, tw, Nr, s);
        case 512: return launch_fft4_rows_t<real, 32, 16, SIGN>(psi, tw, Nr, s);
        case 1024: return launch_fft4_rows_t<real, 32, 32, SIGN>(psi, tw, Nr, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename real, int N1, int N2, int MODE>
static hipError_t launch_fft4_cols_t(typename Vec2<real>::type* psi, const void* tw, const real* mask, const void* rowf,
                                     const void* colf, int Nc, int Nr, hipStream_t s, typename Vec2<real>::type* spec) {
    constexpr int T = N1 > N2 ? N1 : N2, CB = kFft4Threads / (2 * T);
    if (Nc % CB) return hipErrorInvalidValue;
    const size_t lds = (size_t)CB * (N1 * N2 + fft4_col_skew<real>()) * sizeof(Cx<real>);
    static const hipError_t attr = lds > 65536 ? hipFuncSetAttribute((const void*)&fft4_cols<real, N1, N2, CB, MODE>,
                                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)
                                               : hipSuccess; /* beyond 64 KiB of LDS (gfx950 grants up to 160 KiB) */
    if (attr != hipSuccess) return attr;
    const real ntot = (real)((double)Nc * (double)Nr);
    hipLaunchKernelGGL((fft4_cols<real, N1, N2, CB, MODE>), dim3(Nc / CB), dim3(2 * CB * T), lds, s, psi, (const Cx<real>*)tw, mask,
                       (const Cx<real>*)rowf, (const Cx<real>*)colf, Nc, ntot, spec);
    return hipGetLastError();
}

template <typename real, int MODE>
static hipError_t launch_fft4_cols(typename Vec2<real>::type* psi, const void* tw, const real* mask, const void* rowf, const void* colf,
                                   int Nc, int Nr, hipStream_t s, typename Vec2<real>::type* spec) {
    switch (Nr) {
        case 256: return launch_fft4_cols_t<real, 16, 16, MODE>(psi, tw, mask, rowf, colf, Nc, Nr, s, spec);
        case 512: return launch_fft4_cols_t<real, 32, 16, MODE>(psi, tw, mask, rowf, colf, Nc, Nr, s, spec);
        case 1024: return launch_fft4_cols_t<real, 32, 32, MODE>(psi, tw, mask, rowf, colf, Nc, Nr, s, spec);
        default: return hipErrorInvalidValue;
    }
}

/* ------------------------------------------------------------------------------------------
   Direct DFT for lengths that are not powers of two (any even N; the reference validates even
   grid sizes, system_validate_input.cpp:19-56): out[k] = sum_j x[j] w^((j k) mod N) with
   w = e^(-+2 pi i / N) from the twiddle table, one output per thread and pass, the sequence in
   LDS.  O(N) multiply-adds per output value instead of O(log N): a general-size path next to the
   radix-4 Stockham kernels, which stay the fast path for powers of two. */
template <typename real, int SIGN>
__device__ __forceinline__ Cx<real> dft_one(const Cx<real>* x, int stride, const Cx<real>* __restrict__ tw, int N, int k) {
    real sr = 0, si = 0;
    int idx = 0;
    for (int j = 0; j < N; j++) {
        Cx<real> w = tw[idx];
        if (SIGN > 0) w.i = -w.i;
        const Cx<real> v = x[j * stride];
        sr = phx_fma(v.r, w.r, phx_fma(-v.i, w.i, sr));
        si = phx_fma(v.r, w.i, phx_fma(v.i, w.r, si));
        idx += k; /* (j k) mod N, incrementally: idx + k < 2 N */
        if (idx >= N) idx -= N;
    }
    return {sr, si};
}

template <typename real, int SIGN>
__global__ __launch_bounds__(kFftThreads) void dft_rows(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw, int N) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<real>* A = reinterpret_cast<Cx<real>*>(lds_raw);
    typename Vec2<real>::type* row = data + (size_t)blockIdx.x * (size_t)N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const auto v = row[i];
        A[i] = {v.x, v.y};
    }
    __syncthreads();
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        const Cx<real> f = dft_one<real, SIGN>(A, 1, tw, N, k);
        typename Vec2<real>::type v;
        v.x = f.r;
        v.y = f.i;
        row[k] = v;
    }
}

/* CB adjacent columns: forward DFT along the column, the k-space factor (MODE as fft_cols_filter, with
   the same optional spectrum output), inverse DFT; two LDS buffers (the transforms are out of place). */
template <typename real, int CB, int MODE>
__global__ __launch_bounds__(kFftThreads) void dft_cols_filter(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw,
                                                              const real* __restrict__ mask, const Cx<real>* __restrict__ rowf,
                                                              const Cx<real>* __restrict__ colf, int Nc, int Nr, real Ntot,
                                                              typename Vec2<real>::type* spec) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<real>* A = reinterpret_cast<Cx<real>*>(lds_raw);
    Cx<real>* B = A + CB * Nr;
    const int c0 = fft_xcd_group(blockIdx.x, gridDim.x) * CB;
    for (int i = threadIdx.x; i < CB * Nr; i += blockDim.x) {
        const int r = i / CB, j = i - r * CB;
        const auto v = data[(size_t)r * Nc + c0 + j];
        A[j * Nr + r] = {v.x, v.y};
    }
    __syncthreads();
    for (int i = threadIdx.x; i < CB * Nr; i += blockDim.x) {
        const int k = i / CB, j = i - k * CB;
        Cx<real> f = dft_one<real, -1>(A + j * Nr, 1, tw, Nr, k);
        if constexpr (MODE == kColForward) {
            typename Vec2<real>::type o;
            o.x = f.r;
            o.y = f.i;
            data[(size_t)k * Nc + c0 + j] = o; /* A holds the columns: the in-place store is safe */
            continue;
        } else if constexpr (MODE == kColPropagate) {
            f = cmul(cmul(f, colf[c0 + j]), rowf[k]);
        } else {
            /* mask of the quadrant-swapped grid (fft_shift_2D, kernel_fft.cuh:15-32) */
            int rs = k + (Nr >> 1), cs = c0 + j + (Nc >> 1);
            if (rs >= Nr) rs -= Nr;
            if (cs >= Nc) cs -= Nc;
            const real m = mask[(size_t)rs * Nc + cs];
            f.r = f.r / Ntot * m;
            f.i = f.i / Ntot * m;
            if (spec) {
                typename Vec2<real>::type o;
                o.x = f.r;
                o.y = f.i;
                spec[(size_t)k * Nc + c0 + j] = o;
            }
        }
        B[j * Nr + k] = f;
    }
    if constexpr (MODE == kColForward) return;
    __syncthreads();
    for (int i = threadIdx.x; i < CB * Nr; i += blockDim.x) {
        const int r = i / CB, j = i - r * CB;
        const Cx<real> f = dft_one<real, +1>(B + j * Nr, 1, tw, Nr, r);
        typename Vec2<real>::type v;
        v.x = f.r;
        v.y = f.i;
        data[(size_t)r * Nc + c0 + j] = v;
    }
}

/* sizes the direct DFT takes: even, and two column buffers of one column in 64 KiB */
static bool dft_supported(int N, size_t real_size) { return N >= 2 && N % 2 == 0 && (size_t)N * 4 * real_size <= 65536; }

template <typename real, int MODE>
static hipError_t dft_roundtrip(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const real* mask, const void* rowf,
                                const void* colf, int Nc, int Nr, hipStream_t s, typename Vec2<real>::type* spec = nullptr) {
    const auto* twc = (const Cx<real>*)tw_c;
    const auto* twr = (const Cx<real>*)tw_r;
    const auto* rf = (const Cx<real>*)rowf;
    const auto* cf = (const Cx<real>*)colf;
    const real ntot = (real)((double)Nc * (double)Nr);
    hipLaunchKernelGGL((dft_rows<real, -1>), dim3(Nr), dim3(kFftThreads), (size_t)Nc * sizeof(Cx<real>), s, psi, twc, Nc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t per_col = 2 * (size_t)Nr * sizeof(Cx<real>);
    auto fits = [&](int cb) { return per_col * cb <= 65536 && Nc % cb == 0; };
    if (fits(4))
        hipLaunchKernelGGL((dft_cols_filter<real, 4, MODE>), dim3(Nc / 4), dim3(kFftThreads), 4 * per_col, s, psi, twr, mask, rf, cf, Nc, Nr, ntot, spec);
    else if (fits(2))
        hipLaunchKernelGGL((dft_cols_filter<real, 2, MODE>), dim3(Nc / 2), dim3(kFftThreads), 2 * per_col, s, psi, twr, mask, rf, cf, Nc, Nr, ntot, spec);
    else
        hipLaunchKernelGGL((dft_cols_filter<real, 1, MODE>), dim3(Nc), dim3(kFftThreads), per_col, s, psi, twr, mask, rf, cf, Nc, Nr, ntot, spec);
    e = hipGetLastError();
    if (e != hipSuccess || MODE == kColForward) return e;
    hipLaunchKernelGGL((dft_rows<real, +1>), dim3(Nr), dim3(kFftThreads), (size_t)Nc * sizeof(Cx<real>), s, psi, twc, Nc);
    return hipGetLastError();
}

static int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

/* lengths the radix-4 Stockham kernels take: powers of two with one LDS row of N complex values
   and at most max_bfly butterflies per thread */
static bool stockham_supported(int N, size_t real_size) {
    if (N < 4 || (N & (N - 1))) return false;
    const int mb = real_size == 4 ? max_bfly<float>() : max_bfly<double>();
    return (size_t)N * 2 * real_size <= 65536 && N / 4 <= mb * kFftThreads;
}

bool fft_supported(int N, size_t real_size) { return stockham_supported(N, real_size) || dft_supported(N, real_size); }

/* smallest power of two KB with butterflies <= KB * threads */
static int bfly_bound(int butterflies) {
    int kb = 1;
    while (kb * kFftThreads < butterflies) kb <<= 1;
    return kb;
}

/* PHX_FFT_PAD=1 pads the LDS index of the Stockham kernels (one slot per 32) where the padded buffer
   fits the 64 KiB a workgroup may hold: it removes the 4-way bank conflicts of the first passes'
   writes but measured slower at 1024^2 (39.4 vs 37.4 us FP32, 51.8 vs 49.5 us FP64), so it is off */
static bool pad_lds(size_t elems, size_t elem_size) {
    static const bool env_on = [] { const char* e = std::getenv("PHX_FFT_PAD"); return e && std::atoi(e) == 1; }();
    return env_on && lds_elems<true>(elems) * elem_size <= 65536;
}

template <typename real, int SIGN, bool PAD>
static hipError_t launch_rows_p(typename Vec2<real>::type* psi, const void* tw_c, int Nc, int Nr, hipStream_t s) {
    const size_t lds = lds_elems<PAD>((size_t)Nc) * sizeof(Cx<real>);
    const int kb = bfly_bound(Nc / 4);
    const auto* tw = (const Cx<real>*)tw_c;
    const int l = ilog2(Nc);
    if (kb == 1)
        hipLaunchKernelGGL((fft_rows<real, SIGN, 1, PAD>), dim3(Nr), dim3(kFftThreads), lds, s, psi, tw, Nc, l);
    else if (kb == 2)
        hipLaunchKernelGGL((fft_rows<real, SIGN, 2, PAD>), dim3(Nr), dim3(kFftThreads), lds, s, psi, tw, Nc, l);
    else if (kb == 4)
        hipLaunchKernelGGL((fft_rows<real, SIGN, 4, PAD>), dim3(Nr), dim3(kFftThreads), lds, s, psi, tw, Nc, l);
    else if constexpr (max_bfly<real>() >= 8)
        hipLaunchKernelGGL((fft_rows<real, SIGN, 8, PAD>), dim3(Nr), dim3(kFftThreads), lds, s, psi, tw, Nc, l);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename real, int SIGN>
static hipError_t launch_rows(typename Vec2<real>::type* psi, const void* tw_c, int Nc, int Nr, hipStream_t s) {
    if (pad_lds((size_t)Nc, sizeof(Cx<real>))) return launch_rows_p<real, SIGN, true>(psi, tw_c, Nc, Nr, s);
    return launch_rows_p<real, SIGN, false>(psi, tw_c, Nc, Nr, s);
}

template <typename real, int CB, int MODE, bool PAD>
static hipError_t launch_cols_p(typename Vec2<real>::type* psi, const void* tw_r, const real* mask, const void* rowf, const void* colf,
                                int Nc, int Nr, hipStream_t s, typename Vec2<real>::type* spec) {
    const size_t lds = lds_elems<PAD>((size_t)CB * Nr) * sizeof(Cx<real>);
    const int kb = bfly_bound(CB * (Nr / 4));
    const auto* tw = (const Cx<real>*)tw_r;
    const auto* rf = (const Cx<real>*)rowf;
    const auto* cf = (const Cx<real>*)colf;
    const int l = ilog2(Nr);
    const real ntot = (real)((double)Nc * (double)Nr);
    if (kb == 1)
        hipLaunchKernelGGL((fft_cols_filter<real, CB, 1, MODE, PAD>), dim3(Nc / CB), dim3(kFftThreads), lds, s, psi, tw, mask, rf, cf, Nc, Nr, l, ntot, spec);
    else if (kb == 2)
        hipLaunchKernelGGL((fft_cols_filter<real, CB, 2, MODE, PAD>), dim3(Nc / CB), dim3(kFftThreads), lds, s, psi, tw, mask, rf, cf, Nc, Nr, l, ntot, spec);
    else if (kb == 4)
        hipLaunchKernelGGL((fft_cols_filter<real, CB, 4, MODE, PAD>), dim3(Nc / CB), dim3(kFftThreads), lds, s, psi, tw, mask, rf, cf, Nc, Nr, l, ntot, spec);
    else if constexpr (max_bfly<real>() >= 8)
        hipLaunchKernelGGL((fft_cols_filter<real, CB, 8, MODE, PAD>), dim3(Nc / CB), dim3(kFftThreads), lds, s, psi, tw, mask, rf, cf, Nc, Nr, l, ntot, spec);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename real, int CB, int MODE>
static hipError_t launch_cols(typename Vec2<real>::type* psi, const void* tw_r, const real* mask, const void* rowf, const void* colf,
                              int Nc, int Nr, hipStream_t s, typename Vec2<real>::type* spec) {
    if (pad_lds((size_t)CB * Nr, sizeof(Cx<real>)))
        return launch_cols_p<real, CB, MODE, true>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    return launch_cols_p<real, CB, MODE, false>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
}

/* rows forward -> columns (forward, k-space factor, inverse) -> rows inverse; kColForward stops
   after the forward column transform */
template <typename real, int MODE>
static hipError_t fft_roundtrip(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const real* mask, const void* rowf,
                                const void* colf, int Nc, int Nr, hipStream_t s, typename Vec2<real>::type* spec = nullptr) {
    int n1, n2;
    if (fft4_enabled() && fft4_split(Nc, n1, n2) && fft4_split(Nr, n1, n2) && Nc % (kFft4Threads / 32) == 0) {
        /* both row kernels leave each row transformed in natural order, in place */
        const bool rows4 = fft4_rows_enabled();
        hipError_t e = rows4 ? launch_fft4_rows<real, -1>(psi, tw_c, Nc, Nr, s) : launch_rows<real, -1>(psi, tw_c, Nc, Nr, s);
        if (e == hipSuccess) e = launch_fft4_cols<real, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
        if (e != hipSuccess || MODE == kColForward) return e;
        return rows4 ? launch_fft4_rows<real, +1>(psi, tw_c, Nc, Nr, s) : launch_rows<real, +1>(psi, tw_c, Nc, Nr, s);
    }
    hipError_t e = launch_rows<real, -1>(psi, tw_c, Nc, Nr, s);
    if (e != hipSuccess) return e;
    /* column block width: as wide as LDS and registers allow, but no wider than keeps
       512 workgroups in flight (the machine has 256 CUs) */
    const size_t per_col = (size_t)Nr * sizeof(Cx<real>);
    auto fits = [&](int cb) {
        return per_col * cb <= 65536 && Nc % cb == 0 && cb * (Nr / 4) <= max_bfly<real>() * kFftThreads && (cb == 1 || Nc / cb >= 512);
    };
    /* PHX_FFT_CB: column block width for tuning runs (taken when it fits LDS and registers) */
    static const int cb_env = [] { const char* e = std::getenv("PHX_FFT_CB"); return e ? std::atoi(e) : 0; }();
    auto fits_env = [&](int cb) {
        return cb_env == cb && per_col * cb <= 65536 && Nc % cb == 0 && cb * (Nr / 4) <= max_bfly<real>() * kFftThreads;
    };
    if (fits_env(8))
        e = launch_cols<real, 8, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    else if (fits_env(4))
        e = launch_cols<real, 4, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    else if (fits_env(1))
        e = launch_cols<real, 1, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    else if (fits(8))
        e = launch_cols<real, 8, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    else if (fits(4))
        e = launch_cols<real, 4, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    else if (fits(2))
        e = launch_cols<real, 2, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    else
        e = launch_cols<real, 1, MODE>(psi, tw_r, mask, rowf, colf, Nc, Nr, s, spec);
    if (e != hipSuccess || MODE == kColForward) return e;
    return launch_rows<real, +1>(psi, tw_c, Nc, Nr, s);
}

/* powers of two on both axes: Stockham kernels; otherwise the direct DFT kernels for both axes */
template <typename real>
hipError_t launch_fft_filter(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const real* mask, int Nc, int Nr,
                             hipStream_t s, typename Vec2<real>::type* spec) {
    if (stockham_supported(Nc, sizeof(real)) && stockham_supported(Nr, sizeof(real)))
        return fft_roundtrip<real, kColFilter>(psi, tw_c, tw_r, mask, nullptr, nullptr, Nc, Nr, s, spec);
    if (!fft_supported(Nc, sizeof(real)) || !fft_supported(Nr, sizeof(real))) return hipErrorInvalidValue;
    return dft_roundtrip<real, kColFilter>(psi, tw_c, tw_r, mask, nullptr, nullptr, Nc, Nr, s, spec);
}

/* the unnormalised forward 2D transform of the grid, in place */
template <typename real>
hipError_t launch_fft_forward(typename Vec2<real>::type* data, const void* tw_c, const void* tw_r, int Nc, int Nr, hipStream_t s) {
    if (stockham_supported(Nc, sizeof(real)) && stockham_supported(Nr, sizeof(real)))
        return fft_roundtrip<real, kColForward>(data, tw_c, tw_r, nullptr, nullptr, nullptr, Nc, Nr, s);
    if (!fft_supported(Nc, sizeof(real)) || !fft_supported(Nr, sizeof(real))) return hipErrorInvalidValue;
    return dft_roundtrip<real, kColForward>(data, tw_c, tw_r, nullptr, nullptr, nullptr, Nc, Nr, s);
}

template <typename real>
hipError_t launch_fft_propagate(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const void* rowf, const void* colf,
                                int Nc, int Nr, hipStream_t s) {
    if (stockham_supported(Nc, sizeof(real)) && stockham_supported(Nr, sizeof(real)))
        return fft_roundtrip<real, kColPropagate>(psi, tw_c, tw_r, nullptr, rowf, colf, Nc, Nr, s);
    if (!fft_supported(Nc, sizeof(real)) || !fft_supported(Nr, sizeof(real))) return hipErrorInvalidValue;
    return dft_roundtrip<real, kColPropagate>(psi, tw_c, tw_r, nullptr, rowf, colf, Nc, Nr, s);
}

template hipError_t launch_fft_propagate<float>(float2*, const void*, const void*, const void*, const void*, int, int, hipStream_t);
template hipError_t launch_fft_propagate<double>(double2*, const void*, const void*, const void*, const void*, int, int, hipStream_t);
template hipError_t launch_fft_filter<float>(float2*, const void*, const void*, const float*, int, int, hipStream_t, float2*);
template hipError_t launch_fft_filter<double>(double2*, const void*, const void*, const double*, int, int, hipStream_t, double2*);
template hipError_t launch_fft_forward<float>(float2*, const void*, const void*, int, int, hipStream_t);
template hipError_t launch_fft_forward<double>(double2*, const void*, const void*, int, int, hipStream_t);

} // namespace phx
