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
 *
 * The transforms are unnormalised DFTs with the forward sign e^{-2 pi i jk/N}
 * (cuFFT/FFTW convention).  Algorithm: Stockham autosort, radix 4 with a final
 * radix-2 pass when log2 N is odd (textbook form, e.g. Van Loan, "Computational
 * Frameworks for the FFT", 1992).  Twiddles come from a table computed on the
 * host in double precision.  Supported lengths: powers of two from 4 to 4096
 * (FP32) / 2048 (FP64), as two LDS buffers of one row must fit in 64 KiB.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "custom_rhs.h"
#include "phx_kernels.h"

namespace phx {

constexpr int kFftThreads = 256;

template <typename real>
struct Cx {
    real r, i;
};

template <typename real>
__device__ __forceinline__ Cx<real> cmul(Cx<real> a, Cx<real> w) {
    return {phx_fma(a.r, w.r, -(a.i * w.i)), phx_fma(a.r, w.i, a.i * w.r)};
}

/* One Stockham pass group over `count` independent length-N sequences stored
   back to back in src (sequence j at src + j*N).  Returns the buffer holding
   the result.  sign = -1 forward, +1 inverse (conjugated twiddles). */
template <typename real, int SIGN>
__device__ Cx<real>* stockham(Cx<real>* src, Cx<real>* dst, const Cx<real>* __restrict__ tw, int N, int log2N, int count) {
    int n = N, log2n = log2N, s = 1, log2s = 0;
    while (n >= 4) {
        const int m = n >> 2;
        const int stepshift = log2N - log2n; /* twiddle index p * N / n */
        const int nb = N >> 2;              /* butterflies per sequence */
        for (int b = threadIdx.x; b < nb * count; b += blockDim.x) {
            const int seq = b / nb;
            const int bb = b - seq * nb;
            const int p = bb >> log2s;
            const int q = bb & (s - 1);
            const Cx<real>* x = src + seq * N;
            Cx<real>* y = dst + seq * N;
            const Cx<real> a = x[q + s * p], bv = x[q + s * (p + m)], c = x[q + s * (p + 2 * m)], d = x[q + s * (p + 3 * m)];
            Cx<real> w1 = tw[p << stepshift], w2 = tw[(2 * p) << stepshift], w3 = tw[(3 * p) << stepshift];
            if (SIGN > 0) {
                w1.i = -w1.i;
                w2.i = -w2.i;
                w3.i = -w3.i;
            }
            const Cx<real> apc = {a.r + c.r, a.i + c.i}, amc = {a.r - c.r, a.i - c.i};
            const Cx<real> bpd = {bv.r + d.r, bv.i + d.i};
            /* j(b - d) with j = -i (forward) or +i (inverse) */
            const Cx<real> jbmd = SIGN < 0 ? Cx<real>{bv.i - d.i, -(bv.r - d.r)} : Cx<real>{-(bv.i - d.i), bv.r - d.r};
            y[q + s * (4 * p)] = {apc.r + bpd.r, apc.i + bpd.i};
            y[q + s * (4 * p + 1)] = cmul(Cx<real>{amc.r + jbmd.r, amc.i + jbmd.i}, w1);
            y[q + s * (4 * p + 2)] = cmul(Cx<real>{apc.r - bpd.r, apc.i - bpd.i}, w2);
            y[q + s * (4 * p + 3)] = cmul(Cx<real>{amc.r - jbmd.r, amc.i - jbmd.i}, w3);
        }
        __syncthreads();
        Cx<real>* t = src;
        src = dst;
        dst = t;
        n = m;
        log2n -= 2;
        s <<= 2;
        log2s += 2;
    }
    if (n == 2) {
        for (int b = threadIdx.x; b < s * count; b += blockDim.x) {
            const int seq = b >> log2s;
            const int q = b & (s - 1);
            const Cx<real>* x = src + seq * N;
            Cx<real>* y = dst + seq * N;
            const Cx<real> a = x[q], bv = x[q + s];
            y[q] = {a.r + bv.r, a.i + bv.i};
            y[q + s] = {a.r - bv.r, a.i - bv.i};
        }
        __syncthreads();
        Cx<real>* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

template <typename real, int SIGN>
__global__ __launch_bounds__(kFftThreads) void fft_rows(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw, int N, int log2N) {
    extern __shared__ unsigned char lds_raw[];
    Cx<real>* A = reinterpret_cast<Cx<real>*>(lds_raw);
    Cx<real>* B = A + N;
    typename Vec2<real>::type* row = data + (size_t)blockIdx.x * (size_t)N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const auto v = row[i];
        A[i] = {v.x, v.y};
    }
    __syncthreads();
    const Cx<real>* R = stockham<real, SIGN>(A, B, tw, N, log2N, 1);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        typename Vec2<real>::type v;
        v.x = R[i].r;
        v.y = R[i].i;
        row[i] = v;
    }
}

/* CB adjacent columns: forward FFT along the column, F <- F / (N_c N_r) * mask[shifted], inverse FFT. */
template <typename real, int CB>
__global__ __launch_bounds__(kFftThreads) void fft_cols_filter(typename Vec2<real>::type* data, const Cx<real>* __restrict__ tw,
                                                              const real* __restrict__ mask, int Nc, int Nr, int log2Nr, real Ntot) {
    extern __shared__ unsigned char lds_raw[];
    Cx<real>* A = reinterpret_cast<Cx<real>*>(lds_raw);
    Cx<real>* B = A + CB * Nr;
    const int c0 = blockIdx.x * CB;
    for (int i = threadIdx.x; i < CB * Nr; i += blockDim.x) {
        const int r = i / CB, j = i - r * CB; /* consecutive threads read consecutive columns */
        const auto v = data[(size_t)r * Nc + c0 + j];
        A[j * Nr + r] = {v.x, v.y};
    }
    __syncthreads();
    Cx<real>* F = stockham<real, -1>(A, B, tw, Nr, log2Nr, CB);
    Cx<real>* G = F == A ? B : A;
    /* mask of the quadrant-swapped grid (fft_shift_2D, kernel_fft.cuh:15-32), applied as data / N * mask */
    for (int i = threadIdx.x; i < CB * Nr; i += blockDim.x) {
        const int j = i / Nr, r = i - j * Nr;
        const int rs = (r + (Nr >> 1)) & (Nr - 1);
        const int cs = (c0 + j + (Nc >> 1)) & (Nc - 1);
        const real m = mask[(size_t)rs * Nc + cs];
        F[i].r = F[i].r / Ntot * m;
        F[i].i = F[i].i / Ntot * m;
    }
    __syncthreads();
    const Cx<real>* R = stockham<real, +1>(F, G, tw, Nr, log2Nr, CB);
    for (int i = threadIdx.x; i < CB * Nr; i += blockDim.x) {
        const int r = i / CB, j = i - r * CB;
        typename Vec2<real>::type v;
        v.x = R[j * Nr + r].r;
        v.y = R[j * Nr + r].i;
        data[(size_t)r * Nc + c0 + j] = v;
    }
}

static int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

bool fft_supported(int N, size_t real_size) {
    if (N < 4 || (N & (N - 1))) return false;
    return 2 * (size_t)N * 2 * real_size <= 65536;
}

template <typename real, int CB>
static hipError_t launch_cols(typename Vec2<real>::type* psi, const void* tw_r, const real* mask, int Nc, int Nr, hipStream_t s) {
    const size_t lds = 2 * (size_t)CB * Nr * sizeof(Cx<real>);
    hipLaunchKernelGGL((fft_cols_filter<real, CB>), dim3(Nc / CB), dim3(kFftThreads), lds, s, psi, (const Cx<real>*)tw_r, mask, Nc, Nr,
                       ilog2(Nr), (real)((double)Nc * (double)Nr));
    return hipGetLastError();
}

template <typename real>
hipError_t launch_fft_filter(typename Vec2<real>::type* psi, const void* tw_c, const void* tw_r, const real* mask, int Nc, int Nr,
                             hipStream_t s) {
    const size_t row_lds = 2 * (size_t)Nc * sizeof(Cx<real>);
    hipLaunchKernelGGL((fft_rows<real, -1>), dim3(Nr), dim3(kFftThreads), row_lds, s, psi, (const Cx<real>*)tw_c, Nc, ilog2(Nc));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    /* widest column block whose two LDS buffers fit in 64 KiB */
    const size_t per_col = 2 * (size_t)Nr * sizeof(Cx<real>);
    if (per_col * 4 <= 65536 && Nc % 4 == 0)
        e = launch_cols<real, 4>(psi, tw_r, mask, Nc, Nr, s);
    else if (per_col * 2 <= 65536 && Nc % 2 == 0)
        e = launch_cols<real, 2>(psi, tw_r, mask, Nc, Nr, s);
    else
        e = launch_cols<real, 1>(psi, tw_r, mask, Nc, Nr, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fft_rows<real, +1>), dim3(Nr), dim3(kFftThreads), row_lds, s, psi, (const Cx<real>*)tw_c, Nc, ilog2(Nc));
    return hipGetLastError();
}

template hipError_t launch_fft_filter<float>(float2*, const void*, const void*, const float*, int, int, hipStream_t);
template hipError_t launch_fft_filter<double>(double2*, const void*, const void*, const double*, int, int, hipStream_t);

} // namespace phx
