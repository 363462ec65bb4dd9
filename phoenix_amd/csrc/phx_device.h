/*
 * phx_device.h -- gfx950 device helpers shared by the streaming stencil kernels:
 * XCD-aware tile order, DPP wavefront shifts for east/west neighbours, raw
 * buffer loads whose out-of-range offsets read as zero.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phx {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

/* DPP controls of the GFX9 wavefront shifts: lane l of the result reads lane
   l+1 for wave_shl:1 and lane l-1 for wave_shr:1; the edge lane reads 0. */
constexpr int kDppWaveShl1 = 0x130;
constexpr int kDppWaveShr1 = 0x138;
/* quad_perm [1, 0, 3, 2]: lane l reads lane l ^ 1 (adjacent lanes swap) */
constexpr int kDppPairSwap = 0xB1;
/* buffer descriptor word 3 for GFX9-family raw buffers */
constexpr int kBufferFlags = 0x00020000;
/* byte offset beyond any row: raw buffer loads there return 0 */
constexpr uint32_t kOutOfRange = 0x80000000u;

/* Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
   Remap so that each XCD receives a contiguous range of tiles: horizontally
   adjacent tiles then share one L2 for their halo columns, and a tile stays on
   the same XCD from step to step. */
__device__ __forceinline__ int xcd_contiguous(int b, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int x = b & 7;
    const int start = x * q + (x < r ? x : r);
    return start + (b >> 3);
}

/* Periodic row wrap without branches, so that a fully unrolled row loop stays one
   basic block (a branch there splits it, and DPP moves then no longer fold into the
   adds that consume them).  x mod Nr for x = g + bias >= 0 is x - Nr * q with
   q = mulhi(x, ceil(2^32 / Nr)), exact whenever x * Nr < 2^32; g >= -halo, and
   bias = Nr (Nr >= halo) or Nr * halo keeps x below 3 Nr + the tile height. */
struct RowWrap {
    uint32_t nr, magic, bias;
};
__device__ __forceinline__ RowWrap make_row_wrap(int Nr, int halo) {
    RowWrap w;
    w.nr = (uint32_t)Nr;
    w.magic = 0xFFFFFFFFu / (uint32_t)Nr + 1u; /* ceil(2^32 / Nr) for Nr >= 2 */
    w.bias = Nr >= halo ? (uint32_t)Nr : (uint32_t)(Nr * halo);
    return w;
}
/* Row of tile row g: wrapped into [0, Nr) (exact modulo for periodic grids; for a
   zero boundary `in` says whether g lies in the grid, the wrapped row is then unused). */
__device__ __forceinline__ uint32_t tile_row(int g, const RowWrap& w, int periodic, bool& in) {
    in = periodic || (g >= 0 && g < (int)w.nr);
    const uint32_t x = (uint32_t)g + w.bias;
    return x - __umulhi(x, w.magic) * w.nr;
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    int2 x = __builtin_bit_cast(int2, v);
    x.x = __builtin_amdgcn_mov_dpp(x.x, CTRL, 0xF, 0xF, true);
    x.y = __builtin_amdgcn_mov_dpp(x.y, CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, x);
}
template <typename real>
__device__ __forceinline__ real east(real v) { return dpp_move<kDppWaveShl1>(v); }
template <typename real>
__device__ __forceinline__ real west(real v) { return dpp_move<kDppWaveShr1>(v); }

/* keep a uniform value in a vector register (an FMA may read one scalar operand only) */
template <typename T>
__device__ __forceinline__ T in_vgpr(T x) {
    asm volatile("" : "+v"(x));
    return x;
}

using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t row_rsrc(const void* row, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(row), 0, (int)bytes, kBufferFlags);
}

/* The vector results of the raw buffer builtins are reinterpreted as a whole:
   element access on them (v[1]) compiles to a single-dword load on this
   toolchain and returns the first element twice. */
template <typename real>
__device__ __forceinline__ real load_real(rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float load_real<float>(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double load_real<double>(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <typename real>
struct Cplx;
template <>
struct Cplx<float> {
    using type = float2;
};
template <>
struct Cplx<double> {
    using type = double2;
};
template <typename real>
__device__ __forceinline__ typename Cplx<real>::type load_cplx(rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float2 load_cplx<float>(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double2 load_cplx<double>(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

/* Loads and stores with a scalar byte offset (soffset) on top of the per-lane
   offset: one descriptor per array, the row offset in an SGPR.  Lanes whose
   per-lane offset is out of range read 0 and do not write.  CP is the cache policy:
   0 for ordinary cached accesses, kDeviceScope for data that other workgroups of the
   same launch write or read while it runs (no stale line in this CU's or this XCD's
   cache is used, and a completed store is visible to every other CU). */
constexpr int kDeviceScope = 16; /* sc1: device-scope coherence */
template <typename real, int CP = 0>
__device__ __forceinline__ real load_real_s(rsrc_t r, uint32_t off, uint32_t soff) {
    if constexpr (sizeof(real) == 4)
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, CP));
    else
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, CP));
}
template <typename real, int CP = 0>
__device__ __forceinline__ typename Cplx<real>::type load_cplx_s(rsrc_t r, uint32_t off, uint32_t soff) {
    if constexpr (sizeof(real) == 4)
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, CP));
    else
        return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, CP));
}
typedef int phx_v2i __attribute__((ext_vector_type(2)));
typedef int phx_v4i __attribute__((ext_vector_type(4)));
template <int CP = 0>
__device__ __forceinline__ void store_real_s(rsrc_t r, uint32_t off, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, off, soff, CP);
}
template <int CP = 0>
__device__ __forceinline__ void store_real_s(rsrc_t r, uint32_t off, uint32_t soff, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(phx_v2i, v), r, off, soff, CP);
}
template <int CP = 0>
__device__ __forceinline__ void store_cplx_s(rsrc_t r, uint32_t off, uint32_t soff, float2 v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(phx_v2i, v), r, off, soff, CP);
}
/* A buffer store of more than 8 bytes reads its data registers after it issues.  The
   compiler separates it from a following VALU write of those registers only when its
   soffset is not an SGPR; with an SGPR soffset the 16-byte store has written the
   overwriting value on gfx950 (the real part of a complex FP64 cell replaced by the next
   FMA's result, intermittently).  So the row offset joins the per-lane offset and soffset
   stays 0.  Out-of-range lanes stay out of range: kOutOfRange + soff < 2^32 for every
   soff < kMaxArrayBytes, and a descriptor of size 0 drops any offset. */
template <int CP = 0>
__device__ __forceinline__ void store_cplx_s(rsrc_t r, uint32_t off, uint32_t soff, double2 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(phx_v4i, v), r, off + soff, 0, CP);
}

/* Largest array the whole-array descriptors address: offsets are 32-bit and the
   out-of-range marker kOutOfRange must stay beyond every valid offset. */
constexpr size_t kMaxArrayBytes = 0x7FFFFFFFu;

} // namespace phx
