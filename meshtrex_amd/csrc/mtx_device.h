// Device helpers shared by the kernel translation units: voxel keys, the
// activity rule, wave and workgroup sums, XCD-aware work order.
#pragma once

#include "mtx_internal.h"

namespace mtx {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t vkey(float f)
{
    // order-preserving map of IEEE-754 binary32 onto uint32
    uint32_t b = __float_as_uint(f);
    uint32_t m = (uint32_t)((int32_t)b >> 31) | 0x80000000u;
    return b ^ m;
}
__device__ __forceinline__ uint32_t vkey(uint8_t v) { return v; }

__device__ __forceinline__ float unkey(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(b);
}

// A pointer read from a descriptor in memory (a batch unit's volume, pairs or
// map) is generic to the compiler, and loads through it become flat loads,
// which also count against the 4-bit LDS/scalar counter and so cap the loads in
// flight.  Every such pointer is device global memory: rebuilding it as an
// address-space-1 pointer lets the loads derived from it be global ones.
template <typename T>
__device__ __forceinline__ T* as_global(T* p)
{
    using G = __attribute__((address_space(1))) T;
    return (T*)(G*)(uintptr_t)p;
}

// The argument block of a kernel whose only parameter is a T, read again from
// the kernel-argument segment at each use (scalar loads through the constant
// cache).  A long loop that keeps a dozen 64-bit argument pointers live
// otherwise holds them in scalar registers; past the scalar register budget
// they spill to vector-register lanes and every use costs two lane reads on
// the vector ALU.  The segment pointer comes from the dispatch (the only
// parameter sits at its offset 0), never from the parameter's address, which
// would make the compiler copy the block to private memory; the empty asm
// makes each use's address opaque, so the loads stay where they are used.
template <typename T>
struct ArgRef {
    using P = __attribute__((address_space(4))) const T*;
    __device__ __forceinline__ P operator->() const
    {
        P q = (P)__builtin_amdgcn_kernarg_segment_ptr();
        asm volatile("" : "+s"(q));
        return q;
    }
};

__device__ __forceinline__ float vfloat(float f) { return f; }
__device__ __forceinline__ float vfloat(uint8_t v) { return (float)v; }

// Activity of a block from its key pair (CPU rule of src/blockFilteringTestUtils.h:456-468,
// or the live shader's rule of src/shaders/occupiedBlockPrefixSum.comp.glsl:84).
__device__ __forceinline__ bool block_active(uint32_t lo, uint32_t hi, const ActParams& a)
{
    float flo = a.dtype == kU8 ? (float)lo : unkey(lo);
    float fhi = a.dtype == kU8 ? (float)hi : unkey(hi);
    if (a.rule == kRuleGpu) return a.iso >= flo && a.iso <= fhi;
    if (lo == hi) return false;
    if (a.dtype == kU8) return a.iso_r_valid && a.iso_r >= lo && a.iso_r <= hi;
    return a.iso >= flo && a.iso <= fhi;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Wave-wide scans on DPP lane moves (GFX9 controls, which CDNA keeps): shifts
// by 1, 2, 4, 8 lanes inside each row of 16, then the row broadcasts of lane
// 15 (into rows 1 and 3) and lane 31 (into rows 2 and 3).  A lane whose
// source is outside its row, or whose row the row mask leaves out, reads 0,
// the identity of the sums and unsigned maxima below.  The moves run on the
// VALU: no LDS permute and no per-step lane masks held in scalar registers.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <int CTRL, int ROWS, typename V>
__device__ __forceinline__ V dpp_or_zero(V v)
{
    if constexpr (sizeof(V) == 8) {
        const uint64_t u = (uint64_t)v;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xF, false);
        return (V)(((uint64_t)hi << 32) | lo);
    } else {
        return (V)(uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWS, 0xF, false);
    }
}

template <typename V, typename Op>
__device__ __forceinline__ V wave_incl_op(V v, Op op)
{
    v = op(v, dpp_or_zero<kDppRowShr1, 0xF>(v));
    v = op(v, dpp_or_zero<kDppRowShr2, 0xF>(v));
    v = op(v, dpp_or_zero<kDppRowShr4, 0xF>(v));
    v = op(v, dpp_or_zero<kDppRowShr8, 0xF>(v));
    v = op(v, dpp_or_zero<kDppRowBcast15, 0xA>(v));
    v = op(v, dpp_or_zero<kDppRowBcast31, 0xC>(v));
    return v;
}

// Value of lane 63 (a scalar register).
template <typename V>
__device__ __forceinline__ V wave_last(V v)
{
    if constexpr (sizeof(V) == 8) {
        const uint64_t u = (uint64_t)v;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
        return (V)(((uint64_t)hi << 32) | lo);
    } else {
        return (V)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    }
}

// Inclusive prefix sum over the wave's lanes (unsigned 32- or 64-bit; all lanes active).
template <typename V>
__device__ __forceinline__ V wave_incl_scan(V v)
{
    return wave_incl_op(v, [](V x, V y) { return (V)(x + y); });
}

// Maximum of v over the lanes before this one (0 for lane 0).
__device__ __forceinline__ uint32_t wave_excl_max(uint32_t v)
{
    const uint32_t m = wave_incl_op(v, [](uint32_t x, uint32_t y) { return max(x, y); });
    const uint32_t p = __shfl_up(m, 1, 64);
    return lane_id() == 0 ? 0u : p;
}

// Sum over the wave, the same in every lane.
template <typename V>
__device__ __forceinline__ V wave_sum(V v)
{
    return wave_last(wave_incl_scan<V>(v));
}

// XCD-aware remap: the dispatcher deals consecutive workgroup IDs round-robin
// over the 8 XCDs; give each XCD a contiguous range of logical work so that
// neighbouring tiles share one L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t wg, uint32_t nwg)
{
    if (nwg % 8u) return wg;
    return (wg % 8u) * (nwg / 8u) + wg / 8u;
}

// ---------------------------------------------------------------------------
// workgroup sums (256 threads)
// ---------------------------------------------------------------------------

// Sum of v over the workgroup, returned to every thread; s: 4 words of LDS.
template <typename V, typename S>
__device__ __forceinline__ S wg_sum256(V v, S* s)
{
    const S w = wave_sum<S>((S)v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = w;
    __syncthreads();
    const S r = s[0] + s[1] + s[2] + s[3];
    __syncthreads();
    return r;
}

// Exclusive prefix of tile t from the per-tile counts of an earlier launch:
// the workgroup adds up its predecessors' counts itself.
template <typename V>
__device__ __forceinline__ uint64_t tile_prefix(const V* __restrict__ counts, uint32_t t, uint64_t* s)
{
    uint64_t acc = 0;
    for (uint32_t k = threadIdx.x; k < t; k += 256) acc += counts[k];
    return wg_sum256<uint64_t, uint64_t>(acc, s);
}

}  // namespace mtx
