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

template <typename V>
__device__ __forceinline__ V wave_incl_scan(V v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        V t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Maximum of v over the lanes before this one (0 for lane 0).
__device__ __forceinline__ uint32_t wave_excl_max(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v = max(v, t);
    }
    const uint32_t p = __shfl_up(v, 1, 64);
    return lane == 0 ? 0u : p;
}

template <typename V>
__device__ __forceinline__ V wave_sum(V v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
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
