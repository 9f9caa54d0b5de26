// Device helpers shared by the kernel translation units: voxel keys, the
// activity rule, wave scans, XCD-aware work order, decoupled look-back.
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
// decoupled look-back (single-pass scan across workgroups)
// status word: bits 63:62 = 1 aggregate ready, 2 inclusive prefix ready
// ---------------------------------------------------------------------------

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t ld_acq(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by one whole wave; returns the exclusive prefix of `tile` (same in all lanes).
__device__ inline uint64_t lookback(uint64_t* status, uint32_t tile, uint64_t aggregate)
{
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) st_rel(&status[0], kFlagInc | aggregate);
        return 0;
    }
    if (lane == 0) st_rel(&status[tile], kFlagAgg | aggregate);
    uint64_t excl = 0;
    int64_t end = (int64_t)tile - 1;
    for (;;) {
        int64_t idx = end - (int64_t)lane;
        uint64_t s = idx >= 0 ? ld_acq(&status[idx]) : kFlagInc;
        uint64_t flag = s >> 62;
        if (__any(flag == 0)) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t inc = __ballot(flag == 2);
        uint64_t v = s & kValMask;
        if (inc) {
            uint32_t first = __builtin_ctzll(inc);
            excl += wave_sum<uint64_t>(lane <= first ? v : 0ull);
            break;
        }
        excl += wave_sum<uint64_t>(v);
        end -= 64;
    }
    if (lane == 0) st_rel(&status[tile], kFlagInc | (excl + aggregate));
    return excl;
}

}  // namespace mtx
