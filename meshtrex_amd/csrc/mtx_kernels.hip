// gfx950 kernels of the isosurface-extraction path: block min/max, active-block
// compaction, and PMB marching cubes (count, scan, emit).  Written for CDNA4:
// wave64, LDS tiles, XCD-aware work order; no matrix cores (nothing here is a
// dense contraction).
#include "mtx_device.h"
#include "mc_cases.h"

#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace mtx {


// ---------------------------------------------------------------------------
// block min/max, B = 8, overlap 1: the roofline kernel
//
// Lane l of a wave reads 4 consecutive voxels (16 B for f32, 4 B for u8) of a
// row, so one wave load is a fully coalesced 1 KiB (f32) row segment covering
// 32 x-blocks: a nontemporal load for aligned rows, a direct global->LDS load
// (read back aligned) for rows that are not 16-byte aligned.  Lanes 2k, 2k+1 cover x-block k; its overlap column x = 8k+8 is
// lane 2k+2's first voxel (the last block of the segment takes it from one
// wave-uniform load).  A wave walks M consecutive y-blocks within each plane,
// carrying the shared row 8(by+1) from block by into by+1, and sweeps a chunk
// of zc z-blocks, carrying the shared plane 8(bz+1) likewise: inside a wave
// every voxel is read once; only the rows/planes at wave boundaries are read
// twice (1/(8M) + 1/(8zc) of the volume).  The four waves of a workgroup take
// adjacent x segments (4 KiB contiguous per row) or, for narrow volumes,
// adjacent y groups.  Out-of-window rows/planes are clamped onto the last
// in-window one: re-reading it cannot change a min/max, and it keeps every
// load unconditional so a batch of 8 rows is in flight together.
// ---------------------------------------------------------------------------

struct MMAcc {
    uint32_t amin, amax;  // all 4 voxels of this lane
    uint32_t emin, emax;  // first voxel of this lane (overlap column for lane-2)
    uint32_t xmin, xmax;  // wave-uniform extra column
};

__device__ __forceinline__ void mm_merge(MMAcc& a, const MMAcc& b)
{
    a.amin = min(a.amin, b.amin); a.amax = max(a.amax, b.amax);
    a.emin = min(a.emin, b.emin); a.emax = max(a.emax, b.emax);
    a.xmin = min(a.xmin, b.xmin); a.xmax = max(a.xmax, b.xmax);
}

// Direct global -> LDS loads of 4 or 16 bytes per lane: lane l's bytes land at
// dst + l * size; dst is wave-uniform.  Completion is tracked by the
// vector-memory counter.
__device__ __forceinline__ void lds_direct_load4(const void* src, void* dst)
{
    __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(uintptr_t)src,
                                     (__attribute__((address_space(3))) void*)dst, 4, 0, 0);
}
__device__ __forceinline__ void lds_direct_load16(const void* src, void* dst)
{
    __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(uintptr_t)src,
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// 4 voxels of T per lane (16 bytes for f32, 4 for u8); the source needs no
// alignment beyond the voxel's.
template <typename T>
__device__ __forceinline__ void lds_direct_load_vox4(const T* src, T* dst)
{
    if constexpr (sizeof(T) == 4) lds_direct_load16(src, dst);
    else lds_direct_load4(src, dst);
}

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<uint8_t> { using type = uchar4; };

// 4 aligned voxels with a nontemporal load: the volume is streamed once, and
// measured on MI355X nontemporal loads stream faster than plain ones.
typedef float mtx_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load4_nt(const float* p, uint32_t (&k)[4])
{
    const mtx_f4v v = __builtin_nontemporal_load(reinterpret_cast<const mtx_f4v*>(p));
    k[0] = vkey(v.x); k[1] = vkey(v.y); k[2] = vkey(v.z); k[3] = vkey(v.w);
}
__device__ __forceinline__ void load4_nt(const uint8_t* p, uint32_t (&k)[4])
{
    const uint32_t v = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
    k[0] = v & 0xFFu; k[1] = (v >> 8) & 0xFFu; k[2] = (v >> 16) & 0xFFu; k[3] = v >> 24;
}

// Keys of this lane's 4 voxels of a row (neutral-min for voxels past X) and of
// the extra column.
template <typename T, bool ALIGNED>
__device__ __forceinline__ void mm_load(const T* __restrict__ row, uint32_t X, uint32_t x0, bool lane_ok,
                                        uint32_t xe, bool xe_ok, uint32_t (&k)[4], uint32_t& kx)
{
    if (ALIGNED) {
        load4_nt(row + (lane_ok ? x0 : 0u), k);
        if (!lane_ok) k[0] = k[1] = k[2] = k[3] = 0xFFFFFFFFu;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool ok = x0 + j < X;
            k[j] = vkey(row[ok ? x0 + j : 0u]);
            if (!ok) k[j] = 0xFFFFFFFFu;
        }
    }
    kx = vkey(row[xe_ok ? xe : 0u]);
}

template <bool ALIGNED>
__device__ __forceinline__ MMAcc mm_row(const uint32_t (&k)[4], uint32_t kx, uint32_t X, uint32_t x0,
                                        bool lane_ok, bool xe_ok)
{
    // neutral (all-ones) entries must not raise the max: mask them out
    const bool ok0 = ALIGNED ? lane_ok : (x0 < X);
    const bool ok1 = ALIGNED ? lane_ok : (x0 + 1 < X);
    const bool ok2 = ALIGNED ? lane_ok : (x0 + 2 < X);
    const bool ok3 = ALIGNED ? lane_ok : (x0 + 3 < X);
    MMAcc a;
    a.amin = min(min(k[0], k[1]), min(k[2], k[3]));
    a.amax = max(max(ok0 ? k[0] : 0u, ok1 ? k[1] : 0u), max(ok2 ? k[2] : 0u, ok3 ? k[3] : 0u));
    a.emin = k[0];
    a.emax = ok0 ? k[0] : 0u;
    a.xmin = xe_ok ? kx : 0xFFFFFFFFu;
    a.xmax = xe_ok ? kx : 0u;
    return a;
}

// One plane for M y-blocks of NB rows starting at row y0:
// out[k] = (merge ? out[k] (+) : ) rows of block k.
template <typename T, bool ALIGNED, int M, int NB = 8>
__device__ __forceinline__ void mm_plane(MMAcc (&out)[M], bool merge, const T* __restrict__ vol,
                                         uint64_t zoff, uint32_t X, uint32_t Y, uint32_t y0,
                                         uint32_t x0, bool lane_ok, uint32_t xe, bool xe_ok,
                                         T* stage = nullptr, bool zlast = false)
{
    uint32_t k0[4], kx0;
    mm_load<T, ALIGNED>(vol + zoff + (uint64_t)min(y0, Y - 1u) * X, X, x0, lane_ok, xe, xe_ok, k0, kx0);
    MMAcc carry = mm_row<ALIGNED>(k0, kx0, X, x0, lane_ok, xe_ok);
#pragma unroll
    for (int b = 0; b < M; ++b) {
        const uint32_t yb = y0 + (uint32_t)NB * b;
        const uint32_t ylast = min(yb + (uint32_t)NB, Y - 1u);
        uint32_t k[NB][4], kx[NB];
        if (ALIGNED || X < 4u) {
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const uint32_t y = min(yb + 1u + r, ylast);
                mm_load<T, ALIGNED>(vol + zoff + (uint64_t)y * X, X, x0, lane_ok, xe, xe_ok, k[r], kx[r]);
            }
        } else {
            // Unaligned rows through LDS: one 4-voxel direct load per lane and
            // row (the source needs only voxel alignment), read back aligned.
            // Lanes past the row's end read its start (their voxels are
            // neutral); the lane at the end reads on into the next row, except
            // on the volume's last row, where it reads the row's last 4 voxels
            // and shifts them into place.
            const bool tail = lane_ok && x0 + 4u > X;
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const uint32_t y = min(yb + 1u + r, ylast);
                const T* row = vol + zoff + (uint64_t)y * X;
                const bool vend = zlast && y == Y - 1u;
                const uint32_t xs = !lane_ok ? 0u : ((tail && vend) ? X - 4u : x0);
                lds_direct_load_vox4<T>(row + xs, stage + 256u * r);
                kx[r] = vkey(row[xe_ok ? xe : 0u]);
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the rows are in LDS
            using V4 = typename Vec4<T>::type;
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const V4 v = reinterpret_cast<const V4*>(stage + 256u * r)[lane_id()];
                uint32_t l0 = vkey(v.x), l1 = vkey(v.y), l2 = vkey(v.z), l3 = vkey(v.w);
                const uint32_t y = min(yb + 1u + r, ylast);
                if (zlast && y == Y - 1u && tail) {
                    // voxels X-4 .. X-1 were read; this lane's start at x0 - (X - 4)
                    const uint32_t sh = x0 - (X - 4u);
                    l0 = sh == 1u ? l1 : (sh == 2u ? l2 : l3);
                    l1 = sh == 1u ? l2 : l3;
                    l2 = l3;
                }
                k[r][0] = x0 < X ? l0 : 0xFFFFFFFFu;
                k[r][1] = x0 + 1u < X ? l1 : 0xFFFFFFFFu;
                k[r][2] = x0 + 2u < X ? l2 : 0xFFFFFFFFu;
                k[r][3] = x0 + 3u < X ? l3 : 0xFFFFFFFFu;
            }
        }
        MMAcc a = carry;
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const MMAcc row = mm_row<ALIGNED>(k[r], kx[r], X, x0, lane_ok, xe_ok);
            mm_merge(a, row);
            if (r == NB - 1) carry = row;
        }
        if (merge) mm_merge(out[b], a);
        else out[b] = a;
        // one batch of NB rows in flight at a time: keeps the register budget at
        // 4+ waves/SIMD instead of hoisting every block's loads
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <typename T, bool ALIGNED, bool MASK, int M>
__global__ __launch_bounds__(256) void k_minmax_b8(
    const T* __restrict__ vol, uint32_t X, uint32_t Y, uint32_t Z, uint32_t gx, uint32_t gy,
    uint32_t bzb, uint32_t bze, uint32_t ntx, uint32_t wx, uint32_t ntxg, uint32_t nygrp, uint32_t zc,
    uint2* __restrict__ pairs, uint32_t* __restrict__ mask, uint32_t wpr, ActParams act)
{
    // z-blocks [bzb, bze) of the volume (all of it, or one uploaded slab)
    // tile order: x groups fastest, then y groups, then z chunks; XCD-contiguous
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t txg = tile % ntxg;
    const uint32_t rest = tile / ntxg;
    const uint32_t ygrp = rest % nygrp;
    const uint32_t cz = rest / nygrp;
    const uint32_t lane = lane_id();
    // the wave index is made uniform only for the unaligned kernel: the
    // aligned one is measured fastest with the register allocation (2 waves
    // per SIMD) that the plain index gives
    const uint32_t w = ALIGNED ? threadIdx.x >> 6 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // unaligned rows are staged through LDS, 8 rows of 256 voxels per wave
    __shared__ __align__(16) T s_stage[ALIGNED ? 1 : 4][ALIGNED ? 1 : 8 * 256];
    T* stage = s_stage[ALIGNED ? 0u : w];
    const uint32_t wy = 4u / wx;
    const uint32_t tx = txg * wx + w % wx;
    const uint32_t by0 = (ygrp * wy + w / wx) * (uint32_t)M;
    if (tx >= ntx || by0 >= gy) return;
    const uint32_t bz0 = bzb + cz * zc;
    if (bz0 >= bze) return;
    const uint32_t bz1 = min(bz0 + zc, bze);

    const uint32_t x0 = tx * 256u + lane * 4u;
    const bool lane_ok = x0 < X;
    const uint32_t xe = tx * 256u + 256u;
    const bool xe_ok = xe < X;
    const uint32_t y0 = by0 * 8u;
    const uint64_t XY = (uint64_t)X * Y;

    MMAcc first[M];
    mm_plane<T, ALIGNED, M>(first, false, vol, (uint64_t)(bz0 * 8u) * XY, X, Y, y0, x0, lane_ok, xe, xe_ok,
                            stage, bz0 * 8u == Z - 1u);

    const uint32_t bx = tx * 32u + (lane >> 1);
    for (uint32_t bz = bz0; bz < bz1; ++bz) {
        MMAcc acc[M];
#pragma unroll
        for (int b = 0; b < M; ++b) acc[b] = first[b];
        const uint32_t zl = min(bz * 8u + 8u, Z - 1u);
#pragma unroll 1
        for (uint32_t dz = 1; dz < 8u; ++dz) {
            const uint32_t z = min(bz * 8u + dz, zl);
            mm_plane<T, ALIGNED, M>(acc, true, vol, (uint64_t)z * XY, X, Y, y0, x0, lane_ok, xe, xe_ok, stage,
                                    z == Z - 1u);
        }
        mm_plane<T, ALIGNED, M>(first, false, vol, (uint64_t)zl * XY, X, Y, y0, x0, lane_ok, xe, xe_ok, stage,
                                zl == Z - 1u);
#pragma unroll
        for (int b = 0; b < M; ++b) {
            mm_merge(acc[b], first[b]);
            const uint32_t by = by0 + b;
            // x combine: pair (2k, 2k+1) plus lane 2k+2's first voxel
            uint32_t bmin = min(acc[b].amin, __shfl_xor(acc[b].amin, 1, 64));
            uint32_t bmax = max(acc[b].amax, __shfl_xor(acc[b].amax, 1, 64));
            const uint32_t src = min(lane + 2u, 63u);
            uint32_t nmin = __shfl(acc[b].emin, (int)src, 64);
            uint32_t nmax = __shfl(acc[b].emax, (int)src, 64);
            if (lane >= 62u) { nmin = acc[b].xmin; nmax = acc[b].xmax; }
            bmin = min(bmin, nmin);
            bmax = max(bmax, nmax);
            const bool writer = !(lane & 1u) && bx < gx && by < gy;
            const uint64_t id = bx + (uint64_t)gx * (by + (uint64_t)gy * bz);
            if (writer) pairs[id] = make_uint2(bmin, bmax);
            if (MASK && by < gy) {
                uint64_t bal = __ballot(writer && block_active(bmin, bmax, act));
                // compress the even bits (one per x-block) into 32 bits
                uint64_t x = bal & 0x5555555555555555ull;
                x = (x | (x >> 1)) & 0x3333333333333333ull;
                x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
                x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
                x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
                x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
                if (lane == 0) mask[((uint64_t)bz * gy + by) * wpr + tx] = (uint32_t)x;
            }
        }
    }
}

// Volumes too small to fill the GPU with block columns: a workgroup per
// (256-voxel x tile, y block, z block); its four waves take planes 0-2, 3-4,
// 5-6 and 7-8 of the block row's nine and merge them through LDS, so a wave
// waits for at most three planes in a row instead of nine.
template <typename T, bool ALIGNED, bool MASK>
__global__ __launch_bounds__(256) void k_minmax_b8_zsplit(
    const T* __restrict__ vol, uint32_t X, uint32_t Y, uint32_t Z, uint32_t gx, uint32_t gy, uint32_t bzb,
    uint32_t bze, uint32_t ntx, uint2* __restrict__ pairs, uint32_t* __restrict__ mask, uint32_t wpr,
    ActParams act)
{
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = tile % ntx;
    const uint32_t rest = tile / ntx;
    const uint32_t by = rest % gy;
    const uint32_t bz = bzb + rest / gy;
    if (bz >= bze) return;  // the whole workgroup
    const uint32_t lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __shared__ __align__(16) T s_stage[ALIGNED ? 1 : 4][ALIGNED ? 1 : 8 * 256];
    __shared__ MMAcc s_acc[3][64];
    T* stage = s_stage[ALIGNED ? 0u : w];
    const uint32_t x0 = tx * 256u + lane * 4u;
    const bool lane_ok = x0 < X;
    const uint32_t xe = tx * 256u + 256u;
    const bool xe_ok = xe < X;
    const uint32_t y0 = by * 8u;
    const uint64_t XY = (uint64_t)X * Y;
    const uint32_t p0 = w == 0 ? 0u : 2u * w + 1u;  // 0, 3, 5, 7
    const uint32_t p1 = w == 0 ? 3u : p0 + 2u;       // 3, 5, 7, 9
    MMAcc acc[1];
#pragma unroll 1
    for (uint32_t p = p0; p < p1; ++p) {
        const uint32_t z = min(bz * 8u + p, Z - 1u);
        mm_plane<T, ALIGNED, 1>(acc, p != p0, vol, (uint64_t)z * XY, X, Y, y0, x0, lane_ok, xe, xe_ok, stage,
                                z == Z - 1u);
    }
    if (w) s_acc[w - 1][lane] = acc[0];
    __syncthreads();
    if (w) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) mm_merge(acc[0], s_acc[k][lane]);
    // x combine: pair (2k, 2k+1) plus lane 2k+2's first voxel
    const uint32_t bx = tx * 32u + (lane >> 1);
    uint32_t bmin = min(acc[0].amin, __shfl_xor(acc[0].amin, 1, 64));
    uint32_t bmax = max(acc[0].amax, __shfl_xor(acc[0].amax, 1, 64));
    const uint32_t src = min(lane + 2u, 63u);
    uint32_t nmin = __shfl(acc[0].emin, (int)src, 64);
    uint32_t nmax = __shfl(acc[0].emax, (int)src, 64);
    if (lane >= 62u) { nmin = acc[0].xmin; nmax = acc[0].xmax; }
    bmin = min(bmin, nmin);
    bmax = max(bmax, nmax);
    const bool writer = !(lane & 1u) && bx < gx;
    const uint64_t id = bx + (uint64_t)gx * (by + (uint64_t)gy * bz);
    if (writer) pairs[id] = make_uint2(bmin, bmax);
    if (MASK) {
        uint64_t bal = __ballot(writer && block_active(bmin, bmax, act));
        uint64_t x = bal & 0x5555555555555555ull;
        x = (x | (x >> 1)) & 0x3333333333333333ull;
        x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
        x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
        x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
        x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
        if (lane == 0) mask[((uint64_t)bz * gy + by) * wpr + tx] = (uint32_t)x;
    }
}

// Compute units of the current device.
static int device_cus()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    return n;
}

struct MMPlan {
    uint32_t M, wx, ntxg, nygrp, zc, ncz;
    uint64_t nwg;
    double cost;   // modelled seconds
};

// Model of a plan's duration: the larger of its HBM time -- the volume plus
// the rows and planes read twice, 1/(8M) + 1/(8zc) of it, at the measured
// nontemporal read rate -- and its latency time: a wave walks 8 zc + 1
// planes, each M + 1/2 batches of 8 rows in flight one after another, about
// 1.8 us per batch on MI355X (from a 247^3 volume at M = 4: 72.7 us for 41
// batches).  The latency term is what keeps small volumes from running on a
// few dozen workgroups.
constexpr double kMmReadBps = 6.8e12;
constexpr double kMmBatchSec = 1.8e-6;

// Pick y-blocks per wave (M) and the z chunk: every workgroup resident in one
// round (no tail), the smallest modelled duration.
static MMPlan plan_minmax(const Grid& g, uint32_t ntx, const int (&resident)[3], uint32_t rows, double bytes)
{
    const uint32_t wx = ntx >= 4 ? 4u : (ntx >= 2 ? 2u : 1u);
    MMPlan best{};
    double best_cost = 1e30;
    const uint32_t ms[3] = { 1, 2, 4 };
    for (int mi = 0; mi < 3; ++mi) {
        const uint32_t M = ms[mi];
        if (rows && rows != M) continue;
        const uint64_t target = (uint64_t)device_cus() * (uint64_t)(resident[mi] > 0 ? resident[mi] : 1);
        MMPlan p{};
        p.M = M;
        p.wx = wx;
        p.ntxg = (ntx + wx - 1) / wx;
        const uint32_t ygroups = (g.gy + M - 1) / M;
        p.nygrp = (ygroups + (4 / wx) - 1) / (4 / wx);
        const uint64_t cols = (uint64_t)p.ntxg * p.nygrp;
        uint64_t zc = (cols * g.gz + target - 1) / target;
        if (zc < 1) zc = 1;
        if (zc > g.gz) zc = g.gz;
        p.zc = (uint32_t)zc;
        p.ncz = (uint32_t)((g.gz + zc - 1) / zc);
        p.nwg = cols * p.ncz;
        const double t_hbm = bytes * (1.0 + 1.0 / (8.0 * M) + 1.0 / (8.0 * (double)zc)) / kMmReadBps;
        const double t_lat = (8.0 * (double)zc + 1.0) * ((double)M + 0.5) * kMmBatchSec;
        const double cost = t_hbm > t_lat ? t_hbm : t_lat;
        if (cost < best_cost) {
            best_cost = cost;
            best = p;
            best.cost = cost;
        }
    }
    return best;
}

// The plane-split kernel's modelled duration: three planes of 1.5 batches of
// rows per wave, one round per device-full of workgroups; the volume plus the
// shared row and plane of each block read twice.
static double zsplit_cost(const Grid& g, uint32_t ntx, int resident, double bytes)
{
    const uint64_t nwg = (uint64_t)ntx * g.gy * g.gz;
    const uint64_t per_round = (uint64_t)device_cus() * (uint64_t)(resident > 0 ? resident : 1);
    const double rounds = (double)((nwg + per_round - 1) / per_round);
    const double t_hbm = bytes * (1.0 + 1.0 / 8.0 + 1.0 / 8.0) / kMmReadBps;
    const double t_lat = rounds * 3.0 * 1.5 * kMmBatchSec;
    return t_hbm > t_lat ? t_hbm : t_lat;
}

template <typename T, bool ALIGNED, bool MASK, int M>
static hipError_t launch_mm_m(const VolDesc& v, const Grid& g, uint32_t bzb, uint32_t bze, const MMPlan& p,
                              uint32_t ntx, uint32_t* pairs, uint32_t* mask, const ActParams& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_minmax_b8<T, ALIGNED, MASK, M>), dim3((uint32_t)p.nwg), dim3(256), 0, s,
                       static_cast<const T*>(v.data), v.X, v.Y, v.Z, g.gx, g.gy, bzb, bze, ntx, p.wx,
                       p.ntxg, p.nygrp, p.zc, reinterpret_cast<uint2*>(pairs), mask, (g.gx + 31u) / 32u, a);
    return hipGetLastError();
}

template <typename T, bool MASK>
static hipError_t launch_mm_b8_t(const VolDesc& v, const Grid& g, uint32_t bzb, uint32_t bze, uint32_t* pairs,
                                 uint32_t* mask, const ActParams* act, uint32_t rows, hipStream_t s)
{
    const uint32_t ntx = (g.gx + 31u) / 32u;
    const bool aligned = (v.X % 4u) == 0 && (((uintptr_t)v.data) % (4 * sizeof(T))) == 0;
    // workgroups per CU of the aligned and the unaligned (LDS-staged) kernels
    static int resident[2][3] = { { 0, 0, 0 }, { 0, 0, 0 } };
    static int resident_zs[2] = { 0, 0 };
    if (!resident[0][0]) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident_zs[0], k_minmax_b8_zsplit<T, true, MASK>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident_zs[1], k_minmax_b8_zsplit<T, false, MASK>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident[0][0], k_minmax_b8<T, true, MASK, 1>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident[0][1], k_minmax_b8<T, true, MASK, 2>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident[0][2], k_minmax_b8<T, true, MASK, 4>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident[1][0], k_minmax_b8<T, false, MASK, 1>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident[1][1], k_minmax_b8<T, false, MASK, 2>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident[1][2], k_minmax_b8<T, false, MASK, 4>, 256, 0);
        for (auto& rr : resident)
            for (int& r : rr)
                if (r < 1) r = 1;
    }
    Grid gr = g;
    gr.gz = bze - bzb;  // the planned range
    const double bytes = (double)v.X * v.Y * (double)min(8u * gr.gz + 1u, v.Z) * sizeof(T);
    const MMPlan p = plan_minmax(gr, ntx, resident[aligned ? 0 : 1], rows, bytes);
    if (p.nwg == 0 || p.nwg > 0x7FFFFFFFull) return hipErrorInvalidValue;
    ActParams a{};
    if (act) a = *act;
    if (rows == 0 && zsplit_cost(gr, ntx, resident_zs[aligned ? 0 : 1], bytes) < p.cost) {
        const uint64_t nwg = (uint64_t)ntx * g.gy * (bze - bzb);
        if (nwg > 0x7FFFFFFFull) return hipErrorInvalidValue;
        auto* pr = reinterpret_cast<uint2*>(pairs);
        const uint32_t wpr = (g.gx + 31u) / 32u;
        if (aligned)
            hipLaunchKernelGGL((k_minmax_b8_zsplit<T, true, MASK>), dim3((uint32_t)nwg), dim3(256), 0, s,
                               static_cast<const T*>(v.data), v.X, v.Y, v.Z, g.gx, g.gy, bzb, bze, ntx, pr, mask,
                               wpr, a);
        else
            hipLaunchKernelGGL((k_minmax_b8_zsplit<T, false, MASK>), dim3((uint32_t)nwg), dim3(256), 0, s,
                               static_cast<const T*>(v.data), v.X, v.Y, v.Z, g.gx, g.gy, bzb, bze, ntx, pr, mask,
                               wpr, a);
        return hipGetLastError();
    }
#define MTX_MM_CASE(MM)                                                                            \
    if (p.M == MM)                                                                                 \
        return aligned ? launch_mm_m<T, true, MASK, MM>(v, g, bzb, bze, p, ntx, pairs, mask, a, s) \
                       : launch_mm_m<T, false, MASK, MM>(v, g, bzb, bze, p, ntx, pairs, mask, a, s);
    MTX_MM_CASE(1)
    MTX_MM_CASE(2)
    MTX_MM_CASE(4)
#undef MTX_MM_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_minmax_b8(const VolDesc& v, const Grid& g, uint32_t bzb, uint32_t bze, uint32_t* pairs,
                            uint32_t* mask, const ActParams* act, uint32_t rows, hipStream_t s)
{
    if (rows != 0 && rows != 1 && rows != 2 && rows != 4) return hipErrorInvalidValue;
    if (bzb >= bze || bze > g.gz) return hipErrorInvalidValue;
    const bool m = mask != nullptr;
    if (v.dtype == kF32)
        return m ? launch_mm_b8_t<float, true>(v, g, bzb, bze, pairs, mask, act, rows, s)
                 : launch_mm_b8_t<float, false>(v, g, bzb, bze, pairs, mask, act, rows, s);
    return m ? launch_mm_b8_t<uint8_t, true>(v, g, bzb, bze, pairs, mask, act, rows, s)
             : launch_mm_b8_t<uint8_t, false>(v, g, bzb, bze, pairs, mask, act, rows, s);
}

// ---------------------------------------------------------------------------
// batched block min/max, B = 4, 8 or 16, overlap 1
//
// The row walk of k_minmax_b8 generalised to B/4 lanes per x-block (a wave
// segment of 256 voxels holds 64, 32 or 16 x-blocks) and to many volumes per
// launch: each volume has its own launch plan and a contiguous range of
// workgroups; a workgroup finds its volume with one wave ballot per 64
// volumes.  Volumes whose rows are not 16-byte aligned take scalar loads.
// ---------------------------------------------------------------------------

template <typename T, bool ALIGNED, int B, int M>
__device__ __forceinline__ void mm_batch_body(const BatchVolume& V, uint32_t tile, T* stage)
{
    constexpr uint32_t L = B / 4;       // lanes per x-block
    constexpr uint32_t BPW = 64 / L;    // x-blocks per wave segment
    const T* __restrict__ vol = as_global(static_cast<const T*>(V.data));
    const uint32_t X = V.X, Y = V.Y, Z = V.Z;
    const uint32_t txg = tile % V.ntxg;
    const uint32_t rest = tile / V.ntxg;
    const uint32_t ygrp = rest % V.nygrp;
    const uint32_t cz = rest / V.nygrp;
    const uint32_t lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t wy = 4u / V.wx;
    const uint32_t tx = txg * V.wx + w % V.wx;
    const uint32_t by0 = (ygrp * wy + w / V.wx) * (uint32_t)M;
    if (tx >= V.ntx || by0 >= V.gy) return;
    const uint32_t bz0 = cz * V.zc;
    if (bz0 >= V.gz) return;
    const uint32_t bz1 = min(bz0 + V.zc, V.gz);

    const uint32_t x0 = tx * 256u + lane * 4u;
    const bool lane_ok = x0 < X;
    const uint32_t xe = tx * 256u + 256u;
    const bool xe_ok = xe < X;
    const uint32_t y0 = by0 * (uint32_t)B;
    const uint64_t XY = (uint64_t)X * Y;
    uint2* __restrict__ pairs = as_global(reinterpret_cast<uint2*>(V.pairs));

    MMAcc first[M];
    mm_plane<T, ALIGNED, M, B>(first, false, vol, (uint64_t)(bz0 * B) * XY, X, Y, y0, x0, lane_ok, xe, xe_ok,
                               stage, bz0 * B == Z - 1u);
    const uint32_t bx = tx * BPW + lane / L;
    for (uint32_t bz = bz0; bz < bz1; ++bz) {
        MMAcc acc[M];
#pragma unroll
        for (int b = 0; b < M; ++b) acc[b] = first[b];
        const uint32_t zl = min(bz * B + B, Z - 1u);
#pragma unroll 1
        for (uint32_t dz = 1; dz < (uint32_t)B; ++dz) {
            const uint32_t z = min(bz * B + dz, zl);
            mm_plane<T, ALIGNED, M, B>(acc, true, vol, (uint64_t)z * XY, X, Y, y0, x0, lane_ok, xe, xe_ok, stage,
                                       z == Z - 1u);
        }
        mm_plane<T, ALIGNED, M, B>(first, false, vol, (uint64_t)zl * XY, X, Y, y0, x0, lane_ok, xe, xe_ok, stage,
                                   zl == Z - 1u);
#pragma unroll
        for (int b = 0; b < M; ++b) {
            mm_merge(acc[b], first[b]);
            const uint32_t by = by0 + b;
            // x combine: the L lanes of the block plus the first voxel of lane + L
            uint32_t bmin = acc[b].amin, bmax = acc[b].amax;
#pragma unroll
            for (uint32_t d = 1; d < L; d <<= 1) {
                bmin = min(bmin, (uint32_t)__shfl_xor(bmin, (int)d, 64));
                bmax = max(bmax, (uint32_t)__shfl_xor(bmax, (int)d, 64));
            }
            const uint32_t src = min(lane + L, 63u);
            uint32_t nmin = __shfl(acc[b].emin, (int)src, 64);
            uint32_t nmax = __shfl(acc[b].emax, (int)src, 64);
            if (lane >= 64u - L) { nmin = acc[b].xmin; nmax = acc[b].xmax; }
            bmin = min(bmin, nmin);
            bmax = max(bmax, nmax);
            if (lane % L == 0u && bx < V.gx && by < V.gy)
                pairs[bx + (uint64_t)V.gx * (by + (uint64_t)V.gy * bz)] = make_uint2(bmin, bmax);
        }
    }
}

template <typename T, int B, int M>
__global__ __launch_bounds__(256) void k_minmax_batch(const BatchVolume* __restrict__ vols,
                                                      const uint32_t* __restrict__ wg0, uint32_t nvol)
{
    __shared__ uint32_t s_vol;
    // unaligned rows are staged through LDS, B rows of 256 voxels per wave
    __shared__ __align__(16) T s_stage[4][B * 256];
    const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
    if (threadIdx.x < 64) {
        // the last volume whose first workgroup is <= wg (wg0 ascending, wg0[0] = 0)
        uint32_t n = 0;
        for (uint32_t c = 0; c < nvol; c += 64) {
            const uint32_t k = c + threadIdx.x;
            n += (uint32_t)__popcll(__ballot(k < nvol && wg0[k] <= wg));
        }
        if (threadIdx.x == 0) s_vol = n - 1;
    }
    __syncthreads();
    const uint32_t v = __builtin_amdgcn_readfirstlane(s_vol);
    const BatchVolume& V = vols[v];
    const uint32_t tile = wg - V.wg0;
    T* stage = s_stage[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    if (V.aligned) mm_batch_body<T, true, B, M>(V, tile, stage);
    else mm_batch_body<T, false, B, M>(V, tile, stage);
}

template <typename T, int B, int M>
static int mm_batch_resident()
{
    static int r = 0;
    if (!r) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, k_minmax_batch<T, B, M>, 256, 0);
        if (r < 1) r = 1;
    }
    return r;
}

static int mm_batch_resident_for(int dtype, uint32_t B, uint32_t M)
{
#define MTX_RES(BB, MM)                                                                          \
    if (B == BB && M == MM)                                                                      \
        return dtype == kF32 ? mm_batch_resident<float, BB, MM>() : mm_batch_resident<uint8_t, BB, MM>();
    MTX_RES(4, 1) MTX_RES(4, 2) MTX_RES(4, 4)
    MTX_RES(8, 1) MTX_RES(8, 2) MTX_RES(8, 4)
    MTX_RES(16, 1) MTX_RES(16, 2) MTX_RES(16, 4)
#undef MTX_RES
    return 1;
}

uint64_t plan_minmax_batch(BatchVolume* vols, uint32_t nvol, int dtype, uint32_t B, uint32_t M)
{
    if (!batch_supported_block(B) || (M != 1 && M != 2 && M != 4) || nvol == 0) return 0;
    const uint64_t target = (uint64_t)device_cus() * (uint64_t)mm_batch_resident_for(dtype, B, M);
    uint32_t maxgz = 1;
    for (uint32_t i = 0; i < nvol; ++i) {
        BatchVolume& v = vols[i];
        v.ntx = (v.X + 255u) / 256u;
        v.wx = v.ntx >= 4 ? 4u : (v.ntx >= 2 ? 2u : 1u);
        v.ntxg = (v.ntx + v.wx - 1) / v.wx;
        const uint32_t ygroups = (v.gy + M - 1) / M;
        v.nygrp = (ygroups + (4 / v.wx) - 1) / (4 / v.wx);
        maxgz = v.gz > maxgz ? v.gz : maxgz;
    }
    // the smallest z chunk that keeps every workgroup resident in one round
    uint32_t zc = 1;
    for (; zc < maxgz; ++zc) {
        uint64_t n = 0;
        for (uint32_t i = 0; i < nvol; ++i)
            n += (uint64_t)vols[i].ntxg * vols[i].nygrp * ((vols[i].gz + zc - 1) / zc);
        if (n <= target) break;
    }
    uint64_t wg = 0;
    for (uint32_t i = 0; i < nvol; ++i) {
        BatchVolume& v = vols[i];
        v.zc = zc < v.gz ? zc : v.gz;
        v.wg0 = (uint32_t)wg;
        wg += (uint64_t)v.ntxg * v.nygrp * ((v.gz + v.zc - 1) / v.zc);
    }
    return wg > 0x7FFFFFFFull ? 0 : wg;
}

hipError_t launch_minmax_batch(const BatchVolume* vols, const uint32_t* wg0, uint32_t nvol, int dtype,
                               uint32_t B, uint32_t M, uint64_t nwg, hipStream_t s)
{
    if (nwg == 0 || nwg > 0x7FFFFFFFull) return hipErrorInvalidValue;
#define MTX_MMB(BB, MM)                                                                          \
    if (B == BB && M == MM) {                                                                    \
        if (dtype == kF32)                                                                       \
            hipLaunchKernelGGL((k_minmax_batch<float, BB, MM>), dim3((uint32_t)nwg), dim3(256), 0, s, vols, wg0, nvol); \
        else                                                                                     \
            hipLaunchKernelGGL((k_minmax_batch<uint8_t, BB, MM>), dim3((uint32_t)nwg), dim3(256), 0, s, vols, wg0, nvol); \
        return hipGetLastError();                                                                \
    }
    MTX_MMB(4, 1) MTX_MMB(4, 2) MTX_MMB(4, 4)
    MTX_MMB(8, 1) MTX_MMB(8, 2) MTX_MMB(8, 4)
    MTX_MMB(16, 1) MTX_MMB(16, 2) MTX_MMB(16, 4)
#undef MTX_MMB
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// block min/max, any B, overlap 0 or 1: one wave per block
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_minmax_generic(const T* __restrict__ vol, uint32_t X,
                                                        uint32_t Y, uint32_t Z, uint32_t B,
                                                        uint32_t gx, uint32_t gy, uint64_t nb,
                                                        uint32_t overlap, uint2* __restrict__ pairs)
{
    const uint64_t blk = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
    if (blk >= nb) return;
    const uint32_t lane = lane_id();
    const uint32_t bx = (uint32_t)(blk % gx);
    const uint32_t by = (uint32_t)((blk / gx) % gy);
    const uint32_t bz = (uint32_t)(blk / ((uint64_t)gx * gy));
    const uint32_t span = overlap ? B : B - 1u;
    const uint32_t xa = bx * B, ya = by * B, za = bz * B;
    const uint32_t xb = min(xa + span, X - 1u), yb = min(ya + span, Y - 1u), zb = min(za + span, Z - 1u);
    const uint32_t wx = xb - xa + 1u, wy = yb - ya + 1u, wz = zb - za + 1u;
    const uint32_t n = wx * wy * wz;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t i = lane; i < n; i += 64u) {
        uint32_t x = xa + i % wx;
        uint32_t y = ya + (i / wx) % wy;
        uint32_t z = za + i / (wx * wy);
        uint32_t k = vkey(vol[x + (uint64_t)X * (y + (uint64_t)Y * z)]);
        lo = min(lo, k);
        hi = max(hi, k);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor(lo, d, 64));
        hi = max(hi, (uint32_t)__shfl_xor(hi, d, 64));
    }
    if (lane == 0) pairs[blk] = make_uint2(lo, hi);
}

hipError_t launch_minmax_generic(const VolDesc& v, const Grid& g, int overlap, uint32_t* pairs,
                                 hipStream_t s)
{
    const uint64_t nb = g.nb();
    const uint64_t nwg = (nb + 3) / 4;
    if (nwg > 0x7FFFFFFFull) return hipErrorInvalidValue;
    uint2* pr = reinterpret_cast<uint2*>(pairs);
    if (v.dtype == kF32)
        hipLaunchKernelGGL(k_minmax_generic<float>, dim3((uint32_t)nwg), dim3(256), 0, s,
                           static_cast<const float*>(v.data), v.X, v.Y, v.Z, g.B, g.gx, g.gy, nb,
                           (uint32_t)overlap, pr);
    else
        hipLaunchKernelGGL(k_minmax_generic<uint8_t>, dim3((uint32_t)nwg), dim3(256), 0, s,
                           static_cast<const uint8_t*>(v.data), v.X, v.Y, v.Z, g.B, g.gx, g.gy,
                           nb, (uint32_t)overlap, pr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// min/max pyramid: each workgroup reduces an 8x8x8 tile of one level into up
// to three coarser levels (4^3, 2^3, 1) through LDS
// ---------------------------------------------------------------------------

struct PyrLevels {
    uint64_t off[4];
    uint32_t ext[4][3];
    uint32_t n;  // number of destination levels (1..3)
};

__global__ __launch_bounds__(512) void k_pyramid(uint2* __restrict__ pairs, PyrLevels L)
{
    __shared__ uint2 s[512];
    const uint32_t t = threadIdx.x;
    const uint32_t lx = t & 7u, ly = (t >> 3) & 7u, lz = t >> 6;
    const uint32_t ntx = (L.ext[0][0] + 7u) / 8u, nty = (L.ext[0][1] + 7u) / 8u;
    const uint32_t tx = blockIdx.x % ntx, ty = (blockIdx.x / ntx) % nty, tz = blockIdx.x / (ntx * nty);
    {
        uint32_t x = tx * 8u + lx, y = ty * 8u + ly, z = tz * 8u + lz;
        uint2 v = make_uint2(0xFFFFFFFFu, 0u);
        if (x < L.ext[0][0] && y < L.ext[0][1] && z < L.ext[0][2])
            v = pairs[L.off[0] + x + (uint64_t)L.ext[0][0] * (y + (uint64_t)L.ext[0][1] * z)];
        s[t] = v;
    }
    for (uint32_t l = 1; l <= L.n; ++l) {
        const uint32_t size = 8u >> l;
        __syncthreads();
        uint2 r = make_uint2(0xFFFFFFFFu, 0u);
        const bool act = lx < size && ly < size && lz < size;
        if (act) {
#pragma unroll
            for (uint32_t c = 0; c < 8u; ++c) {
                uint32_t cx = 2u * lx + (c & 1u), cy = 2u * ly + ((c >> 1) & 1u), cz = 2u * lz + (c >> 2);
                uint2 v = s[cx + 8u * (cy + 8u * cz)];
                r.x = min(r.x, v.x);
                r.y = max(r.y, v.y);
            }
        }
        __syncthreads();
        if (act) {
            s[lx + 8u * (ly + 8u * lz)] = r;
            uint32_t x = ((tx * 8u) >> l) + lx, y = ((ty * 8u) >> l) + ly, z = ((tz * 8u) >> l) + lz;
            if (x < L.ext[l][0] && y < L.ext[l][1] && z < L.ext[l][2])
                pairs[L.off[l] + x + (uint64_t)L.ext[l][0] * (y + (uint64_t)L.ext[l][1] * z)] = r;
        }
    }
}

hipError_t launch_pyramid(uint32_t* pairs, uint32_t levels, const uint32_t (*ext)[3],
                          const uint64_t* off, hipStream_t s)
{
    uint2* pr = reinterpret_cast<uint2*>(pairs);
    for (uint32_t src = 0; src + 1 < levels; src += 3) {
        PyrLevels L{};
        L.n = min(3u, levels - 1 - src);
        for (uint32_t j = 0; j <= L.n; ++j) {
            L.off[j] = off[src + j];
            L.ext[j][0] = ext[src + j][0];
            L.ext[j][1] = ext[src + j][1];
            L.ext[j][2] = ext[src + j][2];
        }
        uint64_t nwg = (uint64_t)((L.ext[0][0] + 7) / 8) * ((L.ext[0][1] + 7) / 8) * ((L.ext[0][2] + 7) / 8);
        hipLaunchKernelGGL(k_pyramid, dim3((uint32_t)nwg), dim3(512), 0, s, pr, L);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// ordered compaction of active block IDs
//
// Reduce-then-scatter in two launches, with no waits between workgroups: the
// first counts the active blocks of every tile, the second lets each tile sum
// its predecessors' counts itself (at most a few thousand cached words) and
// scatter its IDs.  On a chip of 8 XCDs with separate L2s, polling another
// workgroup's published prefix costs a coherent round trip per look-back
// window; reading the finished counts of a previous launch does not.
// ---------------------------------------------------------------------------

constexpr uint32_t kFilterItems = 8;                     // blocks per thread (pairs path)
constexpr uint32_t kFilterTile = 256 * kFilterItems;     // blocks per tile (pairs path)
constexpr uint32_t kMaskTileWords = 64;                  // mask words per tile (one per lane of wave 0)

uint32_t filter_tiles(uint64_t nb, bool from_mask, uint32_t words_per_row, const Grid& g)
{
    uint64_t tiles;
    if (from_mask) tiles = ((uint64_t)g.gy * g.gz * words_per_row + kMaskTileWords - 1) / kMaskTileWords;
    else tiles = (nb + kFilterTile - 1) / kFilterTile;
    return (uint32_t)(tiles ? tiles : 1);
}

// Coarse reject of the pairs path: the pyramid level holding 4^3-block
// ancestors (level 2), or none.
struct Coarse {
    const uint2* pairs;   // level-2 pairs, or nullptr
    uint32_t gx, gy;      // level-0 grid (x, y)
    uint32_t cx, cy;      // level-2 extent (x, y)
};

// Active flags of this thread's 8 blocks of `tile` as wave ballots, and the
// wave's counts in s_cnt[k * 4 + wave].  With a coarse level, a block under an
// inactive ancestor is inactive without reading its own pair: the ancestor's
// range contains the block's, so under either rule it is active whenever the
// block is (src/shaders/occupiedBlockPrefixSum.comp.glsl:47-63).
__device__ __forceinline__ void tile_ballots(const uint2* __restrict__ pairs, uint64_t nb, uint64_t base,
                                             const ActParams& act, uint64_t (&bal)[kFilterItems],
                                             uint32_t* s_cnt, const Coarse& cr = Coarse{})
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t k = 0; k < kFilterItems; ++k) {
        const uint64_t b = base + k * 256u + threadIdx.x;
        bool a = false;
        if (b < nb) {
            bool live = true;
            if (cr.pairs) {
                const uint32_t bx = (uint32_t)(b % cr.gx), by = (uint32_t)((b / cr.gx) % cr.gy);
                const uint32_t bz = (uint32_t)(b / ((uint64_t)cr.gx * cr.gy));
                const uint2 c = cr.pairs[(bx >> 2) + cr.cx * ((by >> 2) + (uint64_t)cr.cy * (bz >> 2))];
                live = block_active(c.x, c.y, act);
            }
            if (live) {
                const uint2 p = pairs[b];
                a = block_active(p.x, p.y, act);
            }
        }
        bal[k] = __ballot(a);
        if (lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
    }
}

__global__ __launch_bounds__(256) void k_filter_pairs_count(const uint2* __restrict__ pairs, uint64_t nb,
                                                            ActParams act, uint32_t* __restrict__ tile_count,
                                                            Coarse cr)
{
    __shared__ uint32_t s_cnt[kFilterItems * 4];
    uint64_t bal[kFilterItems];
    tile_ballots(pairs, nb, (uint64_t)blockIdx.x * kFilterTile, act, bal, s_cnt, cr);
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t v = threadIdx.x < kFilterItems * 4 ? s_cnt[threadIdx.x] : 0u;
        const uint32_t t = wave_sum<uint32_t>(v);
        if (threadIdx.x == 0) tile_count[blockIdx.x] = t;
    }
}

// Scatter of one tile's IDs (block-ID order: k, wave, lane) at `excl`; writes
// the dense map of the tile's blocks.  s_cnt: the tile's per-(k, wave) counts.
__device__ __forceinline__ void tile_scatter(const uint64_t (&bal)[kFilterItems], uint32_t* s_cnt, uint64_t base,
                                             uint64_t nb, uint64_t excl, uint32_t* __restrict__ ids,
                                             uint32_t* __restrict__ item_unit, uint32_t unit,
                                             uint32_t* __restrict__ map)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t k = 0; k < kFilterItems; ++k) {
        const uint64_t b = base + k * 256u + threadIdx.x;
        uint32_t slot = 0xFFFFFFFFu;
        if ((bal[k] >> lane) & 1ull) {
            const uint64_t pos = excl + s_cnt[k * 4 + w] + (uint32_t)__popcll(bal[k] & lt);
            ids[pos] = (uint32_t)b;
            if (item_unit) item_unit[pos] = unit;
            slot = (uint32_t)pos;
        }
        if (b < nb) map[b] = slot;
    }
}

// Exclusive scan of the tile's 32 (k, wave) counts in place; returns the tile total.
__device__ __forceinline__ uint32_t tile_local_scan(uint32_t* s_cnt)
{
    uint32_t tot = 0;
    if (threadIdx.x < 64) {
        const uint32_t v = threadIdx.x < kFilterItems * 4 ? s_cnt[threadIdx.x] : 0u;
        const uint32_t inc = wave_incl_scan<uint32_t>(v);
        tot = __shfl(inc, 63, 64);
        if (threadIdx.x < kFilterItems * 4) s_cnt[threadIdx.x] = inc - v;
    }
    return tot;
}

__global__ __launch_bounds__(256) void k_filter_pairs_scatter(const uint2* __restrict__ pairs, uint64_t nb,
                                                              ActParams act, uint32_t* __restrict__ ids,
                                                              uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ map,
                                                              const uint32_t* __restrict__ tile_count,
                                                              uint32_t ntiles, Coarse cr)
{
    __shared__ uint32_t s_cnt[kFilterItems * 4];
    __shared__ uint64_t s_red[4];
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * kFilterTile;
    const uint64_t excl = tile_prefix<uint32_t>(tile_count, tile, s_red);
    uint64_t bal[kFilterItems];
    tile_ballots(pairs, nb, base, act, bal, s_cnt, cr);
    __syncthreads();
    const uint32_t tot = tile_local_scan(s_cnt);
    if (threadIdx.x == 0 && tile == ntiles - 1) *count = (uint32_t)(excl + tot);
    __syncthreads();
    tile_scatter(bal, s_cnt, base, nb, excl, ids, nullptr, 0, map);
}

hipError_t launch_filter_pairs(const uint32_t* pairs, const Grid& g, const ActParams& act,
                               uint32_t* ids, uint32_t* count, uint32_t* map, TileSums ts,
                               const uint32_t* coarse, const uint32_t (*coarse_ext)[3], hipStream_t s)
{
    const uint64_t nb = g.nb();
    const uint32_t ntiles = filter_tiles(nb, false, 0, g);
    const uint2* pr = reinterpret_cast<const uint2*>(pairs);
    uint32_t* tc = reinterpret_cast<uint32_t*>(ts.sums);
    Coarse cr{};
    if (coarse) {
        cr.pairs = reinterpret_cast<const uint2*>(coarse);
        cr.gx = g.gx; cr.gy = g.gy;
        cr.cx = (*coarse_ext)[0]; cr.cy = (*coarse_ext)[1];
    }
    hipLaunchKernelGGL(k_filter_pairs_count, dim3(ntiles), dim3(256), 0, s, pr, nb, act, tc, cr);
    hipLaunchKernelGGL(k_filter_pairs_scatter, dim3(ntiles), dim3(256), 0, s, pr, nb, act, ids, count, map, tc,
                       ntiles, cr);
    return hipGetLastError();
}

// From the fused bit mask: one 32-bit word (32 x-blocks of one row) per thread.
__global__ __launch_bounds__(256) void k_mask_count(const uint32_t* __restrict__ mask, uint64_t nwords,
                                                    uint32_t* __restrict__ tile_count)
{
    if (threadIdx.x >= 64) return;
    const uint64_t wi = (uint64_t)blockIdx.x * kMaskTileWords + threadIdx.x;
    const uint32_t c = wi < nwords ? (uint32_t)__popc(mask[wi]) : 0u;
    const uint32_t t = wave_sum<uint32_t>(c);
    if (threadIdx.x == 0) tile_count[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_mask_scatter(const uint32_t* __restrict__ mask, uint64_t nwords,
                                                      uint32_t gx, uint32_t wpr, uint32_t* __restrict__ ids,
                                                      uint32_t* __restrict__ count, uint32_t* __restrict__ map,
                                                      const uint32_t* __restrict__ tile_count, uint32_t ntiles)
{
    __shared__ uint64_t s_red[4];
    __shared__ uint32_t s_word[kMaskTileWords], s_pref[kMaskTileWords];
    const uint32_t tile = blockIdx.x;
    const uint64_t excl = tile_prefix<uint32_t>(tile_count, tile, s_red);
    if (threadIdx.x < 64) {
        // wave 0: the tile's words and the tile-relative index of each word's
        // first active block
        const uint64_t wi = (uint64_t)tile * kMaskTileWords + threadIdx.x;
        const uint32_t word = wi < nwords ? mask[wi] : 0u;
        const uint32_t c = (uint32_t)__popc(word);
        const uint32_t inc = wave_incl_scan<uint32_t>(c);
        s_word[threadIdx.x] = word;
        s_pref[threadIdx.x] = inc - c;
        if (threadIdx.x == 63 && tile == ntiles - 1) *count = (uint32_t)(excl + inc);
    }
    __syncthreads();
    // blocks of the tile in order, 256 at a time: consecutive lanes write
    // consecutive map entries (and the active ones consecutive IDs)
    for (uint32_t t = threadIdx.x; t < kMaskTileWords * 32u; t += 256u) {
        const uint32_t wl = t >> 5, bit = t & 31u;
        const uint64_t wg = (uint64_t)tile * kMaskTileWords + wl;
        if (wg >= nwords) break;
        const uint32_t x = (uint32_t)(wg % wpr) * 32u + bit;
        if (x >= gx) continue;   // padding bits past the row's end
        const uint32_t wd = s_word[wl];
        const uint64_t blk = (wg / wpr) * gx + x;
        uint32_t slot = 0xFFFFFFFFu;
        if ((wd >> bit) & 1u) {
            const uint64_t pos = excl + s_pref[wl] + (uint32_t)__popc(wd & ((1u << bit) - 1u));
            ids[pos] = (uint32_t)blk;
            slot = (uint32_t)pos;
        }
        map[blk] = slot;
    }
}

hipError_t launch_compact_mask(const uint32_t* mask, const Grid& g, uint32_t words_per_row,
                               uint32_t* ids, uint32_t* count, uint32_t* map, TileSums ts,
                               hipStream_t s)
{
    const uint64_t nwords = (uint64_t)g.gy * g.gz * words_per_row;
    const uint32_t ntiles = filter_tiles(g.nb(), true, words_per_row, g);
    uint32_t* tc = reinterpret_cast<uint32_t*>(ts.sums);
    hipLaunchKernelGGL(k_mask_count, dim3(ntiles), dim3(256), 0, s, mask, nwords, tc);
    hipLaunchKernelGGL(k_mask_scatter, dim3(ntiles), dim3(256), 0, s, mask, nwords, g.gx, words_per_row, ids,
                       count, map, tc, ntiles);
    return hipGetLastError();
}

// Batched form: one compaction over the tiles of every unit of a class, in
// unit order, so that the concatenated list is ordered by (unit, block ID).
// Each unit's blocks start on a tile boundary; a tile finds its unit with one
// wave ballot per 64 units.  Activity follows the unit's own isovalue and rule.
__device__ __forceinline__ uint32_t batch_tile_unit(const uint32_t* __restrict__ tile0, uint32_t nunits,
                                                    uint32_t tile, uint32_t* s_unit)
{
    if (threadIdx.x < 64) {
        uint32_t n = 0;
        for (uint32_t c = 0; c < nunits; c += 64) {
            const uint32_t k = c + threadIdx.x;
            n += (uint32_t)__popcll(__ballot(k < nunits && tile0[k] <= tile));
        }
        if (threadIdx.x == 0) *s_unit = n - 1;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_unit);
}

__global__ __launch_bounds__(256) void k_filter_batch_count(const BatchUnit* __restrict__ units,
                                                            const uint32_t* __restrict__ tile0, uint32_t nunits,
                                                            uint32_t* __restrict__ tile_count)
{
    __shared__ uint32_t s_cnt[kFilterItems * 4];
    __shared__ uint32_t s_unit;
    const uint32_t tile = blockIdx.x;
    const uint32_t u = batch_tile_unit(tile0, nunits, tile, &s_unit);
    const BatchUnit& U = units[u];
    uint64_t bal[kFilterItems];
    tile_ballots(as_global(reinterpret_cast<const uint2*>(U.pairs)), U.nb, (uint64_t)(tile - U.tile0) * kFilterTile,
                 U.act, bal, s_cnt);
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t v = threadIdx.x < kFilterItems * 4 ? s_cnt[threadIdx.x] : 0u;
        const uint32_t t = wave_sum<uint32_t>(v);
        if (threadIdx.x == 0) tile_count[tile] = t;
    }
}

__global__ __launch_bounds__(256) void k_filter_batch_scatter(const BatchUnit* __restrict__ units,
                                                              const uint32_t* __restrict__ tile0, uint32_t nunits,
                                                              uint32_t ntiles, uint32_t* __restrict__ ids,
                                                              uint32_t* __restrict__ item_unit,
                                                              uint32_t* __restrict__ unit_first,
                                                              uint32_t* __restrict__ count,
                                                              const uint32_t* __restrict__ tile_count)
{
    __shared__ uint32_t s_cnt[kFilterItems * 4];
    __shared__ uint64_t s_red[4];
    __shared__ uint32_t s_unit;
    const uint32_t tile = blockIdx.x;
    const uint32_t u = batch_tile_unit(tile0, nunits, tile, &s_unit);
    const uint64_t excl = tile_prefix<uint32_t>(tile_count, tile, s_red);
    const BatchUnit& U = units[u];
    const uint64_t base = (uint64_t)(tile - U.tile0) * kFilterTile;
    uint64_t bal[kFilterItems];
    tile_ballots(as_global(reinterpret_cast<const uint2*>(U.pairs)), U.nb, base, U.act, bal, s_cnt);
    __syncthreads();
    const uint32_t tot = tile_local_scan(s_cnt);
    if (threadIdx.x == 0) {
        if (tile == U.tile0) unit_first[u] = (uint32_t)excl;
        if (tile == ntiles - 1) {
            *count = (uint32_t)(excl + tot);
            unit_first[nunits] = (uint32_t)(excl + tot);
        }
    }
    __syncthreads();
    tile_scatter(bal, s_cnt, base, U.nb, excl, ids, item_unit, u, as_global(U.map));
}

uint32_t batch_filter_tiles(uint32_t nb) { return (nb + kFilterTile - 1) / kFilterTile; }

hipError_t launch_filter_batch(const BatchUnit* units, const uint32_t* tile0, uint32_t nunits,
                               uint32_t ntiles, uint32_t* ids, uint32_t* item_unit,
                               uint32_t* unit_first, uint32_t* count, TileSums ts, hipStream_t s)
{
    if (nunits == 0 || ntiles == 0) return hipErrorInvalidValue;
    uint32_t* tc = reinterpret_cast<uint32_t*>(ts.sums);
    hipLaunchKernelGGL(k_filter_batch_count, dim3(ntiles), dim3(256), 0, s, units, tile0, nunits, tc);
    hipLaunchKernelGGL(k_filter_batch_scatter, dim3(ntiles), dim3(256), 0, s, units, tile0, nunits, ntiles, ids,
                       item_unit, unit_first, count, tc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// extraction
//
// Vertices are numbered block by block (ascending active block ID), within a
// block by owner corner in x-fastest order, then axis x, y, z; each active
// edge gets exactly one vertex, owned by the corner at its lower end.  Per
// active block the count pass stores one entry per corner:
// (vertex prefix within the block << 3) | (3 edge flags), which lets the emit
// pass resolve edges owned by the +x/+y/+z neighbour blocks without
// recounting them; the neighbours' active indices come from the dense
// block -> active-index map the compaction writes.
//
// One wavefront owns one active block at a time (no workgroup barriers): it
// stages the block's voxels in its own LDS tile, classifies its B^3 corners
// (B^3/64 consecutive corners per lane, a full x-row for B = 8), scans the
// per-lane counts with wave shuffles and writes its outputs.  Waves of a
// workgroup work on different blocks, so a wave waiting on memory never stalls
// the others.
// ---------------------------------------------------------------------------

template <int B> struct EB {
    static constexpr int NC = B * B * B;
    static constexpr int CPL = (NC + 63) / 64;  // corners per lane
    using Entry = typename std::conditional<(B <= 13), uint16_t, uint32_t>::type;
};

bool extract_supported_block(uint32_t B) { return B >= 1 && B <= 16; }
size_t corner_entry_bytes(uint32_t B) { return B <= 13 ? 2 : 4; }


// Orders LDS accesses between the lanes of one wave: the LDS unit executes a
// wave's instructions in program order, so only compiler reordering must be
// prevented.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Items of this wave: each XCD owns a contiguous eighth of the active list, so
// x-neighbour blocks (consecutive IDs) are processed close together on one L2.
struct WaveItems {
    uint32_t first, stride, end;
};

__device__ __forceinline__ WaveItems wave_items(uint32_t na, uint32_t wpg)
{
    // wave-uniform: item indices, block IDs and tile bases stay in scalar registers
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    WaveItems r;
    if (gridDim.x % 8u) {
        r.first = blockIdx.x * wpg + w;
        r.stride = gridDim.x * wpg;
        r.end = na;
        return r;
    }
    const uint32_t chunk = (na + 7u) / 8u;
    const uint32_t xcd = blockIdx.x % 8u;
    r.first = xcd * chunk + (blockIdx.x / 8u) * wpg + w;
    r.stride = (gridDim.x / 8u) * wpg;
    r.end = min(xcd * chunk + chunk, na);
    return r;
}

// Cost-balanced assignment.  Item i of an emit pass costs kItemCost plus its
// vertex (or triangle) count, so the cost of the items before i is
// kItemCost * i + base[i], base being the scan's exclusive prefix.
constexpr uint64_t kItemCost = 24;

// The first i in [0, na] with kItemCost * i + base[i] >= target (na: none);
// base[na] is never read.  Wave-uniform; each round probes 64 positions of the
// remaining range, so a few rounds cover any list.
__device__ __forceinline__ uint32_t cost_bound(const uint64_t* __restrict__ base, uint32_t na, uint64_t target)
{
    const uint32_t lane = lane_id();
    uint32_t lo = 0, hi = na;
    while (lo < hi) {
        const uint32_t p = lo + (uint32_t)(((uint64_t)(hi - lo) * (lane + 1u)) / 65u);  // in [lo, hi)
        const uint64_t bal = __ballot(kItemCost * p + base[p] >= target);
        if (!bal) {
            lo = __shfl(p, 63, 64) + 1u;
        } else {
            const int f = __builtin_ctzll(bal);
            hi = __shfl(p, f, 64);
            if (f > 0) lo = __shfl(p, f - 1, 64) + 1u;
        }
    }
    return lo;
}

// Active blocks to process: the device count, clamped to the capacity of the
// per-block arrays (a count beyond it means the list does not belong to this
// extraction; the scan flags it).
__device__ __forceinline__ uint32_t active_items(const ExtractArgs& a)
{
    return min(*a.count, a.max_blocks);
}

// This wave's items of an emit pass: strided (wave_items), or with
// a.balanced the contiguous items whose cost prefix lies in [C g / W,
// C (g + 1) / W) for wave g of W (C the pass's whole cost, `total` its count
// total), g numbered XCD by XCD so that neighbouring ranges share an L2.
__device__ __forceinline__ WaveItems cost_items(const ExtractArgs& a, uint32_t na, uint32_t wpg,
                                                const uint64_t* __restrict__ base, uint64_t total)
{
    if (!a.balanced) return wave_items(na, wpg);
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t W = (uint64_t)gridDim.x * wpg;
    const uint64_t g = gridDim.x % 8u ? (uint64_t)blockIdx.x * wpg + w
                                      : ((uint64_t)(blockIdx.x % 8u) * (gridDim.x / 8u) + blockIdx.x / 8u) * wpg + w;
    const uint64_t C = kItemCost * (uint64_t)na + total;
    WaveItems r;
    r.first = cost_bound(base, na, C * g / W);
    r.end = cost_bound(base, na, C * (g + 1) / W);
    r.stride = 1;
    return r;
}

template <typename V>
__device__ __forceinline__ V wave_excl_scan(V v, V& total)
{
    V inc = wave_incl_scan<V>(v);
    total = __shfl(inc, 63, 64);
    return inc - v;
}

// The volume an active block belongs to: the launch's own, or in the batched
// form the block's unit (wave-uniform: one unit per active block).
struct ItemVol {
    const void* data;
    uint32_t X, Y, Z, gx, gy, gz;
    float iso;
    const uint32_t* map;
};

// The volume of a batch unit (`unit` loaded ahead by the caller) or the launch's own.
template <bool BATCH>
__device__ __forceinline__ ItemVol item_vol_of(const ExtractArgs& a, uint32_t unit)
{
    ItemVol r;
    if constexpr (BATCH) {
        const BatchUnit& U = a.units[unit];
        r.data = as_global(U.data);
        r.X = U.X; r.Y = U.Y; r.Z = U.Z;
        r.gx = U.gx; r.gy = U.gy; r.gz = U.gz;
        r.iso = U.act.iso;
        r.map = as_global(U.map);
    } else {
        r.data = a.vol.data;
        r.X = a.vol.X; r.Y = a.vol.Y; r.Z = a.vol.Z;
        r.gx = a.grid.gx; r.gy = a.grid.gy; r.gz = a.grid.gz;
        r.iso = a.iso;
        r.map = a.blk_map;
    }
    return r;
}

template <bool BATCH>
__device__ __forceinline__ ItemVol item_vol(const ExtractArgs& a, uint32_t i, uint32_t& unit)
{
    ItemVol r;
    if constexpr (BATCH) {
        unit = __builtin_amdgcn_readfirstlane(a.item_unit[i]);
        const BatchUnit& U = a.units[unit];
        r.data = as_global(U.data);
        r.X = U.X; r.Y = U.Y; r.Z = U.Z;
        r.gx = U.gx; r.gy = U.gy; r.gz = U.gz;
        r.iso = U.act.iso;
        r.map = as_global(U.map);
    } else {
        unit = 0;
        r.data = a.vol.data;
        r.X = a.vol.X; r.Y = a.vol.Y; r.Z = a.vol.Z;
        r.gx = a.grid.gx; r.gy = a.grid.gy; r.gz = a.grid.gz;
        r.iso = a.iso;
        r.map = a.blk_map;
    }
    return r;
}

// ---------------------------------------------------------------------------
// meshlets (renderer layout)
//
// Per active block, the occupied cells (valid, at least one triangle) in cell
// order are cut greedily into meshlets of at most 64 distinct vertices and 126
// triangles, as src/shaders/marching_cubes_pmb.task.glsl:182-215 cuts them
// (a cell that would take the meshlet past either limit starts the next one);
// each meshlet carries its own copy of every vertex its triangles use, in
// order of first use: cell order, then owner index u (mtx_edge_uidx) within a
// cell.
//
// A cell c of a meshlet that starts at cell S adds the active edges that no
// cell of [S, c) uses.  The other cells using an edge of c are its neighbours
// in three fixed directions; the nearest of them with a smaller index lies
// back(u, F) cells before c (F: which neighbours of c exist in the block, bits
// 2a / 2a+1 for -1 / +1 along axis a).  So c adds exactly the edges with
// back(u, F) > c - S.  back takes one of nine values (1, B-1, B, B+1, B^2-B,
// B^2-1, B^2, B^2+1, B^2+B) or none; the mask of added edges therefore
// depends on F and on the interval of c - S between those values only, and is
// read from a table of 64 x 10 entries.
// ---------------------------------------------------------------------------

// MAX_VERTS_PER_MESHLET, MAX_PRIMS_PER_MESHLET: src/shaders/marching_cubes_pmb.task.glsl:24-25
constexpr uint32_t kMeshletVerts = 64;
constexpr uint32_t kMeshletTris = 126;
constexpr uint32_t kBackNone = 0xFFFFu;
constexpr uint32_t kBackClasses = 10;

template <int B>
__host__ __device__ constexpr uint32_t back_dist(int i)
{
    return i == 0 ? 1u : i == 1 ? B - 1u : i == 2 ? (uint32_t)B : i == 3 ? B + 1u : i == 4 ? (uint32_t)(B * B - B)
         : i == 5 ? (uint32_t)(B * B - 1) : i == 6 ? (uint32_t)(B * B) : i == 7 ? (uint32_t)(B * B + 1)
         : (uint32_t)(B * B + B);
}

// Interval of k among the possible back distances: how many are <= k.
template <int B>
__host__ __device__ constexpr uint32_t back_class(uint32_t k)
{
    uint32_t n = 0;
    for (int i = 0; i < 9; ++i) n += back_dist<B>(i) <= k ? 1u : 0u;
    return n;
}

// A k of class cls, or kBackNone if no k has it.
template <int B>
__host__ __device__ constexpr uint32_t back_class_rep(uint32_t cls)
{
    for (uint32_t k = 0; k <= (uint32_t)(B * B + B); ++k)
        if (back_class<B>(k) == cls) return k;
    return kBackNone;
}

// Neighbour bits of cell (lx, ly, lz) of a block whose valid cells along each
// axis are [0, vx) x [0, vy) x [0, vz): bit 2a, the cell at -1 along axis a
// is in the block; bit 2a+1, the cell at +1 is in the block and valid.
__device__ __forceinline__ uint32_t cell_nbrs(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t vx, uint32_t vy,
                                              uint32_t vz)
{
    return (lx > 0 ? 1u : 0u) | (lx + 1 < vx ? 2u : 0u) | (ly > 0 ? 4u : 0u) | (ly + 1 < vy ? 8u : 0u) |
           (lz > 0 ? 16u : 0u) | (lz + 1 < vz ? 32u : 0u);
}

// The three other cells using edge u of a cell: offsets (xp, xq) along the
// edge's two perpendicular axes p < q, and the edge's owner index in them.
// Visits each with its cell-index delta (negative: an earlier cell).
template <int B, typename F>
__device__ __forceinline__ void for_sharing_cells(uint32_t u, uint32_t nb, F&& visit)
{
    const uint32_t ax = u >> 2, dp = u & 1u, dq = (u >> 1) & 1u;
    const uint32_t p = ax == 0 ? 1u : 0u, q = ax == 2 ? 1u : 2u;
    const int sp = p == 0 ? 1 : B, sq = q == 1 ? B : B * B;
#pragma unroll
    for (uint32_t i = 0; i < 2; ++i)
#pragma unroll
        for (uint32_t j = 0; j < 2; ++j) {
            if (i == dp && j == dq) continue;
            const int xp = (int)dp - (int)i, xq = (int)dq - (int)j;
            const bool okp = xp == 0 || ((nb >> (2 * p + (xp > 0 ? 1u : 0u))) & 1u);
            const bool okq = xq == 0 || ((nb >> (2 * q + (xq > 0 ? 1u : 0u))) & 1u);
            if (okp && okq) visit(xp * sp + xq * sq, 4u * ax + i + 2u * j);
        }
}

// back(u, nb): cells back to the nearest earlier cell using the edge, or kBackNone.
template <int B>
__device__ __forceinline__ uint32_t share_back(uint32_t u, uint32_t nb)
{
    uint32_t best = kBackNone;
    for_sharing_cells<B>(u, nb, [&](int delta, uint32_t) {
        if (delta < 0) best = min(best, (uint32_t)(-delta));
    });
    return best;
}

// The cell of a meshlet that first uses edge u of a cell c (the meshlet holds
// cells c - k .. c): back = c minus that cell, u1 = the edge's owner index there.
template <int B>
__device__ __forceinline__ void first_user(uint32_t u, uint32_t nb, uint32_t k, uint32_t& back, uint32_t& u1)
{
    back = 0;
    u1 = u;
    for_sharing_cells<B>(u, nb, [&](int delta, uint32_t uo) {
        if (delta < 0 && (uint32_t)(-delta) <= k && (uint32_t)(-delta) > back) {
            back = (uint32_t)(-delta);
            u1 = uo;
        }
    });
}

// Workgroup tables: active-edge masks of the 256 cases in owner-index bits,
// the masks of edges a cell adds (64 neighbour patterns x 10 classes) and the
// class of every k below B^2 + B.
template <int B>
__device__ __forceinline__ void meshlet_tables(uint16_t* s_emu, uint16_t* s_newtab, uint8_t* s_cls)
{
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_emu[c] = (uint16_t)mtx_case_umask(c);
    for (uint32_t k = threadIdx.x; k < (uint32_t)(B * B + B); k += blockDim.x) s_cls[k] = (uint8_t)back_class<B>(k);
    for (uint32_t nb = threadIdx.x; nb < 64u; nb += blockDim.x) {
        uint32_t back[12];
#pragma unroll
        for (uint32_t u = 0; u < 12u; ++u) back[u] = share_back<B>(u, nb);
#pragma unroll
        for (uint32_t cls = 0; cls < kBackClasses; ++cls) {
            const uint32_t k = back_class_rep<B>(cls);   // folded: cls is a constant of the unrolled loop
            uint32_t m = 0;
#pragma unroll
            for (uint32_t u = 0; u < 12u; ++u) m |= back[u] > k ? 1u << u : 0u;
            s_newtab[nb * kBackClasses + cls] = (uint16_t)m;
        }
    }
}

// First-user table: for each neighbour pattern nb and edge u of a cell, the
// earlier cells of the block sharing the edge, as three 13-bit fields
// (back distance, 9 bits | the edge's owner index there << 9), farthest
// first; unused fields hold distance 0.  In a meshlet that started k cells
// back the edge's first user is the first field with 0 < distance <= k (none:
// the cell itself).
template <int B>
__device__ __forceinline__ void first_user_table(uint64_t* s_fu)
{
    for (uint32_t x = threadIdx.x; x < 64u * 12u; x += blockDim.x) {
        const uint32_t nb = x / 12u, u = x % 12u;
        uint32_t d[3] = { 0, 0, 0 }, uo[3] = { 0, 0, 0 };
        uint32_t n = 0;
        for_sharing_cells<B>(u, nb, [&](int delta, uint32_t o) {
            if (delta < 0) {
                d[n] = (uint32_t)(-delta);
                uo[n] = o;
                ++n;
            }
        });
        // farthest first
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2 - i; ++j)
                if (d[j] < d[j + 1]) {
                    const uint32_t td = d[j], to = uo[j];
                    d[j] = d[j + 1]; uo[j] = uo[j + 1];
                    d[j + 1] = td; uo[j + 1] = to;
                }
        uint64_t e = 0;
        for (int i = 0; i < 3; ++i) e |= (uint64_t)(d[i] | (uo[i] << 9)) << (13 * i);
        s_fu[x] = e;
    }
}

// The first user of edge u (back distance, owner index there) from its table entry.
__device__ __forceinline__ void first_user_of(uint64_t e, uint32_t u, uint32_t k, uint32_t& back, uint32_t& u1)
{
    back = 0;
    u1 = u;
#pragma unroll
    for (int f = 2; f >= 0; --f) {   // nearest first, so the farthest qualifying field is taken last
        const uint32_t fld = (uint32_t)(e >> (13 * f)) & 0x1FFFu;
        const uint32_t d = fld & 511u;
        if (d != 0u && d <= k) {
            back = d;
            u1 = fld >> 9;
        }
    }
}

// The first-user table by distance class: for each neighbour pattern nb,
// edge u and class cls of the distance k back to the meshlet's start (the
// classes of k between consecutive sharing distances, as s_cls gives them),
// one byte: the index i + 1 of the first user's distance back_dist(i) (0:
// the cell itself) | the edge's owner index there << 4.  s_bd[i + 1] holds
// back_dist(i), s_bd[0] = 0.  The same answer as first_user_of for every k of
// the class (the qualifying distances are the same), in one byte read.
template <int B>
__device__ __forceinline__ void first_user_bytes(uint8_t* s_fu8, uint8_t* s_bd, const uint8_t* s_cls)
{
    static_assert(B >= 3, "the nine sharing distances are distinct");
    constexpr uint32_t KF = (uint32_t)(B * B + B);
    if (threadIdx.x < 16u)
        s_bd[threadIdx.x] = (uint8_t)(threadIdx.x >= 1u && threadIdx.x <= 9u ? back_dist<B>((int)threadIdx.x - 1) : 0u);
    __syncthreads();   // s_cls, written by meshlet_tables before this call
    for (uint32_t x = threadIdx.x; x < 64u * 12u; x += blockDim.x) {
        const uint32_t nb = x / 12u, u = x % 12u;
        // each earlier sharing cell: the class of its distance (the distance's
        // index + 1, as the distances are distinct and sorted) and the edge's
        // owner index there
        uint32_t ci[3] = { 0, 0, 0 }, uo[3] = { 0, 0, 0 };
        uint32_t n = 0;
        for_sharing_cells<B>(u, nb, [&](int delta, uint32_t o) {
            if (delta < 0) {
                const uint32_t d = (uint32_t)(-delta);
                ci[n] = d >= KF ? kBackClasses - 1u : (uint32_t)s_cls[d];
                uo[n] = o;
                ++n;
            }
        });
        // a sharing cell is in a meshlet that started k back iff its
        // distance's class is at most k's; the first user is the farthest
        for (uint32_t cls = 0; cls < kBackClasses; ++cls) {
            uint32_t idx = 0, u1 = u;
            for (int f = 0; f < 3; ++f)
                if (ci[f] != 0u && ci[f] <= cls && ci[f] > idx) {
                    idx = ci[f];
                    u1 = uo[f];
                }
            s_fu8[x * kBackClasses + cls] = (uint8_t)(idx | (u1 << 4));
        }
    }
}

// Vertices a cell adds to a meshlet that started k cells before it.
template <int B>
__device__ __forceinline__ uint32_t added_edges(uint32_t emu, uint32_t nb, uint32_t k, const uint16_t* s_newtab,
                                                const uint8_t* s_cls)
{
    const uint32_t cls = k >= (uint32_t)(B * B + B) ? 9u : (uint32_t)s_cls[k];
    return emu & s_newtab[nb * kBackClasses + cls];
}

// Greedy cut of one block into meshlets.  The wave's lanes hold cells
// lane * CPL + j, j < CPL; cell(j, cs, nb) gives cell j's case (0 if invalid)
// and neighbour bits.  Returns the block's meshlet count and meshlet-vertex
// count (wave-uniform) and this lane's start bits (bit j: a meshlet starts at
// cell lane * CPL + j; cell 0 starts the first one, the cells before the
// first occupied one holding nothing).  Each step cuts one meshlet: the
// vertices and triangles every cell from the meshlet's start on would add,
// a wave scan of them, and the first cell whose running sums pass a limit.
// For up to 8 cells per lane each cell's case-derived data is read once and
// kept packed: edge bits (12) | neighbour bits << 12 | triangles << 18 |
// edges no earlier cell of the block uses << 21.
template <int B, int CPL, typename Cell>
__device__ __forceinline__ void meshlet_cut(Cell&& cell, const uint8_t* s_ntri, const uint16_t* s_emu,
                                            const uint16_t* s_newtab, const uint8_t* s_cls, uint32_t& nm, uint32_t& ns, uint64_t& sbits)
{
    constexpr uint32_t NC = B * B * B;
    constexpr bool KEEP = CPL <= 8;
    constexpr uint32_t KFAR = (uint32_t)(B * B + B);   // at or past it every earlier sharer is in the meshlet
    const uint32_t lane = lane_id();
    nm = 0;
    ns = 0;
    sbits = 0;
    uint32_t info[KEEP ? CPL : 1];
    uint32_t any = 0;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
        uint32_t cs, nb;
        cell(j, cs, nb);
        const uint32_t p = lane * CPL + j < NC ? (uint32_t)s_ntri[cs] : 0u;
        any |= p;
        if constexpr (KEEP) {
            const uint32_t emu = p ? (uint32_t)s_emu[cs] : 0u;
            const uint32_t ninf = __popc(emu & s_newtab[nb * kBackClasses + kBackClasses - 1]);
            info[j] = emu | (nb << 12) | (p << 18) | (ninf << 21);
        }
    }
    if (__ballot(any != 0) == 0) return;
    if (lane == 0) sbits = 1;
    uint32_t S = 0;
    // vertices (v) and triangles (p) cell j adds to the meshlet starting at S
    auto vp = [&](int j, uint32_t& v, uint32_t& p) {
        const uint32_t c = lane * CPL + j;
        v = 0;
        p = 0;
        if (c >= NC || c < S) return;
        const uint32_t k = c - S;
        if constexpr (KEEP) {
            const uint32_t e = info[j];
            p = (e >> 18) & 7u;
            if (k >= KFAR) v = e >> 21;
            else if (p) v = __popc(added_edges<B>(e & 0xFFFu, (e >> 12) & 63u, k, s_newtab, s_cls));
        } else {
            uint32_t cs, nb;
            cell(j, cs, nb);
            p = s_ntri[cs];
            if (p) v = __popc(added_edges<B>(s_emu[cs], nb, k, s_newtab, s_cls));
        }
    };
    for (;;) {
        uint32_t lv = 0, lp = 0, vpk = 0, ppk = 0;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            uint32_t v, p;
            vp(j, v, p);
            lv += v;
            lp += p;
            if constexpr (KEEP) {
                vpk |= v << (4 * j);
                ppk |= p << (3 * j);
            }
        }
        uint32_t tv, tp;
        uint32_t rv = wave_excl_scan<uint32_t>(lv, tv);
        uint32_t rp = wave_excl_scan<uint32_t>(lp, tp);
        uint32_t cut = 0xFFFFFFFFu, before = 0;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            uint32_t v, p;
            if constexpr (KEEP) {
                v = (vpk >> (4 * j)) & 15u;
                p = (ppk >> (3 * j)) & 7u;
            } else {
                vp(j, v, p);
            }
            rv += v;
            rp += p;
            if (cut == 0xFFFFFFFFu && (rv > kMeshletVerts || rp > kMeshletTris)) {
                cut = lane * CPL + j;
                before = rv - v;
            }
        }
        const uint64_t bal = __ballot(cut != 0xFFFFFFFFu);
        ++nm;
        if (!bal) {
            ns += tv;
            break;
        }
        const int src = __builtin_ctzll(bal);
        const uint32_t b = __shfl(cut, src, 64);
        ns += __shfl(before, src, 64);
        if (lane == b / CPL) sbits |= 1ull << (b % CPL);
        S = b;
    }
}

// The same greedy cut over a list of the block's occupied cells (cells with
// triangles; an empty cell adds nothing to a meshlet, so no cut falls on
// one): the lanes write their cells' entries in cell order to s_ent (NC words
// of LDS, this wave's), then each step scans 64 entries per pass instead of
// every cell of the block.  Entry: edge bits (12) | neighbour bits << 12 |
// triangles << 18 | cell << 21.
template <int B, int CPL, typename Cell>
__device__ __forceinline__ void meshlet_cut_list(Cell&& cell, const uint8_t* s_ntri, const uint16_t* s_emu,
                                                 const uint16_t* s_newtab, const uint8_t* s_cls, uint32_t* s_ent,
                                                 uint32_t& nm, uint32_t& ns, uint64_t& sbits)
{
    constexpr uint32_t NC = B * B * B;
    static_assert(NC <= 2048 && CPL <= 32, "the cell index shares the entry word");
    constexpr uint32_t KFAR = (uint32_t)(B * B + B);
    const uint32_t lane = lane_id();
    nm = 0;
    ns = 0;
    sbits = 0;
    uint32_t info[CPL];
    uint32_t occ = 0;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
        uint32_t cs, nb;
        cell(j, cs, nb);
        const uint32_t p = lane * CPL + j < NC ? (uint32_t)s_ntri[cs] : 0u;
        info[j] = p ? (uint32_t)s_emu[cs] | (nb << 12) | (p << 18) | ((lane * CPL + j) << 21) : 0u;
        occ |= (p ? 1u : 0u) << j;
    }
    uint32_t T;
    uint32_t q = wave_excl_scan<uint32_t>((uint32_t)__popc(occ), T);
    if (T == 0) return;
#pragma unroll
    for (int j = 0; j < CPL; ++j)
        if ((occ >> j) & 1u) s_ent[q++] = info[j];
    wave_lds_sync();
    if (lane == 0) sbits = 1;
    uint32_t S = 0, e0 = 0;   // the meshlet's first cell and its first entry
    for (;;) {
        uint32_t cv = 0, cp = 0;   // vertices and triangles of the meshlet's earlier passes
        bool cut = false;
        for (uint32_t q0 = e0 & ~63u; q0 < T; q0 += 64u) {
            const uint32_t e = q0 + lane;
            uint32_t v = 0, p = 0, c = 0;
            if (e >= e0 && e < T) {
                const uint32_t x = s_ent[e];
                c = x >> 21;
                p = (x >> 18) & 7u;
                const uint32_t k = c - S;
                const uint32_t cls = k >= KFAR ? kBackClasses - 1u : (uint32_t)s_cls[k];
                v = __popc(x & 0xFFFu & s_newtab[((x >> 12) & 63u) * kBackClasses + cls]);
            }
            const uint32_t rv = cv + wave_incl_scan<uint32_t>(v), rp = cp + wave_incl_scan<uint32_t>(p);
            const uint64_t bal = __ballot(rv > kMeshletVerts || rp > kMeshletTris);
            if (bal) {
                // the first entry past a limit starts the next meshlet
                const int src = __builtin_ctzll(bal);
                const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)c, src);
                ns += (uint32_t)__builtin_amdgcn_readlane((int)(rv - v), src);
                if (lane == b / CPL) sbits |= 1ull << (b % CPL);
                S = b;
                e0 = q0 + (uint32_t)src;
                cut = true;
                break;
            }
            cv = wave_last(rv);
            cp = wave_last(rp);
        }
        ++nm;
        if (!cut) {
            ns += cv;
            break;
        }
    }
    wave_lds_sync();
}

// Words of a block's meshlet start bits (one bit per cell).
template <int B> struct StartBits {
    static constexpr uint32_t NW = (B * B * B + 31) / 32;
};

// This lane's start bits (cells lane * CPL .. lane * CPL + CPL - 1, CPL <= 64)
// from a block's start words.
template <int B, int CPL>
__device__ __forceinline__ uint64_t lane_start_bits(const uint32_t* __restrict__ w, uint32_t lane)
{
    constexpr uint32_t NW = StartBits<B>::NW, NC = B * B * B;
    const uint32_t first = lane * CPL;
    if (first >= NC) return 0;
    const uint32_t w0 = first >> 5, sh = first & 31u;
    const uint64_t lo = (uint64_t)w[w0] | (w0 + 1 < NW ? (uint64_t)w[w0 + 1] << 32 : 0ull);
    uint64_t v = lo >> sh;
    if (sh && w0 + 2 < NW) v |= (uint64_t)w[w0 + 2] << (64 - sh);
    return CPL >= 64 ? v : (v & ((1ull << CPL) - 1ull));
}

template <int B, typename T, bool BATCH>
__global__ __launch_bounds__(256) void k_extract_count(ExtractArgs a)
{
    constexpr int E = B + 1;
    constexpr int NTILE = E * E * E;
    constexpr int NC = EB<B>::NC;
    constexpr int CPL = EB<B>::CPL;
    using Entry = typename EB<B>::Entry;
    __shared__ uint8_t s_all[4][NTILE];
    __shared__ uint8_t s_ntri[256];
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_ntri[c] = (uint8_t)mtx_word_ntri(mtx_tri_words[c]);
    __syncthreads();
    const uint32_t lane = lane_id();
    uint8_t* s_below = s_all[threadIdx.x >> 6];
    const uint32_t na = active_items(a);
    Entry* __restrict__ table = static_cast<Entry*>(a.corner_table);
    const WaveItems it = wave_items(na, 4);
    for (uint32_t i = it.first; i < it.end; i += it.stride) {
        uint32_t unit;
        const ItemVol V = item_vol<BATCH>(a, i, unit);
        const T* __restrict__ vol = static_cast<const T*>(V.data);
        const uint32_t X = V.X, Y = V.Y, Z = V.Z;
        // a volume thinner than 2 voxels on any axis has no cells, hence no
        // triangles and no vertices (every vertex is a corner of some triangle)
        const bool degen = X < 2 || Y < 2 || Z < 2;
        const uint32_t id = a.ids[i];
        const uint32_t bx = id % V.gx, by = (id / V.gx) % V.gy;
        const uint32_t bz = id / (V.gx * V.gy);
        const uint32_t ox = bx * B, oy = by * B, oz = bz * B;
        // all of a batch's loads are issued before the first is consumed
        constexpr int LPL = (NTILE + 63) / 64, NLD = LPL < 16 ? LPL : 16;
#pragma unroll
        for (int k0 = 0; k0 < LPL; k0 += NLD) {
            T v[NLD];
            bool in[NLD];
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const uint32_t j = lane + 64u * (k0 + k);
                const uint32_t x = ox + j % E, y = oy + (j / E) % E, z = oz + j / (E * E);
                in[k] = j < (uint32_t)NTILE && x < X && y < Y && z < Z;
                v[k] = vol[in[k] ? x + (uint64_t)X * (y + (uint64_t)Y * z) : 0];
            }
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const uint32_t j = lane + 64u * (k0 + k);
                if (j < (uint32_t)NTILE) s_below[j] = (uint8_t)(in[k] && vfloat(v[k]) <= V.iso);
            }
        }
        wave_lds_sync();
        uint32_t nv = 0, nt = 0;
        uint32_t fl[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const uint32_t c = lane * CPL + j;
            fl[j] = 0;
            uint32_t cs = 0;
            if (c < (uint32_t)NC) {
                const uint32_t lx = c % B, ly = (c / B) % B, lz = c / (B * B);
                const uint32_t x = ox + lx, y = oy + ly, z = oz + lz;
                const uint32_t t0 = lx + E * (ly + E * lz);
                if (!degen && x < X && y < Y && z < Z) {
                    const uint32_t b0 = s_below[t0];
                    uint32_t f = 0;
                    if (x + 1 < X) f |= b0 ^ s_below[t0 + 1];
                    if (y + 1 < Y) f |= (b0 ^ s_below[t0 + E]) << 1;
                    if (z + 1 < Z) f |= (b0 ^ s_below[t0 + E * E]) << 2;
                    fl[j] = f;
                    nv += __popc(f);
                    if (x + 1 < X && y + 1 < Y && z + 1 < Z) {
#pragma unroll
                        for (uint32_t q = 0; q < 8; ++q)
                            cs |= (uint32_t)s_below[t0 + mtx_corner_x(q) + E * (mtx_corner_y(q) + E * mtx_corner_z(q))] << q;
                        nt += s_ntri[cs];
                    }
                }
                a.cases[(uint64_t)i * NC + c] = (uint8_t)cs;
            }
        }
        uint32_t tv, tt;
        uint32_t run = wave_excl_scan<uint32_t>(nv, tv);
        nt = wave_sum<uint32_t>(nt);
        tt = nt;
        Entry ent[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            ent[j] = (Entry)((run << 3) | fl[j]);
            run += __popc(fl[j]);
        }
        Entry* dst = table + (uint64_t)i * NC + lane * CPL;
        if constexpr (CPL * sizeof(Entry) == 16) {
            if (lane * CPL < (uint32_t)NC) *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(ent);
        } else {
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                if (lane * CPL + j < (uint32_t)NC) dst[j] = ent[j];
        }
        // meshlet counts: 0 here, the cut kernel writes them when meshlets are made
        if (lane == 0) *reinterpret_cast<uint4*>(a.blk_counts + 4 * (uint64_t)i) = make_uint4(tv, tt, 0u, 0u);
        wave_lds_sync();
    }
}

// Count pass specialised for B = 8: the below-isovalue bits of the 9^3 tile are
// gathered with wave ballots (12 x 64 bits, in LDS), each lane owns one x-row
// of 8 corners (ly = lane & 7, lz = lane >> 3) and classifies it with bit
// operations on its four 9-bit rows (y, y+1) x (z, z+1).  A lane's 12 tile
// coordinates are packed once; the loads take 32-bit offsets from the tile
// origin.
//
// CUT: the block's meshlets are cut from the cases right after they are
// classified (the cut needs nothing else), instead of in a launch of its own.
template <typename T, bool BATCH, bool CUT>
__global__ __launch_bounds__(256) void k_extract_count_b8(ExtractArgs a)
{
    constexpr int B = 8, E = 9, NTILE = 729, LPL = 12, NC = 512;
    __shared__ uint64_t s_bits_all[4][LPL];
    __shared__ uint8_t s_ntri[256];
    // meshlet cut tables (CUT only)
    __shared__ uint16_t s_emu[CUT ? 256 : 1];
    __shared__ uint16_t s_newtab[CUT ? 64 * kBackClasses : 1];
    __shared__ uint8_t s_cls[CUT ? B * B + B : 1];
    __shared__ uint32_t s_ent_all[CUT ? 4 : 1][CUT ? NC : 1];   // each wave's list of occupied cells
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_ntri[c] = (uint8_t)mtx_word_ntri(mtx_tri_words[c]);
    if constexpr (CUT) meshlet_tables<B>(s_emu, s_newtab, s_cls);
    __syncthreads();
    const uint32_t lane = lane_id();
    uint64_t* s_bits = s_bits_all[threadIdx.x >> 6];
    uint32_t* s_ent = s_ent_all[CUT ? threadIdx.x >> 6 : 0];
    // tile coordinates of the voxels this lane loads, j = lane + 64 k:
    // (j % 9) | (j / 9 % 9) << 4 | (j / 81) << 8, and bit 12 past the tile
    uint32_t tcoord[LPL];
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
        const uint32_t j = lane + 64u * k;
        tcoord[k] = j < (uint32_t)NTILE ? (j % E) | (((j / E) % E) << 4) | ((j / (E * E)) << 8) : 0x1000u;
    }
    const uint32_t na = active_items(a);
    uint16_t* __restrict__ table = static_cast<uint16_t*>(a.corner_table);
    const uint32_t ly = lane & 7u, lz = lane >> 3;
    const WaveItems it = wave_items(na, 4);
    // the next block's ID (and batch unit) load while this block is processed:
    // its tile loads then wait on one memory round trip, not two or three
    uint32_t id_next = it.first < it.end ? a.ids[it.first] : 0u;
    uint32_t unit_next = BATCH && it.first < it.end ? a.item_unit[it.first] : 0u;
    for (uint32_t i = it.first; i < it.end; i += it.stride) {
        const uint32_t id = __builtin_amdgcn_readfirstlane(id_next);
        const uint32_t unit = __builtin_amdgcn_readfirstlane(unit_next);
        if (i + it.stride < it.end) {
            id_next = a.ids[i + it.stride];
            if constexpr (BATCH) unit_next = a.item_unit[i + it.stride];
        }
        const ItemVol V = item_vol_of<BATCH>(a, unit);
        const T* __restrict__ vol = static_cast<const T*>(V.data);
        const uint32_t X = V.X, Y = V.Y, Z = V.Z;
        const bool degen = X < 2 || Y < 2 || Z < 2;
        const bool off32 = (uint64_t)X * Y * E < 0x80000000ull;
        const uint32_t bx = id % V.gx, by = (id / V.gx) % V.gy;
        const uint32_t bz = id / (V.gx * V.gy);
        const uint32_t ox = bx * B, oy = by * B, oz = bz * B;
        const uint64_t base = ox + (uint64_t)X * (oy + (uint64_t)Y * oz);
        const T* __restrict__ tile = vol + base;
        const bool interior = off32 && ox + B < X && oy + B < Y && oz + B < Z;
        // the coordinates are re-read per block (opaque to the compiler), so
        // the 36 values derived from them are not kept live across blocks
        uint32_t tc[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
            tc[k] = tcoord[k];
            asm volatile("" : "+v"(tc[k]));
        }
        T v[LPL];
        bool in[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
            const uint32_t tx = tc[k] & 15u, ty = (tc[k] >> 4) & 15u, tz = (tc[k] >> 8) & 15u;
            in[k] = !(tc[k] >> 12) && (interior || (ox + tx < X && oy + ty < Y && oz + tz < Z));
            if (off32) {
                // tile-relative offsets fit 32 bits; voxels outside read the tile origin
                v[k] = tile[in[k] ? tx + X * (ty + Y * tz) : 0u];
            } else {
                v[k] = tile[in[k] ? tx + (uint64_t)X * (ty + (uint64_t)Y * tz) : 0ull];
            }
        }
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
            const uint64_t bal = __ballot(in[k] && vfloat(v[k]) <= V.iso);
            if (lane == 0) s_bits[k] = bal;
        }
        wave_lds_sync();
        // the four 9-bit rows of this lane: (y, z), (y+1, z), (y, z+1), (y+1, z+1)
        auto row = [&](uint32_t r) -> uint32_t {
            const uint32_t o = 9u * r;
            return ((uint32_t)(s_bits[o >> 6] >> (o & 63u)) & 0x1FFu) |
                   (((o & 63u) > 55u) ? (uint32_t)(s_bits[(o >> 6) + 1] << (64u - (o & 63u))) & 0x1FFu : 0u);
        };
        const uint32_t r0 = ly + 9u * lz;
        const uint32_t b00 = row(r0), b10 = row(r0 + 1u), b01 = row(r0 + 9u), b11 = row(r0 + 10u);
        const uint32_t nx = X - ox;  // >= 1
        const uint32_t ex_x = nx >= 8u ? 0xFFu : ((1u << nx) - 1u);
        const uint32_t ed_x = nx >= 9u ? 0xFFu : ((1u << (nx - 1u)) - 1u);
        const bool rowok = !degen && oy + ly < Y && oz + lz < Z;
        const bool yok = oy + ly + 1u < Y, zok = oz + lz + 1u < Z;
        const uint32_t fx = rowok ? (b00 ^ (b00 >> 1)) & ed_x : 0u;
        const uint32_t fy = (rowok && yok) ? (b00 ^ b10) & ex_x : 0u;
        const uint32_t fz = (rowok && zok) ? (b00 ^ b01) & ex_x : 0u;
        const uint32_t cellok = (rowok && yok && zok) ? ed_x : 0u;
        uint32_t nt = 0, cases0 = 0, cases1 = 0;
#pragma unroll
        for (uint32_t lx = 0; lx < 8u; ++lx) {
            const uint32_t cs = ((b00 >> lx) & 1u) | ((b00 >> lx) & 2u) | (((b10 >> lx) & 2u) << 1) |
                                (((b10 >> lx) & 1u) << 3) | (((b01 >> lx) & 1u) << 4) | (((b01 >> lx) & 2u) << 4) |
                                (((b11 >> lx) & 2u) << 5) | (((b11 >> lx) & 1u) << 7);
            const uint32_t c = ((cellok >> lx) & 1u) ? cs : 0u;
            nt += (cellok >> lx) & 1u ? s_ntri[c] : 0u;
            if (lx < 4) cases0 |= c << (8 * lx);
            else cases1 |= c << (8 * (lx - 4));
        }
        const uint32_t nv = __popc(fx) + __popc(fy) + __popc(fz);
        uint32_t tv;
        uint32_t run = wave_excl_scan<uint32_t>(nv, tv);
        const uint32_t tt = wave_sum<uint32_t>(nt);
        uint16_t ent[8];
#pragma unroll
        for (uint32_t lx = 0; lx < 8u; ++lx) {
            const uint32_t f = ((fx >> lx) & 1u) | (((fy >> lx) & 1u) << 1) | (((fz >> lx) & 1u) << 2);
            ent[lx] = (uint16_t)((run << 3) | f);
            run += __popc(f);
        }
        *reinterpret_cast<uint4*>(table + (uint64_t)i * NC + lane * 8u) = *reinterpret_cast<const uint4*>(ent);
        *reinterpret_cast<uint2*>(a.cases + (uint64_t)i * NC + lane * 8u) = make_uint2(cases0, cases1);
        uint32_t nm = 0, ns = 0;
        if constexpr (CUT) {
            // this lane's cells are lane * 8 + lx (lx = 0..7), the cells of its row
            const uint32_t vx = X - 1u > ox ? min(X - 1u - ox, (uint32_t)B) : 0u;
            const uint32_t vy = Y - 1u > oy ? min(Y - 1u - oy, (uint32_t)B) : 0u;
            const uint32_t vz = Z - 1u > oz ? min(Z - 1u - oz, (uint32_t)B) : 0u;
            uint64_t sb;
            meshlet_cut_list<B, 8>(
                [&](int j, uint32_t& cs, uint32_t& nb) {
                    cs = ((j < 4 ? cases0 : cases1) >> (8 * (j & 3))) & 0xFFu;
                    nb = cell_nbrs((uint32_t)j, ly, lz, vx, vy, vz);
                },
                s_ntri, s_emu, s_newtab, s_cls, s_ent, nm, ns, sb);
            reinterpret_cast<uint8_t*>(a.starts)[(uint64_t)i * 64 + lane] = (uint8_t)sb;
        }
        if (lane == 0) *reinterpret_cast<uint4*>(a.blk_counts + 4 * (uint64_t)i) = make_uint4(tv, tt, nm, ns);
        wave_lds_sync();
    }
}

// Exclusive scan of the per-block counts {vertices, triangles, meshlets,
// meshlet vertices} into bases, in two launches (reduce-then-scatter, as the
// compaction): tile sums, then each tile adds up its predecessors' sums and
// writes its blocks' bases.  Tile sums: 4 words per tile.
constexpr uint32_t kScanItems = 4;
constexpr uint32_t kScanQ = 4;   // quantities scanned
uint32_t scan_tile_items() { return 256 * kScanItems; }

__device__ __forceinline__ uint32_t scan_tiles_of(uint32_t na)
{
    const uint32_t n = (na + 256 * kScanItems - 1) / (256 * kScanItems);
    return n ? n : 1;
}

__device__ __forceinline__ void block_counts(const uint32_t* __restrict__ cnt, uint32_t na, uint64_t i,
                                             uint64_t (&q)[kScanQ])
{
    const uint4 c = i < na ? *reinterpret_cast<const uint4*>(cnt + 4 * i) : make_uint4(0, 0, 0, 0);
    q[0] = c.x;
    q[1] = c.y;
    q[2] = c.z;
    q[3] = c.w;
}

__global__ __launch_bounds__(256) void k_extract_tile_sums(const uint32_t* __restrict__ cnt,
                                                           const uint32_t* __restrict__ count,
                                                           uint32_t max_blocks,
                                                           uint64_t* __restrict__ tile_sums,
                                                           uint64_t* __restrict__ totals)
{
    __shared__ uint64_t s_red[4 * kScanQ];
    const uint32_t na = min(*count, max_blocks);
    const uint32_t tile = blockIdx.x;
    if (tile >= scan_tiles_of(na)) return;
    // error flags of the emit pass; bit 1: a count beyond the per-block arrays
    if (tile == 0 && threadIdx.x == 0) totals[2] = *count > max_blocks ? 2ull : 0ull;
    uint64_t sum[kScanQ] = { 0, 0, 0, 0 };
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
        uint64_t q[kScanQ];
        block_counts(cnt, na, (uint64_t)tile * 256 * kScanItems + threadIdx.x * kScanItems + k, q);
#pragma unroll
        for (uint32_t r = 0; r < kScanQ; ++r) sum[r] += q[r];
    }
    const uint32_t w = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t r = 0; r < kScanQ; ++r) {
        sum[r] = wave_sum<uint64_t>(sum[r]);
        if ((threadIdx.x & 63) == 0) s_red[4 * r + w] = sum[r];
    }
    __syncthreads();
    if (threadIdx.x < kScanQ) {
        const uint64_t* r = s_red + 4 * threadIdx.x;
        tile_sums[kScanQ * (uint64_t)tile + threadIdx.x] = r[0] + r[1] + r[2] + r[3];
    }
}

// Bases: vbase, tbase, mbase, sbase (meshlet vertices); totals: [0] vertices,
// [1] triangles, [3] meshlets, [4] meshlet vertices; out_counts: vertexCount,
// indexCount, meshletCount, meshletVertexCount (saturated to 32 bits).
__global__ __launch_bounds__(256) void k_extract_scan(const uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ count,
                                                      uint32_t max_blocks,
                                                      uint64_t* __restrict__ vbase,
                                                      uint64_t* __restrict__ tbase,
                                                      uint64_t* __restrict__ mbase,
                                                      uint64_t* __restrict__ sbase,
                                                      uint64_t* __restrict__ totals,
                                                      uint32_t* __restrict__ out_counts,
                                                      const uint64_t* __restrict__ tile_sums)
{
    __shared__ uint64_t s_w[4 * kScanQ];
    __shared__ uint64_t s_red[4 * kScanQ];
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t na = min(*count, max_blocks);
    const uint32_t ntiles = scan_tiles_of(na);
    const uint32_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    // exclusive prefix of this tile: its predecessors' sums
    uint64_t pre[kScanQ] = { 0, 0, 0, 0 };
    for (uint32_t k = threadIdx.x; k < tile; k += 256) {
#pragma unroll
        for (uint32_t r = 0; r < kScanQ; ++r) pre[r] += tile_sums[kScanQ * (uint64_t)k + r];
    }
#pragma unroll
    for (uint32_t r = 0; r < kScanQ; ++r) {
        pre[r] = wave_sum<uint64_t>(pre[r]);
        if (lane == 0) s_red[4 * r + w] = pre[r];
    }
    uint64_t q[kScanItems][kScanQ];
    uint64_t sum[kScanQ] = { 0, 0, 0, 0 };
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
        block_counts(cnt, na, (uint64_t)tile * 256 * kScanItems + threadIdx.x * kScanItems + k, q[k]);
#pragma unroll
        for (uint32_t r = 0; r < kScanQ; ++r) sum[r] += q[k][r];
    }
    uint64_t inc[kScanQ];
#pragma unroll
    for (uint32_t r = 0; r < kScanQ; ++r) {
        inc[r] = wave_incl_scan<uint64_t>(sum[r]);
        if (lane == 63) s_w[4 * r + w] = inc[r];
    }
    __syncthreads();
    uint64_t run[kScanQ], total[kScanQ];
#pragma unroll
    for (uint32_t r = 0; r < kScanQ; ++r) {
        const uint64_t ex = s_red[4 * r] + s_red[4 * r + 1] + s_red[4 * r + 2] + s_red[4 * r + 3];
        uint64_t ow = 0, all = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            ow += j < w ? s_w[4 * r + j] : 0;
            all += s_w[4 * r + j];
        }
        run[r] = ex + ow + inc[r] - sum[r];
        total[r] = ex + all;
    }
    if (threadIdx.x == 0 && tile == ntiles - 1) {
        totals[0] = total[0];
        totals[1] = total[1];
        totals[3] = total[2];
        totals[4] = total[3];
        if (out_counts) {
            const uint64_t c4[4] = { total[0], 3 * total[1], total[2], total[3] };
#pragma unroll
            for (int r = 0; r < 4; ++r) out_counts[r] = c4[r] > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c4[r];
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
        const uint64_t i = (uint64_t)tile * 256 * kScanItems + threadIdx.x * kScanItems + k;
        if (i < na) {
            vbase[i] = run[0];
            tbase[i] = run[1];
            mbase[i] = run[2];
            sbase[i] = run[3];
        }
#pragma unroll
        for (uint32_t r = 0; r < kScanQ; ++r) run[r] += q[k][r];
    }
}

template <typename LT>
__device__ __forceinline__ void grad_normal(const LT* s, uint32_t t, uint32_t sy, uint32_t sz, float n[3])
{
    // normalize(s(p-x)-s(p+x), s(p-y)-s(p+y), s(p-z)-s(p+z)); (0,1,0) where the
    // gradient vanishes (|g| < 1e-5); sy, sz: the tile's y and z strides
    const float gx = vfloat(s[t - 1]) - vfloat(s[t + 1]);
    const float gy = vfloat(s[t - sy]) - vfloat(s[t + sy]);
    const float gz = vfloat(s[t - sz]) - vfloat(s[t + sz]);
    const float l2 = gx * gx + gy * gy + gz * gz;
    if (l2 < 1e-10f) { n[0] = 0.f; n[1] = 1.f; n[2] = 0.f; return; }
    const float r = rsqrtf(l2);
    n[0] = gx * r; n[1] = gy * r; n[2] = gz * r;
}

// The tile is kept in LDS in the volume's own type (u8 voxels as bytes: a
// quarter of the LDS of f32, so LDS does not bound the kernel's occupancy).
// Its rows are padded to whole 4-voxel segments (pitch RP): interior tiles
// are loaded one segment per lane straight from global memory into LDS
// (16-byte direct loads for f32, 4-byte for u8), ND loads per lane and no
// registers, the last load's spare lanes writing past the tile.
template <int B, bool NORMALS, typename T> struct VertCfg {
    static constexpr int A = NORMALS ? 1 : 0;  // apron for central differences
    static constexpr int E = B + 1 + 2 * A;
    static constexpr int NTILE = E * E * E;
    static constexpr int RP = (E + 3) / 4 * 4;  // LDS row pitch, voxels
    static constexpr int NS = RP / 4;           // 4-voxel segments per row
    static constexpr int NSEG = E * E * NS;
    static constexpr int ND = (NSEG + 63) / 64;
    static constexpr int TILE = ND * 64 * 4;    // LDS voxels incl. the spare lanes' segments
    static constexpr unsigned SEGB = 4 * sizeof(T);   // bytes per direct load and lane
    // vertices listed at a time, and half a round of vertices (32 bytes
    // each) staged for coalesced stores: with the f32 tile that is 30 KB per
    // workgroup, so 5 workgroups fit a CU (the kernel is held to the 102
    // VGPRs of 5 waves per SIMD); 512 entries and whole rounds took 36.9 KB,
    // 4 waves per SIMD, and 23 us more at 1024^3
    static constexpr int LIST = 256;
    static constexpr int WAVE_LDS = TILE * (int)sizeof(T) + LIST * 2 + 16 + 32 * 32;
    static constexpr int MIN_WAVES = B == 8 ? 5 : 1;   // per SIMD; larger tiles stop at fewer
    static constexpr int WPG = WAVE_LDS * 4 <= 65536 ? 4 : (WAVE_LDS * 2 <= 65536 ? 2 : 1);
};

// This lane's corner entries (prefix << 3 | flags) and cell cases from the
// count pass, packed: 16-bit entries two per register, cases four per register.
template <int B>
struct LaneCorners {
    using Entry = typename EB<B>::Entry;
    static constexpr int CPL = EB<B>::CPL, NC = EB<B>::NC;
    static constexpr bool E16 = sizeof(Entry) == 2;
    static constexpr int NEW = E16 ? (CPL + 1) / 2 : CPL, NCW = (CPL + 3) / 4;
    uint32_t entw[NEW], csw[NCW];

    __device__ __forceinline__ void load(const Entry* __restrict__ table, const uint8_t* __restrict__ cases,
                                         uint64_t i, uint32_t lane, bool want_cases)
    {
        if constexpr (CPL == 8 && E16) {
            const uint4 e4 = *reinterpret_cast<const uint4*>(table + i * NC + lane * 8u);
            entw[0] = e4.x; entw[1] = e4.y; entw[2] = e4.z; entw[3] = e4.w;
            if (want_cases) {
                const uint2 c2 = *reinterpret_cast<const uint2*>(cases + i * NC + lane * 8u);
                csw[0] = c2.x; csw[1] = c2.y;
            } else {
                csw[0] = csw[1] = 0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < NEW; ++k) entw[k] = 0;
#pragma unroll
            for (int k = 0; k < NCW; ++k) csw[k] = 0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                const uint32_t e = c < (uint32_t)NC ? (uint32_t)table[i * NC + c] : 0u;
                const uint32_t k = (want_cases && c < (uint32_t)NC) ? (uint32_t)cases[i * NC + c] : 0u;
                if (E16) entw[j / 2] |= e << (16 * (j & 1));
                else entw[j] = e;
                csw[j / 4] |= k << (8 * (j & 3));
            }
        }
    }
    __device__ __forceinline__ uint32_t ent(int j) const
    {
        return E16 ? (entw[j / 2] >> (16 * (j & 1))) & 0xFFFFu : entw[j];
    }
    __device__ __forceinline__ uint32_t cs(int j) const { return (csw[j / 4] >> (8 * (j & 3))) & 0xFFu; }
};

// Vertex pass: per wave and active block, stage the voxel tile [-A, B+A]^3
// (clamped to the volume: clamp-to-edge samples for the normals) in LDS, list
// the block's vertices (corner, axis) in prefix order and let the 64 lanes
// take consecutive entries: balanced work and coalesced writes.
template <int B, typename T, bool NORMALS, bool BATCH>
__global__ __launch_bounds__((64 * VertCfg<B, NORMALS, T>::WPG))
__attribute__((amdgpu_waves_per_eu(VertCfg<B, NORMALS, T>::MIN_WAVES, 8))) void k_extract_vertices(ExtractArgs a)
{
    using C = VertCfg<B, NORMALS, T>;
    constexpr int A = C::A, E = C::E, WPG = C::WPG;
    constexpr uint32_t RP = C::RP, NS = C::NS, NSEG = C::NSEG, ND = C::ND;
    constexpr uint32_t LIST = C::LIST;
    constexpr int CPL = EB<B>::CPL;
    using Entry = typename EB<B>::Entry;
    __shared__ __align__(16) T s_val_all[WPG][C::TILE];
    __shared__ uint16_t s_vl_all[WPG][LIST];
    __shared__ __align__(16) float4 s_vo_all[WPG][64];
    const uint32_t lane = lane_id(), w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T* s_val = s_val_all[w];
    uint16_t* s_vl = s_vl_all[w];
    float4* s_vo = s_vo_all[w];
    const uint32_t na = active_items(a);
    const Entry* __restrict__ table = static_cast<const Entry*>(a.corner_table);
    const WaveItems it = cost_items(a, na, WPG, a.blk_vbase, a.totals[0]);
    // the next block's count, bases, ID, batch unit and corner entries load
    // while this block's tile loads and its vertices are computed: one memory
    // round trip per block on the critical path (the tile) instead of three
    struct Pre {
        uint32_t bnv, id, unit;
        uint64_t vb;
        LaneCorners<B> lc;
    };
    auto fetch = [&](Pre& P, uint32_t i) {
        P.bnv = a.blk_counts[4 * (uint64_t)i];
        P.vb = a.blk_vbase[i];
        P.id = a.ids[i];
        P.unit = BATCH ? a.item_unit[i] : 0u;
        P.lc.load(table, nullptr, i, lane, false);
    };
    uint32_t i = it.first;
    Pre cur;
    if (i < it.end) fetch(cur, i);
    while (i < it.end) {
        const uint32_t inx = i + it.stride;
        Pre nxt = cur;
        if (inx < it.end) fetch(nxt, inx);
        const uint32_t bnv = cur.bnv;
        const uint64_t vb = cur.vb;
        bool work = bnv != 0;
        if (work && vb + bnv > a.vcap) {
            if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&a.totals[2]), 1ull);
            work = false;
        }
        if (work) {
        const ItemVol V = item_vol_of<BATCH>(a, __builtin_amdgcn_readfirstlane(cur.unit));
        const T* __restrict__ vol = static_cast<const T*>(V.data);
        const uint32_t X = V.X, Y = V.Y, Z = V.Z;
        const float iso = V.iso;
        const float inv2x = 2.0f / (float)X, inv2y = 2.0f / (float)Y, inv2z = 2.0f / (float)Z;
        const bool off32 = (uint64_t)X * Y * E < 0x80000000ull;
        const uint32_t id = cur.id;
        const uint32_t bx = id % V.gx, by = (id / V.gx) % V.gy;
        const uint32_t bz = id / (V.gx * V.gy);
        const uint32_t ox = bx * B, oy = by * B, oz = bz * B;
        const LaneCorners<B>& lc = cur.lc;
        // LDS voxel (tx, ty, tz) of the tile from (ox - A, oy - A, oz - A) sits at tx + RP * (ty + E * tz)
        const bool direct = off32 && ox >= (uint32_t)A && oy >= (uint32_t)A && oz >= (uint32_t)A &&
                            ox - A + RP <= X && oy + B + A < Y && oz + B + A < Z;
        if (direct) {
            // every segment lies inside its row of the volume
            const T* __restrict__ t0p = vol + ((ox - A) + (uint64_t)X * ((oy - A) + (uint64_t)Y * (oz - A)));
            // the lane index re-read per block (opaque to the compiler): the
            // segment offsets are not kept live across blocks
            uint32_t ln = lane;
            asm volatile("" : "+v"(ln));
            // tile rows (ty, tz) -- bit ty + E tz -- that the block's vertices
            // read: a vertex at core corner (Y0, Z0) = (ly + 1, lz + 1) reads the
            // rows plus(Y0, Z0) = (Y0, Z0) + {0, +-e_y, +-e_z} around its first
            // endpoint, and plus(Y0 + 1, Z0) or plus(Y0, Z0 + 1) around the other
            // for a y or z edge (an x edge stays in the row).  With B = 8 lane
            // ly + 8 lz holds the corner row (ly, lz), so three ballots (any
            // vertex, a y-edge vertex, a z-edge vertex) give the rows that are
            // read; rows outside the mask are not loaded.
            uint64_t need_lo = ~0ull, need_hi = ~0ull;
            if constexpr (B == 8 && NORMALS) {
                uint32_t f = 0;
#pragma unroll
                for (int j = 0; j < CPL; ++j) f |= lc.ent(j) & 7u;
                const uint64_t rows_a = __ballot(f != 0u);
                const uint64_t rows_y = __ballot((f & 2u) != 0u);
                const uint64_t rows_z = __ballot((f & 4u) != 0u);
                uint32_t mrow[E];
#pragma unroll
                for (int tz = 0; tz < E; ++tz) mrow[tz] = 0;
#pragma unroll
                for (int lz = 0; lz < B; ++lz) {
                    // bit ty = ly + 1 of the corner rows in plane lz
                    const uint32_t a1 = (uint32_t)((rows_a >> (8 * lz)) & 0xFFu) << 1;
                    const uint32_t y1 = (uint32_t)((rows_y >> (8 * lz)) & 0xFFu) << 1;
                    const uint32_t z1 = (uint32_t)((rows_z >> (8 * lz)) & 0xFFu) << 1;
                    mrow[lz] |= a1 | (y1 << 1);
                    mrow[lz + 1] |= a1 | (a1 << 1) | (a1 >> 1) | y1 | (y1 << 1) | (y1 << 2) | z1;
                    mrow[lz + 2] |= a1 | (y1 << 1) | z1 | (z1 << 1) | (z1 >> 1);
                    mrow[lz + 3] |= z1;
                }
                need_lo = need_hi = 0;
#pragma unroll
                for (int tz = 0; tz < E; ++tz) {
                    const int sh = E * tz;
                    if (sh < 64) {
                        need_lo |= (uint64_t)mrow[tz] << sh;
                        if (sh + E > 64) need_hi |= (uint64_t)mrow[tz] >> (64 - sh);
                    } else {
                        need_hi |= (uint64_t)mrow[tz] << (sh - 64);
                    }
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < ND; ++k) {
                const uint32_t m = min(ln + 64u * k, NSEG - 1u);
                const uint32_t r = m / NS;
                const T* src = t0p + (4u * (m % NS) + X * (r % (uint32_t)E + Y * (r / (uint32_t)E)));
                if ((((r < 64u) ? need_lo >> r : need_hi >> (r - 64u)) & 1u) != 0) {
                    if constexpr (C::SEGB == 16) lds_direct_load16(src, s_val + 256u * k);
                    else lds_direct_load4(src, s_val + 256u * k);
                }
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the segments are in LDS
        } else {
            // a block at a face of the volume: the same row segments, each row
            // clamped into the volume (clamp-to-edge); a segment that lies
            // inside its row loads straight to LDS, one that crosses x = 0 or
            // x = X is filled voxel by voxel with x clamped
            uint32_t fix = 0;   // bit k: this lane's segment k crosses a row end
#pragma unroll
            for (uint32_t k = 0; k < ND; ++k) {
                const uint32_t m = min(lane + 64u * k, NSEG - 1u);
                const uint32_t r = m / NS;
                const uint64_t gy = (uint64_t)min(max((int32_t)(oy + r % (uint32_t)E) - A, 0), (int32_t)Y - 1);
                const uint64_t gz = (uint64_t)min(max((int32_t)(oz + r / (uint32_t)E) - A, 0), (int32_t)Z - 1);
                const int32_t gx0 = (int32_t)(ox + 4u * (m % NS)) - A;
                if (gx0 >= 0 && gx0 + 4 <= (int32_t)X) {
                    const T* src = vol + ((uint64_t)gx0 + (uint64_t)X * (gy + (uint64_t)Y * gz));
                    if constexpr (C::SEGB == 16) lds_direct_load16(src, s_val + 256u * k);
                    else lds_direct_load4(src, s_val + 256u * k);
                } else {
                    fix |= 1u << k;
                }
            }
            if (fix) {
#pragma unroll
                for (uint32_t k = 0; k < ND; ++k) {
                    if (!((fix >> k) & 1u)) continue;
                    const uint32_t m = min(lane + 64u * k, NSEG - 1u);
                    const uint32_t r = m / NS;
                    const uint64_t gy = (uint64_t)min(max((int32_t)(oy + r % (uint32_t)E) - A, 0), (int32_t)Y - 1);
                    const uint64_t gz = (uint64_t)min(max((int32_t)(oz + r / (uint32_t)E) - A, 0), (int32_t)Z - 1);
                    const int32_t gx0 = (int32_t)(ox + 4u * (m % NS)) - A;
                    const T* row = vol + (uint64_t)X * (gy + (uint64_t)Y * gz);
                    T v4[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) v4[j] = row[min(max(gx0 + j, 0), (int32_t)X - 1)];
                    // the slot the segment's direct load would have filled
#pragma unroll
                    for (int j = 0; j < 4; ++j) s_val[256u * k + 4u * lane + (uint32_t)j] = v4[j];
                }
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the direct segments are in LDS
        }
        for (uint32_t c0 = 0; c0 < bnv; c0 += LIST) {
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                const uint32_t e = lc.ent(j);
                uint32_t run = e >> 3;
#pragma unroll
                for (uint32_t ax = 0; ax < 3; ++ax)
                    if ((e >> ax) & 1u) {
                        if (run - c0 < LIST) s_vl[run - c0] = (uint16_t)(c | (ax << 12));
                        ++run;
                    }
            }
            wave_lds_sync();
            const uint32_t n = min(LIST, bnv - c0);
            // rounds of 64 vertices, every lane taking part: a lane computes
            // vertex r0 + lane; each half of the round (lanes 0-31, then
            // 32-63) is staged in LDS and written as one contiguous 1 KB run
            // (a 32-byte vertex per lane would take two stores 32 bytes apart,
            // each line written half at a time)
            for (uint32_t r0 = 0; r0 < n; r0 += 64u) {
                const uint32_t v = r0 + lane;
                const bool live = v < n;
                const uint32_t e = s_vl[live ? v : r0];
                const uint32_t c = e & 0xFFFu, ax = e >> 12;
                const uint32_t lx = c % B, ly = (c / B) % B, lz = c / (B * B);
                const uint32_t x = ox + lx, y = oy + ly, z = oz + lz;
                const uint32_t t0 = (lx + A) + RP * ((ly + A) + E * (lz + A));
                const uint32_t st = ax == 0 ? 1u : (ax == 1 ? RP : RP * E);
                const float v0 = vfloat(s_val[t0]), v1 = vfloat(s_val[t0 + st]);
                const float d = v1 - v0;
                // t = 0.5 unless |v1 - v0| > 1e-5 (src/shaders/marching_cubes_pmb.mesh.glsl:129-134);
                // past that bound 1/d cannot overflow, and the reciprocal (1 ulp)
                // keeps t within ~2 ulp of the quotient
                float t = fabsf(d) > 1e-5f ? (iso - v0) * __builtin_amdgcn_rcpf(d) : 0.5f;
                t = fminf(fmaxf(t, 0.0f), 1.0f);
                const float px = (float)x + (ax == 0 ? t : 0.0f);
                const float py = (float)y + (ax == 1 ? t : 0.0f);
                const float pz = (float)z + (ax == 2 ? t : 0.0f);
                const float4 pos = make_float4(px * inv2x - 1.0f, py * inv2y - 1.0f, pz * inv2z - 1.0f, 1.0f);
                float4 nrm = make_float4(0.f, 0.f, 0.f, 0.f);
                if (NORMALS) {
                    float n0[3], n1[3];
                    grad_normal(s_val, t0, RP, RP * E, n0);
                    grad_normal(s_val, t0 + st, RP, RP * E, n1);
                    float m0 = n0[0] * (1.0f - t) + n1[0] * t;
                    float m1 = n0[1] * (1.0f - t) + n1[1] * t;
                    float m2 = n0[2] * (1.0f - t) + n1[2] * t;
                    float l2 = m0 * m0 + m1 * m1 + m2 * m2;
                    if (l2 < 1e-10f) { m0 = n0[0]; m1 = n0[1]; m2 = n0[2]; l2 = m0 * m0 + m1 * m1 + m2 * m2; }
                    const float r = rsqrtf(l2);
                    nrm = make_float4(m0 * r, m1 * r, m2 * r, 0.0f);
                }
                const uint32_t nh = 2 * min(64u, n - r0);   // 16-byte halves of this round
                float4* dst = reinterpret_cast<float4*>(a.vertices) + 2 * (vb + c0 + r0);
                // nontemporal: written once, read by the caller later
                if (lane < 32u) {
                    s_vo[2 * lane] = pos;
                    s_vo[2 * lane + 1] = nrm;
                }
                wave_lds_sync();
                const float4 h0 = s_vo[lane];
                if (lane < nh)
                    __builtin_nontemporal_store(mtx_f4v{h0.x, h0.y, h0.z, h0.w}, reinterpret_cast<mtx_f4v*>(dst + lane));
                wave_lds_sync();
                if (lane >= 32u) {
                    s_vo[2 * (lane - 32u)] = pos;
                    s_vo[2 * (lane - 32u) + 1] = nrm;
                }
                wave_lds_sync();
                const float4 h1 = s_vo[lane];
                if (64 + lane < nh)
                    __builtin_nontemporal_store(mtx_f4v{h1.x, h1.y, h1.z, h1.w},
                                                reinterpret_cast<mtx_f4v*>(dst + 64 + lane));
                wave_lds_sync();
                if (live) {
                    const uint64_t gv = vb + c0 + v;
                    if (a.edge_ids) a.edge_ids[gv] = 3ull * (x + (uint64_t)X * (y + (uint64_t)Y * z)) + ax;
                    if (a.edge_t) a.edge_t[gv] = t;
                    if (a.vox_pos) {
                        a.vox_pos[3 * gv + 0] = px;
                        a.vox_pos[3 * gv + 1] = py;
                        a.vox_pos[3 * gv + 2] = pz;
                    }
                }
            }
            wave_lds_sync();
        }
        }
        cur = nxt;
        i = inx;
    }
}

// Three consecutive uint32 (4-byte aligned) as one nontemporal 12-byte store:
// written once, read by the next stage or the caller, so kept out of the
// caches that the next step's min/max streams through.
typedef uint32_t mtx_u32x3 __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ void store3_nt(uint32_t* p, uint32_t x, uint32_t y, uint32_t z)
{
    const mtx_u32x3 v = { x, y, z };
    __builtin_nontemporal_store(v, reinterpret_cast<mtx_u32x3*>(p));
}

// Per-wave LDS table of the first vertex and the 3 edge flags at every owner
// corner of [0, B]^3, for the triangle passes.  PK (every vertex index below
// 2^28): one word per corner -- index in bits 0..27, flags in 28..30, bit 31
// set when the corner's owner block is not active -- so a triangle vertex
// resolves with one LDS read; otherwise a word of index and a byte of flags
// (0xF0: owner not active).
constexpr uint32_t kPackedIndexLimit = 1u << 28;

template <bool PK>
struct CornerLds {
    uint32_t* idx;
    uint8_t* fl;
    __device__ __forceinline__ void set(uint32_t q, uint32_t index, uint32_t flags) const
    {
        if constexpr (PK) {
            idx[q] = index | (flags << 28);
        } else {
            idx[q] = index;
            fl[q] = (uint8_t)flags;
        }
    }
    __device__ __forceinline__ void set_none(uint32_t q) const
    {
        if constexpr (PK) {
            idx[q] = 0x80000000u;
        } else {
            idx[q] = 0u;
            fl[q] = (uint8_t)0xF0u;
        }
    }
    // index of the vertex on the edge with rank mask m (the owned edges of
    // lower axis) at corner q; bad collects 0x80 (PK) or 0xF0 for a corner
    // whose owner is not active
    __device__ __forceinline__ uint32_t vertex(uint32_t q, uint32_t m, uint32_t& bad) const
    {
        if constexpr (PK) {
            const uint32_t e = idx[q];
            bad |= (e >> 24) & 0x80u;
            return (e & (kPackedIndexLimit - 1u)) + __popc((e >> 28) & m & 3u);
        } else {
            const uint32_t gf = fl[q];
            bad |= gf;
            return __popc(gf & m & 3u) + idx[q];
        }
    }
};

template <int B> struct TriCfg {
    static constexpr int F = B + 1;             // owner corners [0, B]^3
    static constexpr int NF = F * F * F;
    // triangles listed at a time: 512 keeps the B = 8 workgroup at 29 KB of
    // LDS (5 waves per SIMD, as its registers allow); blocks with more
    // triangles (dense noisy surfaces) are listed in several passes
    static constexpr int LIST = 512;
    // list entry: owner-corner index of the cell's (0,0,0) corner (QB bits)
    // | decoded-triangle index (case * 5 + k) << QB
    static constexpr int QB = NF <= 1024 ? 10 : 13;
    // decoded triangle: per vertex, DB bits = offset of its edge's owner
    // corner from the cell's (DB - 2 bits) | rank mask (1 << axis) - 1
    static constexpr int DB = F * F + F < 256 ? 10 : 20;
    using TDec = typename std::conditional<DB == 10, uint32_t, uint64_t>::type;
    static constexpr int WAVE_LDS = NF * 5 + LIST * 4 + 64;
    static constexpr int WG_LDS = 256 + 1280 * (int)sizeof(TDec) + 4 * (NF - B * B * B);
    static constexpr int WPG = WAVE_LDS * 4 + WG_LDS <= 65536 ? 4 : (WAVE_LDS * 2 + WG_LDS <= 65536 ? 2 : 1);
};

// Owners of the 12 cube edges, 5 bits each (dx | dy << 1 | dz << 2 | axis << 3).
constexpr uint64_t edge_owner_table()
{
    uint64_t t = 0;
    for (unsigned e = 0; e < 12u; ++e) t |= (uint64_t)mtx_edge_owner(e) << (5u * e);
    return t;
}
constexpr uint64_t kEdgeOwners = edge_owner_table();

// Face corner f of [0, B]^3 that a +x/+y/+z neighbour owns (a coordinate
// equal to B): x = B (F*F corners), then y = B with x < B (B*F), then z = B
// with x, y < B (B*B): F^3 - B^3 in all.
template <int B>
__device__ __forceinline__ void face_corner(uint32_t f, uint32_t& qx, uint32_t& qy, uint32_t& qz)
{
    constexpr uint32_t F = B + 1;
    if (f < F * F) {
        qx = B; qy = f % F; qz = f / F;
    } else if (f < F * F + B * F) {
        const uint32_t g = f - F * F;
        qx = g % B; qy = B; qz = g / B;
    } else {
        const uint32_t g = f - F * F - B * F;
        qx = g % B; qy = g / B; qz = B;
    }
}

// One block's inputs to the triangle pass.  They are loaded in three stages
// while the wave works on the previous block: (1) the block's ID, counts,
// bases and corner entries/cases; (2) the active indices of its +x/+y/+z
// neighbours (lanes 0..7, neighbour (lane & 1, lane >> 1 & 1, lane >> 2); the
// (1,1,1) one owns no edge a cell of this block uses); (3) the neighbours'
// first vertices and the entries of the face corners they own.
//
// Bases are kept as their low 32 bits (vertex indices and meshlet offsets are
// 32-bit; the index base stays 64-bit for the capacity check).  Face entries
// of 16 bits are packed two to a register; the all-ones entry marks a corner
// whose neighbour is not active (a real entry is at most (3 B^3) << 3 | 7).
template <int B>
struct TriBlock {
    using Entry = typename EB<B>::Entry;
    static constexpr bool E16 = sizeof(Entry) == 2;
    static constexpr int NFACE = (B + 1) * (B + 1) * (B + 1) - B * B * B;
    static constexpr int NPL = (NFACE + 63) / 64;
    static constexpr int NFW = E16 ? (NPL + 1) / 2 : NPL;
    static constexpr uint32_t kNone = E16 ? 0xFFFFu : 0xFFFFFFFFu;
    uint32_t id, bnv, bnt, unit_v, unit, first;
    uint32_t vb, mb, uv, ut, nbase;
    uint64_t tb;
    int32_t r;
    LaneCorners<B> lc;
    uint32_t fentw[NFW];
    __device__ __forceinline__ void set_face(int k, uint32_t e)
    {
        if (E16) fentw[k / 2] = (k & 1) ? (fentw[k / 2] & 0xFFFFu) | (e << 16) : (fentw[k / 2] & 0xFFFF0000u) | e;
        else fentw[k] = e;
    }
    __device__ __forceinline__ uint32_t face(int k) const
    {
        return E16 ? (fentw[k / 2] >> (16 * (k & 1))) & 0xFFFFu : fentw[k];
    }
};

// Triangle pass: per wave and active block, the global index of the first
// vertex at every owner corner of [0, B]^3 and its 3 edge flags go to LDS --
// the block's own corners from its count-pass entries, the corners on the
// +x/+y/+z faces from the neighbours' entries -- then the block's triangles
// are listed (cell, case, k) and resolved three LDS reads each by consecutive
// lanes.  No voxel is read.  The next block's inputs are loaded during this
// one, so no dependent global load waits at a block's start.
template <int B, bool BATCH, bool PK>
__global__ __launch_bounds__((64 * TriCfg<B>::WPG)) void k_extract_triangles(ExtractArgs a)
{
    using C = TriCfg<B>;
    using TB = TriBlock<B>;
    constexpr int WPG = C::WPG, F = C::F, NF = C::NF;
    constexpr uint32_t LIST = C::LIST;
    constexpr int NC = EB<B>::NC;
    constexpr int CPL = EB<B>::CPL;
    constexpr int NFACE = TB::NFACE, NPL = TB::NPL;
    using Entry = typename EB<B>::Entry;
    __shared__ uint32_t s_gidx_all[WPG][NF];
    __shared__ uint8_t s_gfl_all[PK ? 1 : WPG][PK ? 1 : NF];
    __shared__ uint32_t s_tl_all[WPG][LIST];
    using TDec = typename C::TDec;
    constexpr int QB = C::QB, DB = C::DB;
    __shared__ uint8_t s_ntri[256];
    __shared__ TDec s_tdec[256 * 5];
    // face corner f: its index in [0, B]^3 (bits 0..12), the neighbour that
    // owns it (13..15) and its corner in that neighbour's table (16..27)
    __shared__ uint32_t s_face[NFACE];
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_ntri[c] = (uint8_t)mtx_word_ntri(mtx_tri_words[c]);
    for (uint32_t x = threadIdx.x; x < 256u * 5u; x += blockDim.x) {
        const uint64_t wd = mtx_tri_words[x / 5u];
        const uint32_t k = x % 5u;
        TDec d = 0;
        if (k < mtx_word_ntri(wd)) {
            for (uint32_t j = 0; j < 3u; ++j) {
                const uint32_t o = (uint32_t)(kEdgeOwners >> (5u * mtx_word_edge(wd, k, j))) & 0x1Fu;
                const uint32_t off = (o & 1u) + F * (((o >> 1) & 1u) + F * ((o >> 2) & 1u));
                const uint32_t m = (1u << (o >> 3)) - 1u;
                d |= (TDec)(off | (m << (DB - 2))) << (DB * j);
            }
        }
        s_tdec[x] = d;
    }
    for (uint32_t f = threadIdx.x; f < (uint32_t)NFACE; f += blockDim.x) {
        uint32_t qx, qy, qz;
        face_corner<B>(f, qx, qy, qz);
        const uint32_t sx = qx == (uint32_t)B, sy = qy == (uint32_t)B, sz = qz == (uint32_t)B;
        s_face[f] = (qx + F * (qy + F * qz)) | ((sx | (sy << 1) | (sz << 2)) << 13) |
                    (((qx - B * sx) + B * ((qy - B * sy) + B * (qz - B * sz))) << 16);
    }
    __syncthreads();
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const CornerLds<PK> cor{ s_gidx_all[w], s_gfl_all[PK ? 0 : w] };
    uint32_t* s_tl = s_tl_all[w];
    const uint32_t na = active_items(a);
    const Entry* __restrict__ table = static_cast<const Entry*>(a.corner_table);

    // low words of the 64-bit bases
    const uint32_t* __restrict__ vb32 = reinterpret_cast<const uint32_t*>(a.blk_vbase);
    const uint32_t* __restrict__ tb32 = reinterpret_cast<const uint32_t*>(a.blk_tbase);
    const uint32_t* __restrict__ mb32 = reinterpret_cast<const uint32_t*>(a.blk_mbase);
    auto stage1 = [&](TB& P, uint32_t i) {
        P.bnv = a.blk_counts[4 * (uint64_t)i];
        P.bnt = a.blk_counts[4 * (uint64_t)i + 1];
        P.vb = vb32[2 * (uint64_t)i];
        P.tb = a.blk_tbase[i];
        P.mb = mb32[2 * (uint64_t)i];
        P.id = a.ids[i];
        if constexpr (BATCH) P.unit_v = a.item_unit[i];
        P.lc.load(table, a.cases, i, lane, true);
    };
    auto stage2 = [&](TB& P) {
        uint32_t gx, gy, gz;
        const uint32_t* map;
        if constexpr (BATCH) {
            P.unit = __builtin_amdgcn_readfirstlane(P.unit_v);
            const BatchUnit& U = a.units[P.unit];
            gx = U.gx; gy = U.gy; gz = U.gz;
            map = as_global(U.map);
            P.first = a.unit_first[P.unit];
        } else {
            gx = a.grid.gx; gy = a.grid.gy; gz = a.grid.gz;
            map = a.blk_map;
        }
        const uint32_t bx = P.id % gx, by = (P.id / gx) % gy, bz = P.id / (gx * gy);
        const uint32_t nx = bx + (lane & 1u), ny = by + ((lane >> 1) & 1u), nz = bz + ((lane >> 2) & 1u);
        P.r = -1;
        if (lane != 0 && lane < 7 && nx < gx && ny < gy && nz < gz)
            P.r = (int32_t)map[nx + gx * (ny + (uint64_t)gy * nz)];
    };
    auto stage3 = [&](TB& P) {
        if constexpr (BATCH) {
            P.uv = vb32[2 * (uint64_t)P.first];
            P.ut = tb32[2 * (uint64_t)P.first];
        } else {
            P.uv = 0;
            P.ut = 0;
        }
        P.nbase = (lane < 8 && P.r >= 0) ? vb32[2 * (uint64_t)P.r] : 0u;
#pragma unroll
        for (int k = 0; k < TB::NFW; ++k) P.fentw[k] = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < NPL; ++k) {
            const uint32_t f = lane + 64u * k;
            const uint32_t fd = s_face[f < (uint32_t)NFACE ? f : 0u];
            const int32_t jn = __shfl(P.r, (int)((fd >> 13) & 7u), 64);
            if (f < (uint32_t)NFACE && jn >= 0) P.set_face(k, (uint32_t)table[(uint64_t)jn * NC + (fd >> 16)]);
        }
    };

    const WaveItems it = cost_items(a, na, WPG, a.blk_tbase, a.totals[1]);
    TB cur, nxt;
    uint32_t i = it.first;
    if (i < it.end) {
        stage1(cur, i);
        stage2(cur);
        stage3(cur);
    }
    while (i < it.end) {
        const uint32_t inx = i + it.stride;
        const bool more = inx < it.end;
        if (more) stage1(nxt, inx);
        const uint32_t bnt = cur.bnt;
        bool work = bnt != 0;
        if (work && 3 * (cur.tb + bnt) > a.icap) {
            if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&a.totals[2]), 1ull);
            work = false;
        }
        uint32_t nt = 0;
        constexpr bool PACKN = CPL * 3 <= 32;   // per-cell triangle counts kept in one register
        uint32_t npk = 0;
        if (work) {
            // own corners: vertex index = block base + prefix from the count pass
            const uint32_t vrel = cur.vb - cur.uv;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                if (c < (uint32_t)NC) {
                    const uint32_t q = c % B + F * ((c / B) % B + F * (c / (B * B)));
                    cor.set(q, vrel + (cur.lc.ent(j) >> 3), cur.lc.ent(j) & 7u);
                }
                const uint32_t n = s_ntri[cur.lc.cs(j)];
                nt += n;
                if (PACKN) npk |= n << (3 * j);
            }
            // neighbour-owned face corners
#pragma unroll
            for (int k = 0; k < NPL; ++k) {
                const uint32_t f = lane + 64u * k;
                const uint32_t fd = s_face[f < (uint32_t)NFACE ? f : 0u];
                const uint32_t nb = (uint32_t)__shfl((int)cur.nbase, (int)((fd >> 13) & 7u), 64) - cur.uv;
                if (f < (uint32_t)NFACE) {
                    const uint32_t q = fd & 0x1FFFu;
                    const uint32_t e = cur.face(k);
                    if (e != TB::kNone) cor.set(q, nb + (e >> 3), e & 7u);
                    else cor.set_none(q);
                }
            }
        }
        if (more) stage2(nxt);
        if (work) {
            wave_lds_sync();
            const uint64_t tb = cur.tb;
            uint32_t tt;
            const uint32_t et = wave_excl_scan<uint32_t>(nt, tt);
            uint32_t bad = 0;
            for (uint32_t c0 = 0; c0 < tt; c0 += LIST) {
                // list entries in output order: cell by cell, the case's triangles k = 0..4
                uint32_t run = et - c0;
                // re-read per pass (opaque to the compiler): nothing derived
                // from them stays live across passes, which are rare
                LaneCorners<B> lc = cur.lc;
#pragma unroll
                for (int k = 0; k < LaneCorners<B>::NCW; ++k) asm volatile("" : "+v"(lc.csw[k]));
                asm volatile("" : "+v"(npk));
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    const uint32_t c = lane * CPL + j;
                    if (c < (uint32_t)NC) {
                        const uint32_t cs = lc.cs(j);
                        const uint32_t n = PACKN ? (npk >> (3 * j)) & 7u : (uint32_t)s_ntri[cs];
                        const uint32_t ent = (c % B + F * ((c / B) % B + F * (c / (B * B)))) | ((cs * 5u) << QB);
#pragma unroll
                        for (uint32_t k = 0; k < 5u; ++k)
                            if (k < n && run + k < LIST) s_tl[run + k] = ent + (k << QB);
                        run += n;
                    }
                }
                wave_lds_sync();
                const uint32_t n = min(LIST, tt - c0);
                for (uint32_t q = lane; q < n; q += 64u) {
                    const uint32_t e = s_tl[q];
                    const uint32_t q0 = e & ((1u << QB) - 1u);
                    const TDec d = s_tdec[e >> QB];
                    uint32_t idx3[3];
#pragma unroll
                    for (uint32_t e3 = 0; e3 < 3; ++e3) {
                        const uint32_t f = (uint32_t)(d >> (DB * e3));
                        const uint32_t qq = q0 + (f & ((1u << (DB - 2)) - 1u));
                        idx3[e3] = cor.vertex(qq, f >> (DB - 2), bad);
                    }
                    // one 12-byte store per lane: the wave writes 768 contiguous bytes
                    store3_nt(a.indices + 3 * (tb + c0 + q), idx3[0], idx3[1], idx3[2]);
                }
                wave_lds_sync();
            }
            // an active edge whose owner block is not in the active list: the
            // list does not belong to this volume/isovalue
            if (bad & 0xF0u) atomicOr(reinterpret_cast<unsigned long long*>(&a.totals[2]), 2ull);
            wave_lds_sync();
        }
        if (more) stage3(nxt);
        cur = nxt;
        i = inx;
    }
}

// ---------------------------------------------------------------------------
// meshlet passes
//
// Cut (between the count pass and the scan): one wave per active block reads
// the block's case indices, cuts its occupied cells into meshlets and writes
// the block's meshlet and meshlet-vertex counts (the scan turns them into
// bases) and its start bits.  Emit (after the triangle pass): one wave per
// active block rebuilds, from the start bits and the cases, each cell's added
// edges and first meshlet vertex; then lists the block's triangles in output
// order and, for each triangle vertex, finds the cell of the meshlet that
// first uses its edge: the meshlet index is that cell's first meshlet vertex
// plus the edge's rank among the edges the cell adds, and the vertex-map entry
// of that meshlet vertex is the triangle's own (shared-mesh) index.  Finally
// the meshlet descriptors.
// ---------------------------------------------------------------------------

// This lane's CPL cell cases of active block i (0 for cells outside the volume).
template <int B>
struct LaneCases {
    static constexpr int CPL = EB<B>::CPL, NC = EB<B>::NC, NCW = (CPL + 3) / 4;
    uint32_t w[NCW];
    __device__ __forceinline__ void load(const uint8_t* __restrict__ cases, uint64_t i, uint32_t lane)
    {
        if constexpr (CPL == 8) {
            const uint2 c2 = *reinterpret_cast<const uint2*>(cases + i * NC + lane * 8u);
            w[0] = c2.x;
            w[1] = c2.y;
        } else {
#pragma unroll
            for (int k = 0; k < NCW; ++k) w[k] = 0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                if (c < (uint32_t)NC) w[j / 4] |= (uint32_t)cases[i * NC + c] << (8 * (j & 3));
            }
        }
    }
    __device__ __forceinline__ uint32_t cs(int j) const { return (w[j / 4] >> (8 * (j & 3))) & 0xFFu; }
};

// Origin and valid-cell extents (10 bits per axis) of block `id`.
__device__ __forceinline__ void block_cells(const ExtractArgs& a, uint32_t id, uint32_t& vlim)
{
    const uint32_t gx = a.grid.gx, gy = a.grid.gy, B = a.grid.B;
    const uint32_t ox = (id % gx) * B, oy = ((id / gx) % gy) * B, oz = (id / (gx * gy)) * B;
    const uint32_t vx = a.vol.X - 1 > ox ? min(a.vol.X - 1 - ox, B) : 0u;
    const uint32_t vy = a.vol.Y - 1 > oy ? min(a.vol.Y - 1 - oy, B) : 0u;
    const uint32_t vz = a.vol.Z - 1 > oz ? min(a.vol.Z - 1 - oz, B) : 0u;
    vlim = vx | (vy << 10) | (vz << 20);
}

template <int B>
__device__ __forceinline__ uint32_t nbrs_of_cell(uint32_t c, uint32_t vlim)
{
    return cell_nbrs(c % B, (c / B) % B, c / (B * B), vlim & 1023u, (vlim >> 10) & 1023u, vlim >> 20);
}

template <int B>
__global__ __launch_bounds__(256) void k_meshlet_cut(ExtractArgs a)
{
    constexpr int CPL = EB<B>::CPL;
    constexpr uint32_t NW = StartBits<B>::NW;
    __shared__ uint8_t s_ntri[256];
    __shared__ uint16_t s_emu[256];
    __shared__ uint16_t s_newtab[64 * kBackClasses];
    __shared__ uint8_t s_cls[B * B + B];
    __shared__ uint32_t s_sb_all[4][NW];
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_ntri[c] = (uint8_t)mtx_word_ntri(mtx_tri_words[c]);
    meshlet_tables<B>(s_emu, s_newtab, s_cls);
    __syncthreads();
    const uint32_t lane = lane_id();
    uint32_t* s_sb = s_sb_all[threadIdx.x >> 6];
    const uint32_t na = active_items(a);
    const WaveItems it = wave_items(na, 4);
    // the next block's ID and cases load while this block is cut
    uint32_t id_n = 0;
    LaneCases<B> lc_n;
    if (it.first < it.end) {
        id_n = a.ids[it.first];
        lc_n.load(a.cases, it.first, lane);
    }
    for (uint32_t i = it.first; i < it.end; i += it.stride) {
        const uint32_t id = id_n;
        const LaneCases<B> lc = lc_n;
        if (i + it.stride < it.end) {
            id_n = a.ids[i + it.stride];
            lc_n.load(a.cases, i + it.stride, lane);
        }
        uint32_t vlim;
        block_cells(a, id, vlim);
        uint32_t nm, ns;
        uint64_t sb;
        meshlet_cut<B, CPL>(
            [&](int j, uint32_t& cs, uint32_t& nb) {
                cs = lc.cs(j);
                nb = nbrs_of_cell<B>(lane * CPL + j, vlim);
            },
            s_ntri, s_emu, s_newtab, s_cls, nm, ns, sb);
        if constexpr (CPL == 8 && NW == 16) {
            // lane l's 8 cells are byte l of the block's 64
            reinterpret_cast<uint8_t*>(a.starts)[(uint64_t)i * 64 + lane] = (uint8_t)sb;
        } else {
            for (uint32_t k = lane; k < NW; k += 64u) s_sb[k] = 0;
            wave_lds_sync();
            for (uint64_t m = sb; m; m &= m - 1) {
                const uint32_t c = lane * CPL + (uint32_t)__builtin_ctzll(m);
                atomicOr(&s_sb[c >> 5], 1u << (c & 31u));
            }
            wave_lds_sync();
            for (uint32_t k = lane; k < NW; k += 64u) a.starts[(uint64_t)i * NW + k] = s_sb[k];
            wave_lds_sync();
        }
        if (lane == 0) *reinterpret_cast<uint2*>(a.blk_counts + 4 * (uint64_t)i + 2) = make_uint2(nm, ns);
    }
}

template <int B> struct EmitCfg {
    static constexpr int NC = B * B * B;
    static constexpr int LIST = 256;             // triangles listed at a time
    // list entry: cell index (CB bits) | decoded triangle (case * 5 + k) << CB
    // [| B <= 8: the cell's distance back to its meshlet's first cell << CB + 11]
    static constexpr int CB = B <= 8 ? 9 : 12;
    static constexpr bool KIN = B <= 8;
    static constexpr int MAXM = NC / 5 + 2;      // every meshlet but a block's last has >= 5 cells
    static constexpr int TL = MAXM > LIST ? MAXM : LIST;
    static constexpr int WAVE_LDS = NC * 4 + NC + (KIN ? 0 : NC * 2) + TL * 4 + 3 * LIST * 4;
    static constexpr int WG_LDS = 256 + 512 + 64 * (int)kBackClasses * 2 + 256 * 5 * 2 + 64 * 12 * 8 + NC * 0;
    static constexpr int WPG = WAVE_LDS * 4 + WG_LDS <= 65536 ? 4 : (WAVE_LDS * 2 + WG_LDS <= 65536 ? 2 : 1);
};

// One block's inputs to the emit pass (loaded while the previous block runs).
template <int B>
struct EmitBlock {
    uint4 cnt;          // vertices, triangles, meshlets, meshlet vertices
    uint64_t tb;        // first triangle
    uint32_t mb, msb;   // first meshlet, first meshlet vertex
    uint32_t id;
    uint64_t sbits;     // this lane's start bits
    LaneCases<B> lc;
};

template <int B>
__global__ __launch_bounds__(64 * EmitCfg<B>::WPG, B <= 8 ? 4 : 1) void k_meshlet_emit(ExtractArgs a)
{
    using C = EmitCfg<B>;
    using EBk = EmitBlock<B>;
    constexpr int NC = EB<B>::NC, CPL = EB<B>::CPL, CB = C::CB;
    constexpr uint32_t LIST = C::LIST, NW = StartBits<B>::NW;
    constexpr bool KIN = C::KIN;
    constexpr bool PACKN = CPL * 3 <= 32;
    constexpr int WPG = C::WPG;
    __shared__ uint8_t s_ntri[256];
    __shared__ uint16_t s_emu[256];
    __shared__ uint16_t s_newtab[64 * kBackClasses];
    __shared__ uint8_t s_cls[B * B + B];
    __shared__ uint16_t s_tu[256 * 5];          // owner indices of each decoded triangle's vertices
    __shared__ uint64_t s_fu[64 * 12];          // first-user table
    // per wave: each cell's first meshlet vertex in the block (20 bits) | its
    // added edges << 20; the cells' distance back to their meshlet's start when
    // the list entry cannot hold it; the triangle list (then the descriptors'
    // inputs); the listed triangles' shared-mesh indices
    __shared__ uint32_t s_mc_all[WPG][NC];
    __shared__ uint8_t s_nb_all[WPG][NC];      // each cell's neighbour bits
    __shared__ uint16_t s_mk_all[KIN ? 1 : WPG][KIN ? 1 : NC];
    __shared__ uint32_t s_tl_all[WPG][C::TL];
    __shared__ __align__(16) uint32_t s_idx_all[WPG][3 * LIST];
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_ntri[c] = (uint8_t)mtx_word_ntri(mtx_tri_words[c]);
    for (uint32_t x = threadIdx.x; x < 256u * 5u; x += blockDim.x) {
        const uint64_t wd = mtx_tri_words[x / 5u];
        const uint32_t k = x % 5u;
        uint32_t tu = 0;
        if (k < mtx_word_ntri(wd))
            for (uint32_t j = 0; j < 3u; ++j) tu |= mtx_edge_uidx(mtx_word_edge(wd, k, j)) << (4u * j);
        s_tu[x] = (uint16_t)tu;
    }
    meshlet_tables<B>(s_emu, s_newtab, s_cls);
    first_user_table<B>(s_fu);
    __syncthreads();
    const uint32_t lane = lane_id(), w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* s_mc = s_mc_all[w];
    uint8_t* s_nb = s_nb_all[w];
    uint16_t* s_mk = s_mk_all[KIN ? 0 : w];
    uint32_t* s_tl = s_tl_all[w];
    uint32_t* s_idx = s_idx_all[w];
    const uint32_t* __restrict__ mb32 = reinterpret_cast<const uint32_t*>(a.blk_mbase);
    const uint32_t* __restrict__ sb32 = reinterpret_cast<const uint32_t*>(a.blk_sbase);
    auto load = [&](EBk& P, uint32_t i) {
        P.cnt = *reinterpret_cast<const uint4*>(a.blk_counts + 4 * (uint64_t)i);
        P.tb = a.blk_tbase[i];
        P.mb = mb32[2 * (uint64_t)i];
        P.msb = sb32[2 * (uint64_t)i];
        P.id = a.ids[i];
        if constexpr (CPL == 8 && NW == 16)
            P.sbits = reinterpret_cast<const uint8_t*>(a.starts)[(uint64_t)i * 64 + lane];
        else
            P.sbits = lane_start_bits<B, CPL>(a.starts + (uint64_t)i * NW, lane);
        P.lc.load(a.cases, i, lane);
    };
    // the indices of triangles [c0, c0 + n) of a block go to LDS with direct
    // loads (4 bytes per lane, completion on the vector-memory counter)
    auto stage_indices = [&](uint64_t tb, uint32_t c0, uint32_t n) {
        const uint32_t nd = 3 * n;
        const uint32_t* src = a.indices + 3 * (tb + c0);
        for (uint32_t k = 0; k < nd; k += 64u)
            if (k + lane < nd) lds_direct_load4(src + k + lane, s_idx + k);
    };
    const uint32_t na = active_items(a);
    const WaveItems it = wave_items(na, WPG);
    EBk nxt;
    if (it.first < it.end) load(nxt, it.first);
    for (uint32_t i = it.first; i < it.end; i += it.stride) {
        const EBk cur = nxt;
        if (i + it.stride < it.end) load(nxt, i + it.stride);
        const uint32_t bnt = cur.cnt.y, bnm = cur.cnt.z, bns = cur.cnt.w;
        if (bnt == 0) continue;
        const uint64_t tb = cur.tb;
        const uint32_t mb = cur.mb, msb = cur.msb;
        if (3 * (tb + bnt) > a.icap || (uint64_t)msb + bns > a.mvcap || (uint64_t)mb + bnm > a.mcap ||
            bnm > (uint32_t)C::MAXM) {
            if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&a.totals[2]), 1ull);
            continue;
        }
        // the first list's indices load while the cells are set up
        stage_indices(tb, 0, min(LIST, bnt));
        uint32_t vlim;
        block_cells(a, cur.id, vlim);
        const uint64_t sbits = cur.sbits;
        const uint32_t S0 =
            wave_excl_max(sbits ? lane * CPL + 63u - (uint32_t)__builtin_clzll(sbits) : 0u);
        // each cell's added edges and distance back to its meshlet's start,
        // its triangle count; the block's meshlet vertices and triangles before it
        constexpr bool KEEPM = CPL <= 8;   // the added-edge masks stay in registers
        uint32_t S = S0, cnt_v = 0, npk = 0, nt = 0;
        uint32_t mkeep[KEEPM ? CPL : 1];
        auto added = [&](int j, uint32_t& k) -> uint32_t {
            const uint32_t c = lane * CPL + j;
            if ((sbits >> j) & 1ull) S = c;
            k = c - S;
            const uint32_t cs = cur.lc.cs(j);
            if (c >= (uint32_t)NC || !s_ntri[cs]) return 0u;
            return added_edges<B>(s_emu[cs], nbrs_of_cell<B>(c, vlim), k, s_newtab, s_cls);
        };
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            uint32_t k;
            const uint32_t m = added(j, k);
            if constexpr (KEEPM) mkeep[j] = m;
            cnt_v += __popc(m);
            const uint32_t n = s_ntri[cur.lc.cs(j)];
            nt += n;
            if (PACKN) npk |= n << (3 * j);
        }
        uint32_t ns, tt;
        uint32_t slot = wave_excl_scan<uint32_t>(cnt_v, ns);
        const uint32_t et = wave_excl_scan<uint32_t>(nt, tt);
        const uint32_t nm = wave_sum<uint32_t>((uint32_t)__popcll(sbits));
        if (ns != bns || nm != bnm || tt != bnt) {
            // the cut and the counts disagree with this block's data
            if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&a.totals[2]), 4ull);
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the staged loads have landed
            continue;
        }
        S = S0;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const uint32_t c = lane * CPL + j;
            uint32_t k, m;
            if constexpr (KEEPM) {
                if ((sbits >> j) & 1ull) S = c;
                k = c - S;
                m = mkeep[j];
            } else {
                m = added(j, k);
            }
            if (c < (uint32_t)NC) {
                s_mc[c] = slot | (m << 20);
                s_nb[c] = (uint8_t)nbrs_of_cell<B>(c, vlim);
                if constexpr (!KIN) s_mk[c] = (uint16_t)k;
            }
            slot += __popc(m);
        }
        for (uint32_t c0 = 0; c0 < tt; c0 += LIST) {
            const uint32_t n = min(LIST, tt - c0);
            if (c0) stage_indices(tb, c0, n);
            // list entries in output order: cell by cell, the case's triangles k = 0..4
            uint32_t run = et - c0;
            S = S0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                if ((sbits >> j) & 1ull) S = c;
                if (c < (uint32_t)NC) {
                    const uint32_t cs = cur.lc.cs(j);
                    const uint32_t nn = PACKN ? (npk >> (3 * j)) & 7u : (uint32_t)s_ntri[cs];
                    uint32_t ent = c | ((cs * 5u) << CB);
                    if constexpr (KIN) ent |= (c - S) << (CB + 11);
#pragma unroll
                    for (uint32_t k = 0; k < 5u; ++k)
                        if (k < nn && run + k < LIST) s_tl[run + k] = ent + (k << CB);
                    run += nn;
                }
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the indices are in LDS
            wave_lds_sync();
            // one triangle vertex per lane: its edge's first user in the
            // meshlet and the edge's rank there give the meshlet vertex; that
            // vertex holds this triangle vertex (every triangle using it writes
            // the same shared-mesh index)
            uint32_t* dst = a.mindices + 3 * (tb + c0);
            for (uint32_t v = lane; v < 3 * n; v += 64u) {
                const uint32_t q = v / 3u, e3 = v - 3u * q;
                const uint32_t e = s_tl[q];
                const uint32_t c = e & ((1u << CB) - 1u);
                const uint32_t u = ((uint32_t)s_tu[(e >> CB) & 2047u] >> (4 * e3)) & 15u;
                const uint32_t kk = KIN ? e >> (CB + 11) : (uint32_t)s_mk[c];
                uint32_t back, u1;
                first_user_of(s_fu[(uint32_t)s_nb[c] * 12u + u], u, kk, back, u1);
                const uint32_t mc = s_mc[c - back];
                const uint32_t mv = msb + (mc & 0xFFFFFu) + __popc((mc >> 20) & ((1u << u1) - 1u));
                a.mvmap[mv] = s_idx[v];
                __builtin_nontemporal_store(mv, dst + v);
            }
            wave_lds_sync();
        }
        // descriptors: each meshlet's first meshlet vertex and first triangle
        // within the block (16 bits each), staged at its index
        uint32_t nmt;
        uint32_t m = wave_excl_scan<uint32_t>((uint32_t)__popcll(sbits), nmt);
        uint32_t tp = et;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const uint32_t c = lane * CPL + j;
            if ((sbits >> j) & 1ull) s_tl[m++] = (s_mc[c] & 0xFFFFu) | (tp << 16);
            tp += PACKN ? (npk >> (3 * j)) & 7u : (uint32_t)s_ntri[cur.lc.cs(j)];
        }
        wave_lds_sync();
        const uint32_t tb_rel = (uint32_t)tb;
        for (uint32_t k = lane; k < nm; k += 64u) {
            const uint32_t A = s_tl[k];
            const uint32_t Bn = k + 1 < nm ? s_tl[k + 1] : (bns | (bnt << 16));
            a.meshlets[(uint64_t)mb + k] = make_uint4(msb + (A & 0xFFFFu), 3u * (tb_rel + (A >> 16)),
                                                      (Bn & 0xFFFFu) - (A & 0xFFFFu), (Bn >> 16) - (A >> 16));
        }
        wave_lds_sync();
    }
}

// Triangle and meshlet pass in one launch (single volume, meshlets on).  Per
// wave and active block the block's triangles are listed once, and every
// triangle vertex is resolved twice: to the shared mesh (the first vertex at
// its edge's owner corner, from the block's and its +x/+y/+z neighbours'
// count-pass entries, as k_extract_triangles does) and to its meshlet (the
// edge's first user in the meshlet and the edge's rank among that cell's
// added edges, as k_meshlet_emit does).  One pass writes the index buffer,
// the meshlet indices, the meshlet vertex map and the descriptors; the
// separate passes listed the triangles twice and the emit read the index
// buffer back.  The next block's inputs load during this one.
template <int B> struct TriMeshCfg {
    static constexpr int NC = B * B * B;
    static constexpr int F = B + 1, NF = F * F * F, NFACE = NF - NC;
    static constexpr int LIST = 256;             // triangles listed at a time
    // list entry: cell index (CB bits) | decoded triangle (case * 5 + k) << CB
    // [| B <= 8: the cell's distance back to its meshlet's first cell << CB + 11]
    static constexpr int CB = B <= 8 ? 9 : 12;
    static constexpr bool KIN = B <= 8;
    static constexpr int MAXM = NC / 5 + 2;      // every meshlet but a block's last has >= 5 cells
    static constexpr int TL = MAXM > LIST ? MAXM : LIST;
    using TDec = typename TriCfg<B>::TDec;
    static constexpr int DB = TriCfg<B>::DB;
    // B = 8: the block's occupied cells are listed once and every per-cell
    // step after the corner tables runs over the list, 64 cells per pass
    static constexpr bool LISTED = B == 8;
    static constexpr int WAVE_LDS = NF * 5 + NC * 4 + (KIN ? 0 : NC * 2) + TL * 4 + (LISTED ? NC * 4 : 0);
    static constexpr int WG_LDS = 256 + 1280 * 8 + 4 * NFACE + 512 + 64 * (int)kBackClasses * 2 + (B * B + B) +
                                  (LISTED ? 64 * 12 * (int)kBackClasses + 16 : 64 * 12 * 8);
    static constexpr int WPG = WAVE_LDS * 4 + WG_LDS <= 65536 ? 4 : (WAVE_LDS * 2 + WG_LDS <= 65536 ? 2 : 1);
    static constexpr bool FITS = WAVE_LDS * WPG + WG_LDS <= 65536 && sizeof(TDec) == 4;
};

template <int B>
struct TriMeshBlock : TriBlock<B> {
    uint32_t bnm, bns, msb;
    uint64_t sbits;   // this lane's meshlet start bits
};

template <int B, bool PK>
__global__ __launch_bounds__(64 * TriMeshCfg<B>::WPG) void k_tri_meshlets(ExtractArgs a)
{
    using C = TriMeshCfg<B>;
    using TB = TriMeshBlock<B>;
    constexpr int WPG = C::WPG, F = C::F, NF = C::NF, NC = C::NC, CB = C::CB, DB = C::DB;
    constexpr int CPL = EB<B>::CPL;
    constexpr uint32_t LIST = C::LIST, NW = StartBits<B>::NW;
    constexpr bool KIN = C::KIN, LISTED = C::LISTED;
    // B = 8 with packed corners: the first-user table by distance class
    constexpr bool FU8 = LISTED && PK;
    constexpr bool PACKN = CPL * 3 <= 32;
    constexpr int NFACE = C::NFACE, NPL = TriBlock<B>::NPL;
    using TDec = typename C::TDec;
    static_assert(sizeof(TDec) == 4, "the decoded triangle shares a 64-bit word with its owner indices");
    using Entry = typename EB<B>::Entry;
    __shared__ uint8_t s_ntri[256];
    // per decoded triangle: owner-corner offsets and rank masks (low word),
    // the owner indices of its vertices (bits 32..43)
    __shared__ uint64_t s_tdu[256 * 5];
    __shared__ uint32_t s_face[NFACE];
    __shared__ uint16_t s_emu[256];
    __shared__ uint16_t s_newtab[64 * kBackClasses];
    __shared__ uint8_t s_cls[B * B + B];
    __shared__ uint64_t s_fu[FU8 ? 1 : 64 * 12];                  // first-user table
    __shared__ uint8_t s_fu8[FU8 ? 64 * 12 * kBackClasses : 1];   // ... by distance class
    __shared__ uint8_t s_bd[FU8 ? 16 : 1];
    // per wave: the first vertex and edge flags of every owner corner of
    // [0, B]^3; each cell's first meshlet vertex in the block (20 bits) | its
    // added edges << 20, and neighbour bits; the cells' distance back to their
    // meshlet's start when the list entry cannot hold it; the triangle list
    // (then the descriptors' inputs)
    __shared__ uint32_t s_gidx_all[WPG][NF];
    __shared__ uint8_t s_gfl_all[PK ? 1 : WPG][PK ? 1 : NF];
    __shared__ __align__(16) uint32_t s_mc_all[WPG][NC];
    __shared__ uint16_t s_mk_all[KIN ? 1 : WPG][KIN ? 1 : NC];
    __shared__ uint32_t s_tl_all[WPG][C::TL];
    // B = 8: the occupied cells in cell order: cell | case << 9 | distance back
    // to its meshlet's start << 17 | triangles << 26 | starts a meshlet << 29
    __shared__ uint32_t s_ls_all[LISTED ? WPG : 1][LISTED ? NC : 1];
    for (uint32_t c = threadIdx.x; c < 256u; c += blockDim.x) s_ntri[c] = (uint8_t)mtx_word_ntri(mtx_tri_words[c]);
    for (uint32_t x = threadIdx.x; x < 256u * 5u; x += blockDim.x) {
        const uint64_t wd = mtx_tri_words[x / 5u];
        const uint32_t k = x % 5u;
        TDec d = 0;
        uint32_t tu = 0;
        if (k < mtx_word_ntri(wd)) {
            for (uint32_t j = 0; j < 3u; ++j) {
                const uint32_t e = mtx_word_edge(wd, k, j);
                const uint32_t o = (uint32_t)(kEdgeOwners >> (5u * e)) & 0x1Fu;
                const uint32_t off = (o & 1u) + F * (((o >> 1) & 1u) + F * ((o >> 2) & 1u));
                const uint32_t m = (1u << (o >> 3)) - 1u;
                d |= (TDec)(off | (m << (DB - 2))) << (DB * j);
                tu |= mtx_edge_uidx(e) << (4u * j);
            }
        }
        s_tdu[x] = (uint64_t)d | ((uint64_t)tu << 32);
    }
    for (uint32_t f = threadIdx.x; f < (uint32_t)NFACE; f += blockDim.x) {
        uint32_t qx, qy, qz;
        face_corner<B>(f, qx, qy, qz);
        const uint32_t sx = qx == (uint32_t)B, sy = qy == (uint32_t)B, sz = qz == (uint32_t)B;
        s_face[f] = (qx + F * (qy + F * qz)) | ((sx | (sy << 1) | (sz << 2)) << 13) |
                    (((qx - B * sx) + B * ((qy - B * sy) + B * (qz - B * sz))) << 16);
    }
    meshlet_tables<B>(s_emu, s_newtab, s_cls);
    if constexpr (FU8) first_user_bytes<B>(s_fu8, s_bd, s_cls);
    else first_user_table<B>(s_fu);
    __syncthreads();
    const uint32_t lane = lane_id(), w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const CornerLds<PK> cor{ s_gidx_all[w], s_gfl_all[PK ? 0 : w] };
    uint32_t* s_mc = s_mc_all[w];
    uint16_t* s_mk = s_mk_all[KIN ? 0 : w];
    uint32_t* s_tl = s_tl_all[w];
    uint32_t* s_ls = s_ls_all[LISTED ? w : 0];
    const uint32_t na = active_items(a);
    // argument pointers re-read where used (the block loop's scalar registers
    // are taken by the two blocks in flight)
    const ArgRef<ExtractArgs> AR;
    auto table_ = [&]() { return static_cast<const Entry*>(AR->corner_table); };
    auto vb32_ = [&]() { return reinterpret_cast<const uint32_t*>(AR->blk_vbase); };
    const uint32_t gx = a.grid.gx, gy = a.grid.gy, gz = a.grid.gz;
    auto stage1 = [&](TB& P, uint32_t i) {
        const uint4 cnt = *reinterpret_cast<const uint4*>(AR->blk_counts + 4 * (uint64_t)i);
        P.bnv = cnt.x;
        P.bnt = cnt.y;
        P.bnm = cnt.z;
        P.bns = cnt.w;
        P.vb = vb32_()[2 * (uint64_t)i];
        P.tb = AR->blk_tbase[i];
        P.mb = reinterpret_cast<const uint32_t*>(AR->blk_mbase)[2 * (uint64_t)i];
        P.msb = reinterpret_cast<const uint32_t*>(AR->blk_sbase)[2 * (uint64_t)i];
        P.id = AR->ids[i];
        if constexpr (CPL == 8 && NW == 16)
            P.sbits = reinterpret_cast<const uint8_t*>(AR->starts)[(uint64_t)i * 64 + lane];
        else
            P.sbits = lane_start_bits<B, CPL>(AR->starts + (uint64_t)i * NW, lane);
        P.lc.load(table_(), AR->cases, i, lane, true);
    };
    auto stage2 = [&](TB& P) {
        const uint32_t bx = P.id % gx, by = (P.id / gx) % gy, bz = P.id / (gx * gy);
        const uint32_t nx = bx + (lane & 1u), ny = by + ((lane >> 1) & 1u), nz = bz + ((lane >> 2) & 1u);
        P.r = -1;
        if (lane != 0 && lane < 7 && nx < gx && ny < gy && nz < gz)
            P.r = (int32_t)AR->blk_map[nx + gx * (ny + (uint64_t)gy * nz)];
    };
    auto stage3 = [&](TB& P) {
        P.nbase = (lane < 8 && P.r >= 0) ? vb32_()[2 * (uint64_t)P.r] : 0u;
        const Entry* __restrict__ table = table_();
#pragma unroll
        for (int k = 0; k < TB::NFW; ++k) P.fentw[k] = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < NPL; ++k) {
            const uint32_t f = lane + 64u * k;
            const uint32_t fd = s_face[f < (uint32_t)NFACE ? f : 0u];
            const int32_t jn = __shfl(P.r, (int)((fd >> 13) & 7u), 64);
            if (f < (uint32_t)NFACE && jn >= 0) P.set_face(k, (uint32_t)table[(uint64_t)jn * NC + (fd >> 16)]);
        }
    };

    const WaveItems it = wave_items(na, WPG);
    TB cur, nxt;
    uint32_t i = it.first;
    if (i < it.end) {
        stage1(cur, i);
        stage2(cur);
        stage3(cur);
    }
    while (i < it.end) {
        const uint32_t inx = i + it.stride;
        const bool more = inx < it.end;
        if (more) stage1(nxt, inx);
        const uint32_t bnt = cur.bnt, bnm = cur.bnm, bns = cur.bns;
        const uint64_t tb = cur.tb;
        const uint32_t mb = cur.mb, msb = cur.msb;
        bool work = bnt != 0;
        if (work && (3 * (tb + bnt) > AR->icap || (uint64_t)msb + bns > AR->mvcap || (uint64_t)mb + bnm > AR->mcap ||
                     bnm > (uint32_t)C::MAXM)) {
            if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&AR->totals[2]), 1ull);
            work = false;
        }
        uint32_t vlim = 0, S0 = 0, slot = 0, nt = 0, npk = 0, cnt_v = 0, T = 0;
        const uint64_t sbits = cur.sbits;
        if (work) {
            // own corners: vertex index = block base + prefix from the count pass
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                if (c < (uint32_t)NC) {
                    const uint32_t q = c % B + F * ((c / B) % B + F * (c / (B * B)));
                    cor.set(q, cur.vb + (cur.lc.ent(j) >> 3), cur.lc.ent(j) & 7u);
                }
            }
            // neighbour-owned face corners
#pragma unroll
            for (int k = 0; k < NPL; ++k) {
                const uint32_t f = lane + 64u * k;
                const uint32_t fd = s_face[f < (uint32_t)NFACE ? f : 0u];
                const uint32_t nbv = (uint32_t)__shfl((int)cur.nbase, (int)((fd >> 13) & 7u), 64);
                if (f < (uint32_t)NFACE) {
                    const uint32_t q = fd & 0x1FFFu;
                    const uint32_t e = cur.face(k);
                    if (e != TriBlock<B>::kNone) cor.set(q, nbv + (e >> 3), e & 7u);
                    else cor.set_none(q);
                }
            }
            // each cell's added edges and distance back to its meshlet's start,
            // its triangle count
            block_cells(a, cur.id, vlim);
            S0 = wave_excl_max(sbits ? lane * CPL + 63u - (uint32_t)__builtin_clzll(sbits) : 0u);
            uint32_t S = S0;
            if constexpr (LISTED) {
                static_assert(CPL == 8 && NC == 512, "the list entry holds a 9-bit cell index");
                // every cell's meshlet-vertex word starts at 0 (slot 0, no
                // added edges); the slot pass rewrites the occupied cells
                *reinterpret_cast<uint4*>(s_mc + lane * 8u) = make_uint4(0u, 0u, 0u, 0u);
                *reinterpret_cast<uint4*>(s_mc + lane * 8u + 4u) = make_uint4(0u, 0u, 0u, 0u);
                uint32_t ent[CPL], occ = 0;
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    const uint32_t c = lane * CPL + j;
                    const uint32_t st = (uint32_t)(sbits >> j) & 1u;
                    if (st) S = c;
                    const uint32_t cs = cur.lc.cs(j);
                    const uint32_t n = s_ntri[cs];
                    ent[j] = c | (cs << 9) | ((c - S) << 17) | (n << 26) | (st << 29);
                    occ |= (n ? 1u : 0u) << j;
                    nt += n;
                }
                uint32_t q = wave_excl_scan<uint32_t>((uint32_t)__popc(occ), T);
#pragma unroll
                for (int j = 0; j < CPL; ++j)
                    if ((occ >> j) & 1u) s_ls[q++] = ent[j];
            } else {
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const uint32_t c = lane * CPL + j;
                if ((sbits >> j) & 1ull) S = c;
                const uint32_t cs = cur.lc.cs(j);
                const uint32_t n = s_ntri[cs];
                uint32_t m = 0;
                if (c < (uint32_t)NC) {
                    const uint32_t nb = nbrs_of_cell<B>(c, vlim);
                    if (n) m = added_edges<B>(s_emu[cs], nb, c - S, s_newtab, s_cls);
                    s_mc[c] = m << 20;   // the slot is added after the scan
                    if constexpr (!KIN) s_mk[c] = (uint16_t)(c - S);
                }
                cnt_v += __popc(m);
                nt += n;
                if (PACKN) npk |= n << (3 * j);
            }
            }
        }
        if (more) stage2(nxt);
        if (work) {
            uint32_t ns, tt, et = 0;
            if constexpr (LISTED) {
                wave_lds_sync();
                tt = wave_sum<uint32_t>(nt);
                // meshlet-vertex slots: the occupied cells' added edges, in cell order
                uint32_t run = 0;
                for (uint32_t e0 = 0; e0 < T; e0 += 64u) {
                    const uint32_t e = e0 + lane;
                    const uint32_t x = e < T ? s_ls[e] : 0u;
                    const uint32_t c = x & 511u;
                    uint32_t m = 0;
                    if (e < T)
                        m = added_edges<B>(s_emu[(x >> 9) & 255u], nbrs_of_cell<B>(c, vlim), (x >> 17) & 511u,
                                           s_newtab, s_cls);
                    uint32_t tot;
                    const uint32_t sl = run + wave_excl_scan<uint32_t>((uint32_t)__popc(m), tot);
                    if (e < T) s_mc[c] = sl | (m << 20);
                    run += tot;
                }
                ns = run;
            } else {
                slot = wave_excl_scan<uint32_t>(cnt_v, ns);
                et = wave_excl_scan<uint32_t>(nt, tt);
            }
            const uint32_t nm = wave_sum<uint32_t>((uint32_t)__popcll(sbits));
            if (ns != bns || nm != bnm || tt != bnt) {
                // the cut and the counts disagree with this block's data
                if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&AR->totals[2]), 4ull);
            } else {
                if constexpr (!LISTED) {
#pragma unroll
                    for (int j = 0; j < CPL; ++j) {
                        const uint32_t c = lane * CPL + j;
                        if (c < (uint32_t)NC) {
                            const uint32_t m = s_mc[c];
                            s_mc[c] = m | slot;
                            slot += __popc(m >> 20);
                        }
                    }
                }
                uint32_t bad = 0;
                for (uint32_t c0 = 0; c0 < tt; c0 += LIST) {
                    const uint32_t n = min(LIST, tt - c0);
                    if constexpr (LISTED) {
                        // list entries in output order: cell by cell, the case's triangles k = 0..4
                        wave_lds_sync();
                        uint32_t trun = 0;
                        for (uint32_t e0 = 0; e0 < T; e0 += 64u) {
                            const uint32_t e = e0 + lane;
                            const uint32_t x = e < T ? s_ls[e] : 0u;
                            const uint32_t nn = (x >> 26) & 7u;
                            uint32_t tot;
                            const uint32_t r = trun + wave_excl_scan<uint32_t>(nn, tot) - c0;
                            const uint32_t ent =
                                (x & 511u) | ((((x >> 9) & 255u) * 5u) << CB) | (((x >> 17) & 511u) << (CB + 11));
#pragma unroll
                            for (uint32_t k = 0; k < 5u; ++k)
                                if (k < nn && r + k < LIST) s_tl[r + k] = ent + (k << CB);
                            trun += tot;
                        }
                    } else {
                    // list entries in output order: cell by cell, the case's triangles k = 0..4
                    uint32_t run = et - c0;
                    uint32_t S = S0;
                    // re-read per pass (opaque to the compiler): nothing derived
                    // from them stays live across passes, which are rare
                    LaneCorners<B> lc = cur.lc;
#pragma unroll
                    for (int k = 0; k < LaneCorners<B>::NCW; ++k) asm volatile("" : "+v"(lc.csw[k]));
#pragma unroll
                    for (int j = 0; j < CPL; ++j) {
                        const uint32_t c = lane * CPL + j;
                        if ((sbits >> j) & 1ull) S = c;
                        if (c < (uint32_t)NC) {
                            const uint32_t cs = lc.cs(j);
                            const uint32_t nn = PACKN ? (npk >> (3 * j)) & 7u : (uint32_t)s_ntri[cs];
                            uint32_t ent = c | ((cs * 5u) << CB);
                            if constexpr (KIN) ent |= (c - S) << (CB + 11);
#pragma unroll
                            for (uint32_t k = 0; k < 5u; ++k)
                                if (k < nn && run + k < LIST) s_tl[run + k] = ent + (k << CB);
                            run += nn;
                        }
                    }
                    }
                    wave_lds_sync();
                    // one triangle per lane, its three vertices' lookups side by side
                    for (uint32_t q = lane; q < n; q += 64u) {
                        const uint32_t e = s_tl[q];
                        const uint32_t c = e & ((1u << CB) - 1u);
                        const uint32_t dec = (e >> CB) & 2047u;
                        const uint64_t du = s_tdu[dec];
                        const TDec d = (TDec)(uint32_t)du;
                        const uint32_t tu = (uint32_t)(du >> 32);
                        const uint32_t nbc = nbrs_of_cell<B>(c, vlim);
                        const uint32_t kk = KIN ? e >> (CB + 11) : (uint32_t)s_mk[c];
                        // the distance class of kk (FU8)
                        const uint32_t fcls = kk >= (uint32_t)(B * B + B) ? kBackClasses - 1u : (uint32_t)s_cls[kk];
                        const uint32_t q0 = c % B + F * ((c / B) % B + F * (c / (B * B)));
                        uint32_t gi[3], mv[3];
#pragma unroll
                        for (uint32_t e3 = 0; e3 < 3; ++e3) {
                            // shared mesh: the owner corner's first vertex plus the edge's rank there
                            const uint32_t f = (uint32_t)(d >> (DB * e3));
                            const uint32_t qq = q0 + (f & ((1u << (DB - 2)) - 1u));
                            gi[e3] = cor.vertex(qq, f >> (DB - 2), bad);
                            // meshlet: the edge's first user in the meshlet and its rank there
                            const uint32_t u = (tu >> (4 * e3)) & 15u;
                            uint32_t back, u1;
                            if constexpr (FU8) {
                                const uint32_t fe = s_fu8[(nbc * 12u + u) * kBackClasses + fcls];
                                back = s_bd[fe & 15u];
                                u1 = fe >> 4;
                            } else {
                                first_user_of(s_fu[nbc * 12u + u], u, kk, back, u1);
                            }
                            const uint32_t mc = s_mc[c - back];
                            mv[e3] = msb + (mc & 0xFFFFFu) + __popc((mc >> 20) & ((1u << u1) - 1u));
                        }
#pragma unroll
                        for (uint32_t e3 = 0; e3 < 3; ++e3) AR->mvmap[mv[e3]] = gi[e3];
                        // a triangle's three entries as one 12-byte store per
                        // lane (the wave writes 768 contiguous bytes; three
                        // strided 4-byte stores rewrote each line three times)
                        store3_nt(AR->indices + 3 * (tb + c0 + q), gi[0], gi[1], gi[2]);
                        store3_nt(AR->mindices + 3 * (tb + c0 + q), mv[0], mv[1], mv[2]);
                    }
                    wave_lds_sync();
                }
                // an active edge whose owner block is not in the active list: the
                // list does not belong to this volume/isovalue
                if (bad & 0xF0u) atomicOr(reinterpret_cast<unsigned long long*>(&AR->totals[2]), 2ull);
                // descriptors: each meshlet's first meshlet vertex and first
                // triangle within the block (16 bits each), staged at its index
                if constexpr (LISTED) {
                    // cell 0 starts the first meshlet (offsets 0 and 0 when it is
                    // empty); every other start is an occupied cell
                    uint32_t srun = 0, trun = 0;
                    if ((s_ls[0] & 511u) != 0u) {
                        if (lane == 0) s_tl[0] = 0u;
                        srun = 1;
                    }
                    for (uint32_t e0 = 0; e0 < T; e0 += 64u) {
                        const uint32_t e = e0 + lane;
                        const uint32_t x = e < T ? s_ls[e] : 0u;
                        const uint32_t sf = (x >> 29) & 1u;
                        uint32_t tot, stot;
                        const uint32_t tp = trun + wave_excl_scan<uint32_t>((x >> 26) & 7u, tot);
                        const uint32_t mi = srun + wave_excl_scan<uint32_t>(sf, stot);
                        if (sf && mi < (uint32_t)C::TL) s_tl[mi] = (s_mc[x & 511u] & 0xFFFFu) | (tp << 16);
                        trun += tot;
                        srun += stot;
                    }
                } else {
                uint32_t nmt;
                uint32_t m = wave_excl_scan<uint32_t>((uint32_t)__popcll(sbits), nmt);
                uint32_t tp = et;
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    const uint32_t c = lane * CPL + j;
                    if ((sbits >> j) & 1ull) s_tl[m++] = (s_mc[c] & 0xFFFFu) | (tp << 16);
                    tp += PACKN ? (npk >> (3 * j)) & 7u : (uint32_t)s_ntri[cur.lc.cs(j)];
                }
                }
                wave_lds_sync();
                const uint32_t tb_rel = (uint32_t)tb;
                for (uint32_t k = lane; k < nm; k += 64u) {
                    const uint32_t A = s_tl[k];
                    const uint32_t Bn = k + 1 < nm ? s_tl[k + 1] : (bns | (bnt << 16));
                    AR->meshlets[(uint64_t)mb + k] = make_uint4(msb + (A & 0xFFFFu), 3u * (tb_rel + (A >> 16)),
                                                              (Bn & 0xFFFFu) - (A & 0xFFFFu), (Bn >> 16) - (A >> 16));
                }
            }
            wave_lds_sync();
        }
        if (more) stage3(nxt);
        cur = nxt;
        i = inx;
    }
}

// Meshlet vertices: copies of the shared mesh's vertices through the meshlet
// vertex map, two lanes per vertex (one 16-byte half each), so that a wave
// reads its map entries and writes its copies as contiguous 1 KiB runs.  A map
// entry past the vertex buffer (a block flagged over capacity) gives zeros.
__global__ __launch_bounds__(256) void k_meshlet_vertices(const float4* __restrict__ vertices,
                                                          const uint32_t* __restrict__ map,
                                                          float4* __restrict__ mverts,
                                                          const uint64_t* __restrict__ totals, uint64_t mvcap,
                                                          uint64_t vcap)
{
    const uint64_t n = 2 * min(totals[4], mvcap);
    for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (uint64_t)gridDim.x * 256) {
        const uint64_t j = t >> 1;
        const uint32_t v = map[j];
        float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
        if (v < vcap) x = vertices[2 * (uint64_t)v + (t & 1)];
        __builtin_nontemporal_store(mtx_f4v{x.x, x.y, x.z, x.w}, reinterpret_cast<mtx_f4v*>(mverts + t));
    }
}

// Persistent grids of the extraction kernels, in workgroups per CU (a
// multiple of 8 workgroups for the XCD split, never more waves than blocks).
// Measured on MI355X: a single volume is fastest with whole rounds of the
// resident workgroups -- one round for the triangle passes (each workgroup
// sets up LDS tables; same-box sweep of the shared-mesh pass at 1024^3: one
// round of 6 per CU 8 us faster than 10 per CU), 64 for the vertex
// pass (5 resident at B = 8: with many more workgroups than fit, the
// dispatcher hands them to CUs as others finish, which balances the passes'
// uneven blocks and fills the CUs the triangle + meshlet pass frees; same-box
// sweep at 1024^3: 10 -> 64 per CU took 28 us off both extractions), for
// the B = 8 count pass 24, or on grids of 2^20
// blocks and more 96 (48 with the meshlet cut: the extra rounds even out the
// blocks' load times; measured on one GPU box against 24 at 1024^3: -16 and
// -2.5 us), 8 for the other block sizes' count pass; fewer
// leave CUs idle while long blocks finish, more add per-workgroup set-up.  A
// batch, whose blocks differ far more in work (dense and sparse surfaces of
// many units), is fastest with finer static shares, 32-128 per CU.
constexpr uint32_t kWgPerCuSingle = 8;
constexpr uint32_t kWgPerCuSingleVertices = 64;
constexpr uint32_t kWgPerCuSingleCount = 24;
constexpr uint32_t kWgPerCuLargeCount = 96;
constexpr uint32_t kWgPerCuLargeCountCut = 48;
constexpr uint64_t kLargeGridBlocks = 1ull << 20;
constexpr uint32_t kWgPerCuBatch = 64;
constexpr uint32_t kWgPerCuMeshlet = 8;

// Workgroups of `kernel` one CU holds, queried once per kernel (the first
// launch of a batch is its planning run, outside graph capture).
static int resident_wgs(const void* kernel, uint32_t threads)
{
    static std::mutex mu;
    static std::vector<std::pair<const void*, int>> known;
    std::lock_guard<std::mutex> lk(mu);
    for (const auto& k : known)
        if (k.first == kernel) return k.second;
    int r = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, kernel, (int)threads, 0) != hipSuccess) r = 0;
    known.emplace_back(kernel, r);
    return r;
}

// per_cu_single = 0: one round of the workgroups a CU holds.
template <typename K>
static hipError_t launch_persistent(K kernel, uint32_t threads, const ExtractArgs& a, hipStream_t s,
                                    uint32_t per_cu_single = kWgPerCuSingle)
{
    if (per_cu_single == 0) {
        const int r = resident_wgs(reinterpret_cast<const void*>(kernel), threads);
        per_cu_single = r > 0 ? (uint32_t)r : kWgPerCuSingle;
    }
    uint64_t n = (uint64_t)device_cus() * (a.wg_per_cu ? a.wg_per_cu : (a.units ? kWgPerCuBatch : per_cu_single));
    if (a.balanced) {
        // equal-cost ranges assume every wave runs at once: exactly one round
        // of the workgroups the CUs hold
        const int r = resident_wgs(reinterpret_cast<const void*>(kernel), threads);
        if (r > 0) n = (uint64_t)device_cus() * (uint64_t)r;
    }
    const uint64_t need = ((uint64_t)a.max_blocks + threads / 64 - 1) / (threads / 64);
    if (need < n) n = need;
    n = n < 8 ? 8 : n / 8 * 8;
    hipLaunchKernelGGL(kernel, dim3((uint32_t)n), dim3(threads), 0, s, a);
    return hipGetLastError();
}

template <int B, bool BATCH>
static hipError_t launch_count_b(const ExtractArgs& a, hipStream_t s)
{
    if (B == 8) {
        const bool large = a.max_blocks >= kLargeGridBlocks;
        if (!BATCH && a.starts) {   // with the meshlet cut
            const uint32_t g = large ? kWgPerCuLargeCountCut : kWgPerCuSingleCount;
            if (a.vol.dtype == kF32) return launch_persistent(k_extract_count_b8<float, false, true>, 256, a, s, g);
            return launch_persistent(k_extract_count_b8<uint8_t, false, true>, 256, a, s, g);
        }
        const uint32_t g = large ? kWgPerCuLargeCount : kWgPerCuSingleCount;
        if (a.vol.dtype == kF32) return launch_persistent(k_extract_count_b8<float, BATCH, false>, 256, a, s, g);
        return launch_persistent(k_extract_count_b8<uint8_t, BATCH, false>, 256, a, s, g);
    }
    if (a.vol.dtype == kF32) return launch_persistent(k_extract_count<B, float, BATCH>, 256, a, s);
    return launch_persistent(k_extract_count<B, uint8_t, BATCH>, 256, a, s);
}

template <int B, bool BATCH>
static hipError_t launch_vertices_b(const ExtractArgs& a, hipStream_t s)
{
    if (a.vol.dtype == kF32) {
        return a.normals ? launch_persistent(k_extract_vertices<B, float, true, BATCH>,
                                             64 * VertCfg<B, true, float>::WPG, a, s, kWgPerCuSingleVertices)
                         : launch_persistent(k_extract_vertices<B, float, false, BATCH>,
                                             64 * VertCfg<B, false, float>::WPG, a, s, kWgPerCuSingleVertices);
    }
    return a.normals ? launch_persistent(k_extract_vertices<B, uint8_t, true, BATCH>,
                                         64 * VertCfg<B, true, uint8_t>::WPG, a, s, kWgPerCuSingleVertices)
                     : launch_persistent(k_extract_vertices<B, uint8_t, false, BATCH>,
                                         64 * VertCfg<B, false, uint8_t>::WPG, a, s, kWgPerCuSingleVertices);
}

// Vertex pass, then triangles (and meshlets).  The two depend only on the
// count pass and the scan, so with a side stream the vertex pass (bound by
// its tile fetch) runs there beside the triangle pass (bound by its vector
// ALU) and `s` joins it before anything that reads the vertices.  They share
// the CUs only as far as LDS allows (three triangle + meshlet workgroups
// fill a CU's), so the overlap is partial: 16 us off the extraction at
// 1024^3 (A/B on one GPU box).
template <int B, bool BATCH>
static hipError_t launch_emit_b(const ExtractArgs& a, hipStream_t s)
{
    hipError_t e = hipSuccess;
    const bool side = !BATCH && a.side;
    if (side) {
        e = hipEventRecord(a.side_ev[0], s);
        if (e == hipSuccess) e = hipStreamWaitEvent(a.side, a.side_ev[0], 0);
        if (e == hipSuccess) e = launch_vertices_b<B, BATCH>(a, a.side);
        if (e == hipSuccess) e = hipEventRecord(a.side_ev[1], a.side);
    } else {
        e = launch_vertices_b<B, BATCH>(a, s);
    }
    if (e != hipSuccess) return e;
    bool fused = false;
    if constexpr (!BATCH && TriMeshCfg<B>::FITS) {
        if (a.starts) {
            // triangles and meshlets in one pass; one round of resident
            // workgroups: each sets up about 20 KB of LDS tables, so fewer,
            // longer-lived workgroups are faster
            e = a.vcap < kPackedIndexLimit
                    ? launch_persistent(k_tri_meshlets<B, true>, 64 * TriMeshCfg<B>::WPG, a, s, 0)
                    : launch_persistent(k_tri_meshlets<B, false>, 64 * TriMeshCfg<B>::WPG, a, s, 0);
            fused = true;
        }
    }
    if (!fused) {
        e = a.vcap < kPackedIndexLimit
                ? launch_persistent(k_extract_triangles<B, BATCH, true>, 64 * TriCfg<B>::WPG, a, s, 0)
                : launch_persistent(k_extract_triangles<B, BATCH, false>, 64 * TriCfg<B>::WPG, a, s, 0);
        // meshlets (not in the batched form)
        if (e == hipSuccess && !BATCH && a.starts)
            e = launch_persistent(k_meshlet_emit<B>, 64 * EmitCfg<B>::WPG, a, s, kWgPerCuMeshlet);
    }
    if (e == hipSuccess && side) e = hipStreamWaitEvent(s, a.side_ev[1], 0);
    if (e != hipSuccess || BATCH || !a.starts) return e;
    hipLaunchKernelGGL(k_meshlet_vertices, dim3((uint32_t)device_cus() * 8), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(a.vertices), a.mvmap, reinterpret_cast<float4*>(a.mverts),
                       a.totals, a.mvcap, a.vcap);
    return hipGetLastError();
}

template <int B, bool BATCH>
static hipError_t launch_cut_b(const ExtractArgs& a, hipStream_t s)
{
    return launch_persistent(k_meshlet_cut<B>, 256, a, s, kWgPerCuMeshlet);
}

#define MTX_DISPATCH_B(FN, ...)                                   \
    switch (a.grid.B) {                                           \
    case 1: return FN<1, false>(__VA_ARGS__);                     \
    case 2: return FN<2, false>(__VA_ARGS__);                     \
    case 3: return FN<3, false>(__VA_ARGS__);                     \
    case 4: return FN<4, false>(__VA_ARGS__);                     \
    case 5: return FN<5, false>(__VA_ARGS__);                     \
    case 6: return FN<6, false>(__VA_ARGS__);                     \
    case 7: return FN<7, false>(__VA_ARGS__);                     \
    case 8: return FN<8, false>(__VA_ARGS__);                     \
    case 9: return FN<9, false>(__VA_ARGS__);                     \
    case 10: return FN<10, false>(__VA_ARGS__);                   \
    case 11: return FN<11, false>(__VA_ARGS__);                   \
    case 12: return FN<12, false>(__VA_ARGS__);                   \
    case 13: return FN<13, false>(__VA_ARGS__);                   \
    case 14: return FN<14, false>(__VA_ARGS__);                   \
    case 15: return FN<15, false>(__VA_ARGS__);                   \
    case 16: return FN<16, false>(__VA_ARGS__);                   \
    default: return hipErrorInvalidValue;                         \
    }

// The batched form exists for the block edges a batch can be tuned to.
#define MTX_DISPATCH_BATCH_B(FN, ...)                             \
    switch (a.grid.B) {                                           \
    case 4: return FN<4, true>(__VA_ARGS__);                      \
    case 8: return FN<8, true>(__VA_ARGS__);                      \
    case 16: return FN<16, true>(__VA_ARGS__);                    \
    default: return hipErrorInvalidValue;                         \
    }

bool batch_supported_block(uint32_t B) { return B == 4 || B == 8 || B == 16; }

static hipError_t launch_count_single(const ExtractArgs& a, hipStream_t s)
{
    MTX_DISPATCH_B(launch_count_b, a, s)
}

static hipError_t launch_cut(const ExtractArgs& a, hipStream_t s)
{
    MTX_DISPATCH_B(launch_cut_b, a, s)
}

hipError_t launch_extract_count(const ExtractArgs& a, hipStream_t s)
{
    if (a.units) { MTX_DISPATCH_BATCH_B(launch_count_b, a, s) }
    const hipError_t e = launch_count_single(a, s);
    if (e != hipSuccess || !a.starts || a.grid.B == 8) return e;   // B = 8: the count pass cuts
    return launch_cut(a, s);   // meshlet counts and start bits, from the cases
}

hipError_t launch_extract_emit(const ExtractArgs& a, hipStream_t s)
{
    if (a.units) { MTX_DISPATCH_BATCH_B(launch_emit_b, a, s) }
    MTX_DISPATCH_B(launch_emit_b, a, s)
}

hipError_t launch_extract_scan(const ExtractArgs& a, TileSums ts, hipStream_t s)
{
    uint64_t ntiles = ((uint64_t)a.max_blocks + 256 * kScanItems - 1) / (256 * kScanItems);
    if (ntiles == 0) ntiles = 1;
    if (ntiles > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_extract_tile_sums, dim3((uint32_t)ntiles), dim3(256), 0, s, a.blk_counts, a.count,
                       a.max_blocks, ts.sums, a.totals);
    hipLaunchKernelGGL(k_extract_scan, dim3((uint32_t)ntiles), dim3(256), 0, s, a.blk_counts, a.count,
                       a.max_blocks, a.blk_vbase, a.blk_tbase, a.blk_mbase, a.blk_sbase, a.totals, a.out_counts,
                       ts.sums);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// adjoint of the vertex positions
//
// p = owner + t * e_axis (voxel units), t = (iso - v0) / (v1 - v0) with the
// topology of the forward pass frozen:
//   dt/dv0 = (t - 1) / (v1 - v0),  dt/dv1 = -t / (v1 - v0),  dt/diso = 1 / (v1 - v0).
// Every voxel belongs to one home block (the block holding it as a corner).
// A voxel's gradient gathers the active edges it ends: its own +x/+y/+z edges
// (it is v0) and the edges of its -x/-y/-z neighbour voxels (it is v1), so only
// the home blocks of active blocks and of their +x/+y/+z neighbours can be
// nonzero.  Three launches:
//   k_adjoint_zero  dL/dvolume = 0 everywhere, 16-byte stores;
//   k_adjoint       a wave per active block gathers the gradient of every voxel
//                   of the home blocks it answers for (its own, and each
//                   inactive +x/+y/+z neighbour not claimed by an active block
//                   nearer in the order -x, -y, -z), looking the vertices up in
//                   the forward pass's corner tables and recomputing t from the
//                   volume; each such voxel is written once, lanes on
//                   consecutive x (no atomics, a deterministic result);
//   k_adjoint_iso   the per-workgroup dL/diso partials added in a fixed order.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_adjoint_zero(float* __restrict__ out, uint64_t n)
{
    const uint64_t i0 = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    if (reinterpret_cast<uintptr_t>(out) % 16u == 0) {
        float4* __restrict__ o4 = reinterpret_cast<float4*>(out);
        const uint64_t n4 = n / 4;
        for (uint64_t i = i0; i < n4; i += stride) o4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (uint64_t i = 4 * n4 + i0; i < n; i += stride) out[i] = 0.0f;
    } else {
        for (uint64_t i = i0; i < n; i += stride) out[i] = 0.0f;
    }
}

// The gradient of every voxel of home block b (its corner tables and those of
// its -x, -y, -z neighbours give the vertices; the volume gives t).  Loads are
// issued in three rounds so that each round's latency is paid once: the tile
// (direct to LDS for f32), the corner-table entries and the four vertex bases;
// then every dL/dp the chunk's cells need; then the arithmetic.
template <int B, typename T>
__device__ __forceinline__ void adjoint_home(const AdjointArgs& a, float* s_val, uint32_t b, double& diso)
{
    constexpr int E = B + 2;  // voxels [-1, B] per axis
    constexpr int NTILE = E * E * E;
    constexpr int KT = (NTILE + 63) / 64;
    constexpr int NC = EB<B>::NC;
    constexpr int CPL = EB<B>::CPL;
    constexpr int KC = CPL < 4 ? CPL : 4;  // cells per lane in one round of loads
    using Entry = typename EB<B>::Entry;
    const uint32_t lane = lane_id();
    const ExtractArgs& f = a.fwd;
    const T* __restrict__ vol = static_cast<const T*>(f.vol.data);
    const uint32_t X = f.vol.X, Y = f.vol.Y, Z = f.vol.Z;
    const float iso = f.iso;
    const Entry* __restrict__ table = static_cast<const Entry*>(f.corner_table);
    const uint32_t gx = f.grid.gx, gy = f.grid.gy;
    const uint32_t bx = b % gx, by = (b / gx) % gy, bz = b / (gx * gy);
    const uint32_t ox = bx * B, oy = by * B, oz = bz * B;
    // this block and its -x, -y, -z neighbours: active index (-1: inactive), vertex base
    int32_t nb[4];
    nb[0] = (int32_t)f.blk_map[b];
    nb[1] = bx > 0 ? (int32_t)f.blk_map[b - 1] : -1;
    nb[2] = by > 0 ? (int32_t)f.blk_map[b - gx] : -1;
    nb[3] = bz > 0 ? (int32_t)f.blk_map[b - gx * gy] : -1;
    uint64_t vb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) vb[q] = nb[q] >= 0 ? f.blk_vbase[nb[q]] : 0;
    // the tile [-1, B]^3, clamped to the volume: f32 straight into LDS, u8
    // through registers 8 loads at a time
    auto tile_src = [&](uint32_t j) {
        const int32_t x = min(max((int32_t)(ox + j % E) - 1, 0), (int32_t)X - 1);
        const int32_t y = min(max((int32_t)(oy + (j / E) % E) - 1, 0), (int32_t)Y - 1);
        const int32_t z = min(max((int32_t)(oz + j / (E * E)) - 1, 0), (int32_t)Z - 1);
        return vol + ((uint32_t)x + (uint64_t)X * ((uint32_t)y + (uint64_t)Y * (uint32_t)z));
    };
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const uint32_t j = 64u * (uint32_t)k + lane;
            if (j < (uint32_t)NTILE) lds_direct_load4(tile_src(j), s_val + 64 * k);
        }
    } else {
#pragma unroll 1
        for (int k0 = 0; k0 < KT; k0 += 8) {
            float tv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t j = 64u * (uint32_t)(k0 + k) + lane;
                tv[k] = j < (uint32_t)NTILE ? vfloat(*tile_src(j)) : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t j = 64u * (uint32_t)(k0 + k) + lane;
                if (j < (uint32_t)NTILE) s_val[j] = tv[k];
            }
        }
    }
    bool tile_waited = false;
#pragma unroll 1
    for (int k0 = 0; k0 < CPL; k0 += KC) {
        // round 2: corner-table entries -- the cell's own (it is v0 of its
        // edges) and those of its -x, -y, -z neighbour voxels' owners (it is v1)
        uint32_t eo[KC], en[KC][3];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const uint32_t c = lane + 64u * (uint32_t)(k0 + k);
            const uint32_t lx = c % B, ly = (c / B) % B, lz = c / (B * B);
            const bool ok = c < (uint32_t)NC && ox + lx < X && oy + ly < Y && oz + lz < Z;
            eo[k] = ok && nb[0] >= 0 ? (uint32_t)table[(uint64_t)nb[0] * NC + c] : 0u;
#pragma unroll
            for (uint32_t ax = 0; ax < 3; ++ax) {
                const uint32_t l = ax == 0 ? lx : (ax == 1 ? ly : lz);
                const uint32_t gcoord = ax == 0 ? ox + lx : (ax == 1 ? oy + ly : oz + lz);
                const int32_t n = l == 0 ? nb[1 + ax] : nb[0];
                const uint32_t oc = l == 0 ? (ax == 0 ? (B - 1) + B * (ly + B * lz)
                                                      : (ax == 1 ? lx + B * ((B - 1) + B * lz)
                                                                 : lx + B * (ly + B * (B - 1))))
                                           : c - (ax == 0 ? 1u : (ax == 1 ? (uint32_t)B : (uint32_t)(B * B)));
                en[k][ax] = ok && gcoord > 0 && n >= 0 ? (uint32_t)table[(uint64_t)n * NC + oc] : 0u;
            }
        }
        // round 3: the dL/dp of every edge found
        float go[KC][3], gn[KC][3];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const uint32_t c = lane + 64u * (uint32_t)(k0 + k);
            const uint32_t lx = c % B, ly = (c / B) % B, lz = c / (B * B);
#pragma unroll
            for (uint32_t ax = 0; ax < 3; ++ax) {
                const uint32_t e = eo[k];
                go[k][ax] = (e >> ax) & 1u
                                ? a.dl_dpos[3 * (vb[0] + (e >> 3) + __popc(e & 7u & ((1u << ax) - 1u))) + ax]
                                : 0.0f;
                const uint32_t l = ax == 0 ? lx : (ax == 1 ? ly : lz);
                const uint32_t en_ = en[k][ax];
                const uint64_t vbn = l == 0 ? vb[1 + ax] : vb[0];
                gn[k][ax] = (en_ >> ax) & 1u
                                ? a.dl_dpos[3 * (vbn + (en_ >> 3) + __popc(en_ & 7u & ((1u << ax) - 1u))) + ax]
                                : 0.0f;
            }
        }
        if (!tile_waited) {
            if constexpr (sizeof(T) == 4) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the tile is in LDS
            wave_lds_sync();
            tile_waited = true;
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const uint32_t c = lane + 64u * (uint32_t)(k0 + k);   // consecutive lanes: consecutive x
            if (c >= (uint32_t)NC) break;
            const uint32_t lx = c % B, ly = (c / B) % B, lz = c / (B * B);
            const uint32_t x = ox + lx, y = oy + ly, z = oz + lz;
            if (!(x < X && y < Y && z < Z)) continue;
            const uint32_t t0 = (lx + 1) + E * ((ly + 1) + E * (lz + 1));
            const float vg = s_val[t0];
            float g = 0.0f;
#pragma unroll
            for (uint32_t ax = 0; ax < 3; ++ax) {
                const uint32_t st = ax == 0 ? 1u : (ax == 1 ? (uint32_t)E : (uint32_t)(E * E));
                if ((eo[k] >> ax) & 1u) {
                    const float d = s_val[t0 + st] - vg;
                    if (fabsf(d) > 1e-5f) {  // else t = 0.5: no dependence on the values
                        const float G = go[k][ax];
                        const float t = (iso - vg) / d;
                        g += G * (t - 1.0f) / d;
                        diso += (double)(G / d);
                    }
                }
                if ((en[k][ax] >> ax) & 1u) {
                    const float v0 = s_val[t0 - st];
                    const float d = vg - v0;
                    if (fabsf(d) > 1e-5f) {
                        const float t = (iso - v0) / d;
                        g += -gn[k][ax] * t / d;
                    }
                }
            }
            a.dl_dvol[x + (uint64_t)X * (y + (uint64_t)Y * z)] = g;
        }
    }
    wave_lds_sync();
}

template <int B, typename T>
__global__ __launch_bounds__(256) void k_adjoint(AdjointArgs a)
{
    constexpr int E = B + 2;
    __shared__ float s_all[4][E * E * E];
    __shared__ double s_part[4];
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    float* s_val = s_all[w];
    const ExtractArgs& f = a.fwd;
    const uint32_t gx = f.grid.gx, gy = f.grid.gy, gz = f.grid.gz;
    const uint32_t* __restrict__ map = f.blk_map;
    const uint32_t na = active_items(f);
    // one item per (active block, home block): its own, +x, +y, +z; four
    // blocks in a row when 4 * na would not fit 32 bits
    const bool split = na <= 0x3FFFFFFFu;
    const WaveItems it = wave_items(split ? 4u * na : na, 4);
    const uint32_t pz = gx * gy;
    double diso = 0.0;
    for (uint32_t q = it.first; q < it.end; q += it.stride) {
        const uint32_t A = f.ids[split ? q >> 2 : q];
        const uint32_t bx = A % gx, by = (A / gx) % gy, bz = A / (gx * gy);
        auto off = [&](uint32_t b) { return (int32_t)map[b] < 0; };
        const uint32_t k0 = split ? (q & 3u) : 0u, k1 = split ? k0 + 1u : 4u;
        for (uint32_t k = k0; k < k1; ++k) {
            // an inactive neighbour's home block goes to the first active
            // block of its -x, -y, -z neighbours, in that order
            uint32_t b = A;
            bool mine = true;
            if (k == 1) {
                b = A + 1;
                mine = bx + 1 < gx && off(b);
            } else if (k == 2) {
                b = A + gx;
                mine = by + 1 < gy && off(b) && (bx == 0 || off(b - 1));
            } else if (k == 3) {
                b = A + pz;
                mine = bz + 1 < gz && off(b) && (bx == 0 || off(b - 1)) && (by == 0 || off(b - gx));
            }
            if (mine) adjoint_home<B, T>(a, s_val, b, diso);
        }
    }
    // fixed order: the wave's shuffle tree, then the workgroup's waves 0..3
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) diso += __shfl_xor(diso, d, 64);
    if (lane == 0) s_part[w] = diso;
    __syncthreads();
    if (threadIdx.x == 0) a.iso_partials[blockIdx.x] = ((s_part[0] + s_part[1]) + s_part[2]) + s_part[3];
}

__global__ __launch_bounds__(256) void k_adjoint_iso(const double* __restrict__ partials, uint32_t n,
                                                     float* __restrict__ out)
{
    // thread-strided sums, then a fixed tree: the same order on every run
    __shared__ double s[256];
    double v = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256u) v += partials[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)s[0];
}

template <int B>
static hipError_t launch_adjoint_b(const AdjointArgs& a, hipStream_t s)
{
    const uint64_t n = (uint64_t)a.fwd.vol.X * a.fwd.vol.Y * a.fwd.vol.Z;
    uint64_t zg = (n / 4 + 255) / 256;
    zg = zg < 1 ? 1 : (zg > 8192 ? 8192 : zg);
    hipLaunchKernelGGL(k_adjoint_zero, dim3((uint32_t)zg), dim3(256), 0, s, a.dl_dvol, n);
    if (a.fwd.vol.dtype == kF32)
        hipLaunchKernelGGL((k_adjoint<B, float>), dim3(a.n_wg), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_adjoint<B, uint8_t>), dim3(a.n_wg), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adjoint_iso, dim3(1), dim3(256), 0, s, a.iso_partials, a.n_wg, a.dl_diso);
    return hipGetLastError();
}

uint32_t adjoint_workgroups(uint64_t nb)
{
    uint64_t need = (nb + 3) / 4;
    uint32_t n = 2048;
    if (need < n) n = need < 8 ? 8 : (uint32_t)(((need + 7) / 8) * 8);
    return n;
}

hipError_t launch_adjoint(const AdjointArgs& a, hipStream_t s)
{
    switch (a.fwd.grid.B) {
    case 1: return launch_adjoint_b<1>(a, s);
    case 2: return launch_adjoint_b<2>(a, s);
    case 3: return launch_adjoint_b<3>(a, s);
    case 4: return launch_adjoint_b<4>(a, s);
    case 5: return launch_adjoint_b<5>(a, s);
    case 6: return launch_adjoint_b<6>(a, s);
    case 7: return launch_adjoint_b<7>(a, s);
    case 8: return launch_adjoint_b<8>(a, s);
    case 9: return launch_adjoint_b<9>(a, s);
    case 10: return launch_adjoint_b<10>(a, s);
    case 11: return launch_adjoint_b<11>(a, s);
    case 12: return launch_adjoint_b<12>(a, s);
    case 13: return launch_adjoint_b<13>(a, s);
    case 14: return launch_adjoint_b<14>(a, s);
    case 15: return launch_adjoint_b<15>(a, s);
    case 16: return launch_adjoint_b<16>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mtx
