// gfx950 kernels of the batched form: ordered compaction over many
// (volume, isovalue) units in one pass, and the per-unit output ranges.
#include "../../include/meshtrex_amd/mtx.h"
#include "mtx_device.h"

namespace mtx {

constexpr uint32_t kBatchFilterItems = 8;                           // blocks per thread
constexpr uint32_t kBatchFilterTile = 256 * kBatchFilterItems;      // blocks per tile

uint32_t batch_filter_tiles(uint32_t nb) { return (nb + kBatchFilterTile - 1) / kBatchFilterTile; }

// One decoupled look-back scan over the tiles of every unit of a class, in
// unit order, so that the concatenated active list is ordered by (unit, block
// ID).  Each unit's blocks start on a tile boundary; a tile finds its unit with
// one wave ballot per 64 units.  Activity follows the unit's own isovalue and
// rule (src/blockFilteringTestUtils.h:456-468 or
// src/shaders/occupiedBlockPrefixSum.comp.glsl:84).
__global__ __launch_bounds__(256) void k_filter_batch(const BatchUnit* __restrict__ units,
                                                      const uint32_t* __restrict__ tile0, uint32_t nunits,
                                                      uint32_t ntiles, uint32_t* __restrict__ ids,
                                                      uint32_t* __restrict__ item_unit,
                                                      uint32_t* __restrict__ unit_first,
                                                      uint32_t* __restrict__ count, uint64_t* status,
                                                      uint32_t* ticket)
{
    __shared__ uint32_t s_tile, s_unit;
    __shared__ uint32_t s_cnt[kBatchFilterItems * 4];
    __shared__ uint64_t s_excl;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (w == 0) {
        uint32_t n = 0;
        for (uint32_t c = 0; c < nunits; c += 64) {
            const uint32_t k = c + lane;
            n += (uint32_t)__popcll(__ballot(k < nunits && tile0[k] <= tile));
        }
        if (lane == 0) s_unit = n - 1;
    }
    __syncthreads();
    const uint32_t u = __builtin_amdgcn_readfirstlane(s_unit);
    const BatchUnit& U = units[u];
    const uint64_t nb = U.nb;
    const uint2* __restrict__ pairs = reinterpret_cast<const uint2*>(U.pairs);
    const ActParams act = U.act;
    const uint64_t base = (uint64_t)(tile - U.tile0) * kBatchFilterTile;
    uint64_t bal[kBatchFilterItems];
#pragma unroll
    for (uint32_t k = 0; k < kBatchFilterItems; ++k) {
        const uint64_t b = base + k * 256u + threadIdx.x;
        bool a = false;
        if (b < nb) {
            const uint2 p = pairs[b];
            a = block_active(p.x, p.y, act);
        }
        bal[k] = __ballot(a);
        if (lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
    }
    __syncthreads();
    if (w == 0) {
        const uint32_t v = lane < kBatchFilterItems * 4 ? s_cnt[lane] : 0u;
        const uint32_t inc = wave_incl_scan<uint32_t>(v);
        const uint32_t agg = __shfl(inc, 63, 64);
        if (lane < kBatchFilterItems * 4) s_cnt[lane] = inc - v;
        const uint64_t ex = lookback(status, tile, agg);
        if (lane == 0) {
            s_excl = ex;
            if (tile == U.tile0) unit_first[u] = (uint32_t)ex;
            if (tile == ntiles - 1) {
                *count = (uint32_t)(ex + agg);
                unit_first[nunits] = (uint32_t)(ex + agg);
            }
        }
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t k = 0; k < kBatchFilterItems; ++k) {
        const uint64_t b = base + k * 256u + threadIdx.x;
        uint32_t slot = 0xFFFFFFFFu;
        if ((bal[k] >> lane) & 1ull) {
            const uint64_t pos = s_excl + s_cnt[k * 4 + w] + (uint32_t)__popcll(bal[k] & lt);
            ids[pos] = (uint32_t)b;
            item_unit[pos] = u;
            slot = (uint32_t)pos;
        }
        if (b < nb) U.map[b] = slot;
    }
}

hipError_t launch_filter_batch(const BatchUnit* units, const uint32_t* tile0, uint32_t nunits,
                               uint32_t ntiles, uint32_t* ids, uint32_t* item_unit,
                               uint32_t* unit_first, uint32_t* count, ScanState st, hipStream_t s)
{
    if (nunits == 0 || ntiles == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_filter_batch, dim3(ntiles), dim3(256), 0, s, units, tile0, nunits, ntiles, ids,
                       item_unit, unit_first, count, st.status, st.ticket);
    return hipGetLastError();
}

// One thread per unit of the class: its slices of the batch buffers.
__global__ __launch_bounds__(256) void k_batch_ranges(BatchRangeArgs a)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.nunits) return;
    const uint32_t na = *a.count;
    const uint32_t f = a.unit_first[u], l = a.unit_first[u + 1];
    auto at = [&](const uint64_t* base, uint32_t i, uint64_t total) { return i < na ? base[i] : total; };
    const uint64_t v0 = at(a.vbase, f, a.totals[0]), v1 = at(a.vbase, l, a.totals[0]);
    const uint64_t t0 = at(a.tbase, f, a.totals[1]), t1 = at(a.tbase, l, a.totals[1]);
    const uint64_t m0 = at(a.mbase, f, a.totals[3]), m1 = at(a.mbase, l, a.totals[3]);
    mtx_batch_range r;
    r.vertex_offset = a.voff + v0;
    r.vertex_count = v1 - v0;
    r.index_offset = 3 * (a.toff + t0);
    r.index_count = 3 * (t1 - t0);
    r.meshlet_offset = a.moff + m0;
    r.meshlet_count = m1 - m0;
    r.active_offset = (uint32_t)(a.aoff + f);
    r.active_count = l - f;
    r.block_dim = a.block_dim;
    // bit 0: this unit's output did not fit the batch's capacity (blocks past it
    // were skipped); bit 1: a triangle edge's owner block was missing
    r.flags = (v1 > a.vcap || 3 * t1 > a.icap || m1 > a.mcap) ? 1u : 0u;
    r.flags |= (uint32_t)(a.totals[2] & 2ull);
    static_cast<mtx_batch_range*>(a.ranges)[a.unit_global[u]] = r;
}

hipError_t launch_batch_ranges(const BatchRangeArgs& a, hipStream_t s)
{
    if (a.nunits == 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_ranges, dim3((a.nunits + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace mtx
