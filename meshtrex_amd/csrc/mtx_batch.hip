// gfx950 kernel of the batched form's last step: the per-unit output ranges.
#include "../../include/meshtrex_amd/mtx.h"
#include "mtx_device.h"

namespace mtx {

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
