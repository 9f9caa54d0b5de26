// C-ABI host layer: argument checking, device memory, launch sequencing.
#include "mtx_host.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

using namespace mtx;

// Bytes behind a filter's count word for the compaction's tile counts (grown
// for grids of more than 2^17 tiles).
constexpr uint64_t kFilterScratch = 4ull * 131072;

// mtx_extraction_output.flags bit: the buffers include the meshlet outputs.
constexpr uint32_t kWantMeshlets = 16u;

// Scratch of mtx_filter_active_blocks_noalloc: the compaction's per-tile
// active counts (one uint32 per tile, written by the count launch and read by
// the scatter launch, so nothing has to start at zero).  2^18 tiles cover
// grids of up to 2^29 blocks on the pairs path (2048 blocks per tile) and
// 2^24 words of the fused mask (64 words per tile).
constexpr uint32_t kNoallocTiles = 1u << 18;

}  // namespace

// One copy per GPU, held until the process ends; every call on a GPU shares
// it, so calls that overlap on one GPU race.  In host code `scratch` is only a
// stand-in for the symbol: its device address comes from hipGetSymbolAddress
// (scratch_address below) and only that address reaches a kernel.
static __device__ __align__(256) uint32_t scratch[kNoallocTiles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// This device's address of `scratch`, resolved once per device with the
// device current (the first mtx_context_create on it).
hipError_t scratch_address(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

mtx_options defaults()
{
    mtx_options o;
    mtx_options_default(&o);
    return o;
}

// Block planes of the last uploaded slab of a host volume: min/max streams
// the volume about 100x faster than PCIe delivers it (6 TB/s against
// 30-60 GB/s), so the min/max of the head finishes while the tail (1/50 of
// the block planes, at least one) uploads, and only the tail's own min/max is
// left after the last byte lands.  Measured on MI355X, the head's min/max
// does not slow the concurrent copy (a 4 GiB copy: 74.93 ms alone, 75.02 ms
// beside 43 ms of back-to-back min/max), while every extra copy adds a
// start-up, so two slabs.
uint32_t tail_block_planes(uint32_t gz)
{
    return gz < 4 ? gz : (gz + 49) / 50;
}

// Whether `p` is host memory the HIP runtime allocated (pinned).
bool is_pinned(const void* p)
{
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, p);
    (void)hipGetLastError();   // plain host memory reports an error
    return e == hipSuccess && at.type == hipMemoryTypeHost;
}

// Host volume -> device in two slabs of whole block planes; once a slab (and
// the one plane it shares with the next) is resident, the B = 8 min/max runs
// over its block planes on `s`.  A synchronous call (sync) from pinned memory
// copies with blocking copies -- measured on MI355X, 74.5 ms for 4 GiB of
// pinned memory against 74.9 ms for the same copy enqueued on a stream --
// and enqueues the head's min/max before starting the tail's copy; otherwise
// the copies go on the context's copy stream (for pageable memory the
// runtime's staging is faster that way: 143 against 150 ms), the min/max on
// `s` waiting for each.
hipError_t upload_and_minmax(mtx_context* ctx, hipStream_t s, bool sync, const void* host, void* dev,
                             const VolDesc& v, const Grid& g, uint64_t esz, uint32_t* pairs, uint32_t* mask,
                             const ActParams* act, uint32_t rows)
{
    const uint64_t plane = (uint64_t)v.X * v.Y * esz;
    const uint32_t tail = tail_block_planes(g.gz);
    const uint32_t cuts[2] = { g.gz - tail, g.gz };   // block-plane ends of the slabs
    const uint32_t first = cuts[0] ? 0u : 1u;         // gz < 4: the tail is the whole volume
    hipError_t e = hipSuccess;
    sync = sync && is_pinned(host);
    if (sync) {
        // the buffer is free: earlier work on `s` has finished
        e = hipStreamSynchronize(s);
    } else {
        if (!ctx->copy) e = hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking);
        while (e == hipSuccess && ctx->slab_events.size() < 2) {
            hipEvent_t ev = nullptr;
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e == hipSuccess) ctx->slab_events.push_back(ev);
        }
        // the copy stream starts after earlier work on `s` (the buffer may be in use)
        if (e == hipSuccess) e = hipEventRecord(ctx->slab_events[0], s);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->copy, ctx->slab_events[0], 0);
    }
    uint64_t copied = 0;  // planes resident
    uint32_t bzb = 0;
    for (uint32_t k = first; e == hipSuccess && k < 2; ++k) {
        const uint32_t bze = cuts[k];
        uint64_t need = (uint64_t)bze * 8 + 1;  // planes [0, need) for these blocks
        if (need > v.Z) need = v.Z;
        if (need > copied) {
            char* d = static_cast<char*>(dev) + copied * plane;
            const char* h = static_cast<const char*>(host) + copied * plane;
            const uint64_t n = (need - copied) * plane;
            e = sync ? hipMemcpy(d, h, n, hipMemcpyHostToDevice)
                     : hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, ctx->copy);
            copied = need;
        }
        if (!sync) {
            if (e == hipSuccess) e = hipEventRecord(ctx->slab_events[k], ctx->copy);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->slab_events[k], 0);
        }
        if (e == hipSuccess) e = launch_minmax_b8(v, g, bzb, bze, pairs, mask, act, rows, s);
        bzb = bze;
    }
    return e;
}

// Pyramid layout, src/minMaxManager.cpp:85-92,147-151.
uint32_t pyramid_layout(const uint32_t G[3], uint32_t (*ext)[3], uint64_t* off)
{
    uint32_t m = G[0] > G[1] ? G[0] : G[1];
    m = m > G[2] ? m : G[2];
    uint32_t L = 1;
    while ((m >> L) > 0) ++L;
    uint64_t o = 0;
    uint32_t e[3] = { G[0], G[1], G[2] };
    for (uint32_t l = 0; l < L; ++l) {
        ext[l][0] = e[0];
        ext[l][1] = e[1];
        ext[l][2] = e[2];
        off[l] = o;
        o += (uint64_t)e[0] * e[1] * e[2];
        for (int a = 0; a < 3; ++a) e[a] = (e[a] + 1) >> 1;
    }
    off[L] = o;
    return L;
}

// Argument checks shared by the filter entry points; fills the options and
// the block grid.
mtx_status filter_checks(mtx_context* ctx, const mtx_minmax_output* mm, const mtx_push_constants* pc,
                         const mtx_options* opt, mtx_options& o, Grid& g)
{
    if (!mm || !pc) return fail(ctx, MTX_EINVAL, "null argument");
    if (!mm->minmax) return fail(ctx, MTX_ESTATE, "min/max output is empty");
    o = opt ? *opt : defaults();
    if (o.rule > 1) return fail(ctx, MTX_EINVAL, "unknown activity rule %u", o.rule);
    g = Grid{ mm->block_dim, mm->grid[0], mm->grid[1], mm->grid[2] };
    if (pc->blockGridDim[0] != g.gx || pc->blockGridDim[1] != g.gy || pc->blockGridDim[2] != g.gz ||
        pc->blockDim[0] != g.B)
        return fail(ctx, MTX_ESTATE, "push constants do not match the min/max output");
    return MTX_OK;
}

// Whether the min/max pass's fused activity mask holds this isovalue and rule
// (then the filter compacts the mask instead of re-reading the pairs).
bool mask_usable(const mtx_minmax_output* mm, const mtx_push_constants* pc, uint32_t rule)
{
    return mm->active_mask && mm->mask_valid && mm->mask_rule == rule &&
           (mm->mask_isovalue == pc->isovalue || (std::isnan(mm->mask_isovalue) && std::isnan(pc->isovalue)));
}

// The filter's two launches (tile counts, then the ordered scatter of IDs and
// the map); the tile counts go to `ts`.
hipError_t launch_filter(const mtx_minmax_output* mm, const Grid& g, const mtx_push_constants* pc,
                         const mtx_options& o, bool use_mask, uint32_t* ids, uint32_t* count, uint32_t* map,
                         TileSums ts, hipStream_t s)
{
    if (use_mask) return launch_compact_mask(mm->active_mask, g, mm->mask_words_per_row, ids, count, map, ts, s);
    // coarse reject from pyramid level 2 (an ancestor's range contains its blocks')
    const bool coarse = o.coarse_reject && mm->levels_built > 2;
    return launch_filter_pairs(mm->minmax, g, make_act((int)mm->dtype, pc->isovalue, (int)o.rule), ids, count, map,
                               ts, coarse ? mm->minmax + 2 * mm->level_offset[2] : nullptr,
                               coarse ? &mm->level_extent[2] : nullptr, s);
}

// The filter's list and map describe the min/max output's block grid.
bool same_grid(const mtx_minmax_output* mm, const mtx_filtering_output* f)
{
    return f->block_dim == mm->block_dim && f->grid[0] == mm->grid[0] && f->grid[1] == mm->grid[1] &&
           f->grid[2] == mm->grid[2] && f->capacity >= (uint64_t)mm->grid[0] * mm->grid[1] * mm->grid[2];
}


}  // namespace

// ---------------------------------------------------------------------------

extern "C" {

int mtx_abi_version(void) { return MTX_ABI_VERSION; }

const char* mtx_status_string(mtx_status s)
{
    switch (s) {
    case MTX_OK: return "MTX_OK";
    case MTX_EINVAL: return "MTX_EINVAL";
    case MTX_ENOMEM: return "MTX_ENOMEM";
    case MTX_EDEVICE: return "MTX_EDEVICE";
    case MTX_ECAPACITY: return "MTX_ECAPACITY";
    case MTX_ESTATE: return "MTX_ESTATE";
    }
    return "MTX_UNKNOWN";
}

void mtx_options_default(mtx_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->overlap = 1;
    o->rule = MTX_RULE_CPU;
    o->build_pyramid = 1;
    o->fuse_filter = 1;
    o->normals = 1;
    o->meshlets = 1;
}

void mtx_push_constants_init(mtx_push_constants* pc, const uint32_t dims[3], uint32_t block_dim,
                             float isovalue)
{
    if (!pc) return;
    std::memset(pc, 0, sizeof *pc);
    for (int a = 0; a < 3; ++a) {
        pc->volumeDim[a] = dims[a];
        pc->blockDim[a] = block_dim;
        pc->blockGridDim[a] = block_dim ? (dims[a] + block_dim - 1) / block_dim : 0;
    }
    pc->volumeDim[3] = 1;
    pc->blockDim[3] = 1;
    pc->blockGridDim[3] = 1;
    pc->isovalue = isovalue;
}

mtx_status mtx_context_create(int device, mtx_context** out)
{
    if (!out) return MTX_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MTX_EDEVICE;
    if (device < 0 || device >= n) return MTX_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MTX_EDEVICE;
    mtx_context* c = new (std::nothrow) mtx_context;
    if (!c) return MTX_ENOMEM;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MTX_EDEVICE;
    }
    // set-up of mtx_filter_active_blocks_noalloc: this device's scratch address
    if (scratch_address(device, &c->scratch) != hipSuccess) {
        hipStreamDestroy(c->own);
        delete c;
        return MTX_EDEVICE;
    }
    *out = c;
    return MTX_OK;
}

void mtx_context_destroy(mtx_context* ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->own) {
        hipStreamSynchronize(ctx->own);
        hipStreamDestroy(ctx->own);
    }
    if (ctx->copy) {
        hipStreamSynchronize(ctx->copy);
        hipStreamDestroy(ctx->copy);
    }
    if (ctx->side) {
        hipStreamSynchronize(ctx->side);
        hipStreamDestroy(ctx->side);
    }
    for (hipEvent_t e : ctx->side_ev)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : ctx->slab_events) hipEventDestroy(e);
    delete ctx;
}

const char* mtx_context_last_error(const mtx_context* ctx)
{
    return ctx ? ctx->last_error.c_str() : "no context";
}

uint32_t mtx_key_f32(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

float mtx_unkey_f32(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// ---- min/max -------------------------------------------------------------------

mtx_status mtx_compute_min_max(mtx_context* ctx, const mtx_volume* volume,
                               const mtx_push_constants* pc, const mtx_options* opt, void* stream,
                               mtx_minmax_output* out)
{
    if (!ctx) return MTX_EINVAL;
    if (!volume || !pc || !out) return fail(ctx, MTX_EINVAL, "null argument");
    if (!volume->data) return fail(ctx, MTX_EINVAL, "volume data is null");
    const mtx_options o = opt ? *opt : defaults();
    const uint32_t X = volume->dims[0], Y = volume->dims[1], Z = volume->dims[2];
    const uint32_t B = pc->blockDim[0];
    if (volume->dtype != MTX_U8 && volume->dtype != MTX_F32)
        return fail(ctx, MTX_EINVAL, "unsupported dtype %u", volume->dtype);
    if (X == 0 || Y == 0 || Z == 0) return fail(ctx, MTX_EINVAL, "empty volume %ux%ux%u", X, Y, Z);
    if ((uint64_t)X * Y > 0xFFFFFFFFull || (uint64_t)X * Y * Z > (1ull << 40))
        return fail(ctx, MTX_EINVAL, "volume %ux%ux%u too large", X, Y, Z);
    if (B < 1 || B > 16) return fail(ctx, MTX_EINVAL, "block edge %u outside 1..16", B);
    if (pc->blockDim[1] != B || pc->blockDim[2] != B)
        return fail(ctx, MTX_EINVAL, "blocks must be cubic (blockDim %u,%u,%u)", B, pc->blockDim[1],
                    pc->blockDim[2]);
    if (pc->volumeDim[0] != X || pc->volumeDim[1] != Y || pc->volumeDim[2] != Z)
        return fail(ctx, MTX_EINVAL, "push constants volumeDim does not match the volume");
    if (o.overlap > 1) return fail(ctx, MTX_EINVAL, "overlap must be 0 or 1");
    if (o.minmax_rows != 0 && o.minmax_rows != 1 && o.minmax_rows != 2 && o.minmax_rows != 4)
        return fail(ctx, MTX_EINVAL, "minmax_rows must be 0, 1, 2 or 4");
    Grid g{ B, (X + B - 1) / B, (Y + B - 1) / B, (Z + B - 1) / B };
    if (pc->blockGridDim[0] != g.gx || pc->blockGridDim[1] != g.gy || pc->blockGridDim[2] != g.gz)
        return fail(ctx, MTX_EINVAL, "blockGridDim must be ceil(volumeDim / blockDim)");
    if (g.nb() > 0xFFFFFFFFull) return fail(ctx, MTX_EINVAL, "too many blocks");

    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;

    const uint64_t esz = volume->dtype == MTX_F32 ? 4 : 1;
    const uint64_t vbytes = (uint64_t)X * Y * Z * esz;
    const bool fast = B == 8 && o.overlap == 1;
    const bool fuse = fast && o.fuse_filter;
    // reuse: same problem geometry as the buffers in `out` -> no allocation
    const bool reuse = o.reuse && out->minmax && out->dtype == volume->dtype && out->dims[0] == X &&
                       out->dims[1] == Y && out->dims[2] == Z && out->block_dim == B &&
                       (volume->on_device || out->owned_volume) && (!fuse || out->active_mask);
    if (!reuse) {
        if (o.reuse && out->minmax) mtx_minmax_release(ctx, out);
        std::memset(out, 0, sizeof *out);
        out->dtype = volume->dtype;
        out->dims[0] = X; out->dims[1] = Y; out->dims[2] = Z;
        out->block_dim = B;
        out->grid[0] = g.gx; out->grid[1] = g.gy; out->grid[2] = g.gz;
        out->num_levels = pyramid_layout(out->grid, out->level_extent, out->level_offset);
        if (out->num_levels > MTX_MAX_LEVELS) return fail(ctx, MTX_EINVAL, "pyramid too deep");
        if (!volume->on_device) MTX_HIP(ctx, hipMalloc(&out->owned_volume, vbytes));
        hipError_t e = hipMalloc(&out->minmax, 8 * out->level_offset[out->num_levels]);
        if (e == hipSuccess && fuse) {
            out->mask_words_per_row = (g.gx + 31) / 32;
            e = hipMalloc(&out->active_mask, 4 * (uint64_t)out->mask_words_per_row * g.gy * g.gz);
        }
        if (e != hipSuccess) {
            mtx_minmax_release(ctx, out);
            return fail(ctx, MTX_ENOMEM, "min/max buffers");
        }
    }
    out->overlap = o.overlap;
    out->mask_valid = 0;
    out->levels_built = 1;
    out->volume = volume->on_device ? volume->data : out->owned_volume;
    VolDesc v{ out->volume, X, Y, Z, (int)volume->dtype };
    ActParams act = make_act((int)volume->dtype, pc->isovalue, (int)o.rule);
    uint32_t* mask = fuse ? out->active_mask : nullptr;
    hipError_t e;
    if (volume->on_device) {
        e = fast ? launch_minmax_b8(v, g, 0, g.gz, out->minmax, mask, &act, o.minmax_rows, bd.s)
                 : launch_minmax_generic(v, g, (int)o.overlap, out->minmax, bd.s);
    } else if (fast) {
        e = upload_and_minmax(ctx, bd.s, bd.sync, volume->data, out->owned_volume, v, g, esz, out->minmax, mask,
                              &act, o.minmax_rows);
    } else {
        e = hipMemcpyAsync(out->owned_volume, volume->data, vbytes, hipMemcpyHostToDevice, bd.s);
        if (e == hipSuccess) e = launch_minmax_generic(v, g, (int)o.overlap, out->minmax, bd.s);
    }
    if (e == hipSuccess && o.build_pyramid && out->num_levels > 1) {
        e = launch_pyramid(out->minmax, out->num_levels, out->level_extent, out->level_offset, bd.s);
        out->levels_built = out->num_levels;
    }
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "min/max launch: %s", hipGetErrorString(e));
    if (fuse) {
        out->mask_isovalue = pc->isovalue;
        out->mask_rule = o.rule;
        out->mask_valid = 1;
    }
    return finish(bd);
}

mtx_status mtx_build_pyramid(mtx_context* ctx, mtx_minmax_output* mm, void* stream)
{
    if (!ctx) return MTX_EINVAL;
    if (!mm || !mm->minmax) return fail(ctx, MTX_ESTATE, "min/max output is empty");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;
    if (mm->num_levels > 1) {
        hipError_t e = launch_pyramid(mm->minmax, mm->num_levels, mm->level_extent, mm->level_offset, bd.s);
        if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "pyramid launch: %s", hipGetErrorString(e));
    }
    mm->levels_built = mm->num_levels;
    return finish(bd);
}

void mtx_minmax_release(mtx_context* ctx, mtx_minmax_output* out)
{
    if (!out) return;
    if (ctx) hipSetDevice(ctx->device);
    if (out->minmax) hipFree(out->minmax);
    if (out->active_mask) hipFree(out->active_mask);
    if (out->owned_volume) hipFree(out->owned_volume);
    std::memset(out, 0, sizeof *out);
}

// ---- filter ------------------------------------------------------------------

mtx_status mtx_filter_active_blocks(mtx_context* ctx, const mtx_minmax_output* mm,
                                    const mtx_push_constants* pc, const mtx_options* opt,
                                    void* stream, mtx_filtering_output* out)
{
    if (!ctx) return MTX_EINVAL;
    if (!out) return fail(ctx, MTX_EINVAL, "null argument");
    mtx_options o;
    Grid g;
    if (mtx_status st = filter_checks(ctx, mm, pc, opt, o, g)) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;

    const bool reuse = out->compacted_block_ids && out->block_to_active && out->capacity >= g.nb();
    if (!reuse) {
        if (out->compacted_block_ids || out->done) mtx_filtering_release(ctx, out);
        std::memset(out, 0, sizeof *out);
        MTX_HIP(ctx, hipMalloc(&out->compacted_block_ids, 4 * (g.nb() ? g.nb() : 1)));
        hipError_t e = hipMalloc(&out->active_block_count, 256 + kFilterScratch);
        if (e == hipSuccess) e = hipMalloc(&out->block_to_active, 4 * (g.nb() ? g.nb() : 1));
        if (e != hipSuccess) {
            mtx_filtering_release(ctx, out);
            return fail(ctx, MTX_ENOMEM, "filter buffers");
        }
        out->capacity = (uint32_t)g.nb();
    }
    out->activeBlockCount = 0;
    out->isovalue = pc->isovalue;
    out->rule = o.rule;
    out->block_dim = g.B;
    out->grid[0] = g.gx; out->grid[1] = g.gy; out->grid[2] = g.gz;
    const bool use_mask = mask_usable(mm, pc, o.rule);
    const uint32_t tiles = filter_tiles(g.nb(), use_mask, mm->mask_words_per_row, g);
    if (4ull * tiles > kFilterScratch) {
        // the tile counts live behind the count word; grow them for very large grids
        uint32_t* big = nullptr;
        if (hipMalloc(&big, 256 + 4ull * tiles) != hipSuccess)
            return fail(ctx, MTX_ENOMEM, "filter tile counts");
        hipFree(out->active_block_count);
        out->active_block_count = big;
    }
    TileSums ts{ reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(out->active_block_count) + 256) };
    const hipError_t e = launch_filter(mm, g, pc, o, use_mask, out->compacted_block_ids, out->active_block_count,
                                       out->block_to_active, ts, bd.s);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "filter launch: %s", hipGetErrorString(e));
    if (mtx_status st = mark_done(bd, &out->done)) return st;
    return finish(bd);
}

mtx_status mtx_filter_active_blocks_noalloc(mtx_context* ctx, const mtx_minmax_output* mm,
                                            const mtx_push_constants* pc, const mtx_options* opt,
                                            void* stream, uint32_t* compacted_block_ids,
                                            uint32_t* active_block_count, uint32_t* block_to_active)
{
    if (!ctx) return MTX_EINVAL;
    if (!compacted_block_ids || !active_block_count || !block_to_active)
        return fail(ctx, MTX_EINVAL, "null output buffer");
    mtx_options o;
    Grid g;
    if (mtx_status st = filter_checks(ctx, mm, pc, opt, o, g)) return st;
    const bool use_mask = mask_usable(mm, pc, o.rule);
    const uint32_t tiles = filter_tiles(g.nb(), use_mask, mm->mask_words_per_row, g);
    if (tiles > kNoallocTiles)
        return fail(ctx, MTX_EINVAL, "grid of %llu blocks needs %u compaction tiles; the static scratch holds %u",
                    (unsigned long long)g.nb(), tiles, kNoallocTiles);
    if (!ctx->scratch) return fail(ctx, MTX_EDEVICE, "no scratch address for device %d", ctx->device);
    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;
    const TileSums ts{ static_cast<uint64_t*>(ctx->scratch) };
    const hipError_t e = launch_filter(mm, g, pc, o, use_mask, compacted_block_ids, active_block_count,
                                       block_to_active, ts, bd.s);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "filter launch: %s", hipGetErrorString(e));
    return finish(bd);
}

mtx_status mtx_read_active_count(mtx_context* ctx, mtx_filtering_output* out)
{
    if (!ctx) return MTX_EINVAL;
    if (!out || !out->active_block_count) return fail(ctx, MTX_ESTATE, "filtering output is empty");
    std::lock_guard<std::mutex> lk(ctx->mu);
    uint32_t n = 0;
    if (mtx_status st = read_after(ctx, out->done, &n, out->active_block_count, 4)) return st;
    out->activeBlockCount = n;
    return MTX_OK;
}

void mtx_filtering_release(mtx_context* ctx, mtx_filtering_output* out)
{
    if (!out) return;
    if (ctx) hipSetDevice(ctx->device);
    if (out->compacted_block_ids) hipFree(out->compacted_block_ids);
    if (out->active_block_count) hipFree(out->active_block_count);
    if (out->block_to_active) hipFree(out->block_to_active);
    destroy_event(&out->done);
    std::memset(out, 0, sizeof *out);
}

// ---- extraction ------------------------------------------------------------------

// The vertex pass runs on a side stream only for large extractions: the
// fork and join cost a few microseconds of cross-queue waits, which a small
// volume's passes do not win back (measured on MI355X: 225 small units went
// from 14.3 to 17.4 ms with the side stream; the 1024^3 headline step gains
// about 25 us).  The preallocated form knows only the block grid, the
// allocating form the vertex count.
constexpr uint64_t kSideMinBlocks = 1ull << 20;
constexpr uint64_t kSideMinVertices = 1ull << 21;

// The context's side stream and events for the vertex pass (created on first
// use), set in `a`.
static hipError_t side_stream(mtx_context* ctx, ExtractArgs& a)
{
    hipError_t e = hipSuccess;
    if (!ctx->side) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
        if (!ctx->side_ev[k]) e = hipEventCreateWithFlags(&ctx->side_ev[k], hipEventDisableTiming);
    if (e != hipSuccess) return e;
    a.side = ctx->side;
    a.side_ev[0] = ctx->side_ev[0];
    a.side_ev[1] = ctx->side_ev[1];
    return hipSuccess;
}

static ExtractArgs make_extract_args(const mtx_minmax_output* mm, const mtx_filtering_output* filt,
                                     const mtx_push_constants* pc, mtx_extraction_output* out,
                                     const WsLayout& w)
{
    ExtractArgs a{};
    a.vol = VolDesc{ mm->volume, mm->dims[0], mm->dims[1], mm->dims[2], (int)mm->dtype };
    a.grid = Grid{ mm->block_dim, mm->grid[0], mm->grid[1], mm->grid[2] };
    a.iso = pc->isovalue;
    a.ids = filt->compacted_block_ids;
    a.count = filt->active_block_count;
    a.blk_map = filt->block_to_active;
    char* ws = static_cast<char*>(out->workspace);
    a.blk_counts = reinterpret_cast<uint32_t*>(ws + w.counts);
    a.blk_vbase = reinterpret_cast<uint64_t*>(ws + w.vbase);
    a.blk_tbase = reinterpret_cast<uint64_t*>(ws + w.tbase);
    a.blk_mbase = reinterpret_cast<uint64_t*>(ws + w.mbase);
    a.blk_sbase = reinterpret_cast<uint64_t*>(ws + w.sbase);
    a.corner_table = ws + w.table;
    a.totals = reinterpret_cast<uint64_t*>(ws + w.totals);
    a.cases = reinterpret_cast<uint8_t*>(ws + w.cases);
    a.vertices = reinterpret_cast<float*>(out->vertices);
    a.indices = out->indices;
    a.edge_ids = out->vertex_edge_ids;
    a.vox_pos = out->vertex_voxel_pos;
    a.edge_t = out->vertex_t;
    // meshlets when the output was made for them
    if (out->flags & kWantMeshlets) {
        a.starts = reinterpret_cast<uint32_t*>(ws + w.starts);
        a.meshlets = reinterpret_cast<uint4*>(out->meshlets);
        a.mverts = reinterpret_cast<float*>(out->meshlet_vertices);
        a.mindices = out->meshlet_indices;
        a.mvmap = out->meshlet_vertex_map;
        a.mcap = out->meshlet_capacity;
        a.mvcap = out->meshlet_vertex_capacity;
    }
    a.vcap = out->vertex_capacity;
    a.icap = out->index_capacity;
    a.out_counts = out->counts;
    a.max_blocks = out->max_blocks;
    return a;
}

mtx_status mtx_extract(mtx_context* ctx, const mtx_minmax_output* mm,
                       const mtx_filtering_output* filt, const mtx_push_constants* pc,
                       const mtx_options* opt, void* stream, mtx_extraction_output* out)
{
    if (!ctx) return MTX_EINVAL;
    if (!mm || !filt || !pc || !out) return fail(ctx, MTX_EINVAL, "null argument");
    if (!mm->minmax || !mm->volume) return fail(ctx, MTX_ESTATE, "min/max output is empty");
    if (!filt->compacted_block_ids) return fail(ctx, MTX_ESTATE, "filtering output is empty");
    if (mm->overlap != 1)
        return fail(ctx, MTX_ESTATE, "extraction needs overlap-1 block windows");
    if (!extract_supported_block(mm->block_dim))
        return fail(ctx, MTX_EINVAL, "block edge %u not supported", mm->block_dim);
    if (pc->blockDim[0] != mm->block_dim || pc->blockGridDim[0] != mm->grid[0] ||
        pc->blockGridDim[1] != mm->grid[1] || pc->blockGridDim[2] != mm->grid[2])
        return fail(ctx, MTX_ESTATE, "push constants do not match the min/max output");
    if (!same_grid(mm, filt))
        return fail(ctx, MTX_ESTATE, "filtering output was made for another block grid");
    if (filt->isovalue != pc->isovalue && !(std::isnan(filt->isovalue) && std::isnan(pc->isovalue)))
        return fail(ctx, MTX_ESTATE, "filter was run for isovalue %g, extraction asks %g",
                    (double)filt->isovalue, (double)pc->isovalue);
    const mtx_options o = opt ? *opt : defaults();
    const uint32_t B = mm->block_dim;
    const uint64_t nb = (uint64_t)mm->grid[0] * mm->grid[1] * mm->grid[2];
    const uint32_t want = (o.normals ? 1u : 0u) | (o.edge_ids ? 2u : 0u) |
                          (o.voxel_positions ? 8u : 0u) | (o.meshlets ? kWantMeshlets : 0u);
    const WsLayout w = ws_layout(nb, B);

    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;

    if (o.reuse) {
        if (!out->workspace || out->workspace_bytes < w.bytes || out->block_dim != B ||
            out->max_blocks < nb || (out->flags & want) != want || !out->vertices || !out->indices ||
            out->grid[0] != mm->grid[0] || out->grid[1] != mm->grid[1] || out->grid[2] != mm->grid[2])
            return fail(ctx, MTX_ESTATE, "reuse: output buffers were made for another problem");
        mtx_extraction_output view = *out;   // this call's options decide what is written
        if (!o.meshlets) view.flags &= ~kWantMeshlets;
        ExtractArgs a = make_extract_args(mm, filt, pc, &view, w);
        a.normals = o.normals ? 1 : 0;
        if (!o.edge_ids) {
            a.edge_ids = nullptr;
            a.edge_t = nullptr;
        }
        if (!o.voxel_positions) a.vox_pos = nullptr;
        TileSums ts{ reinterpret_cast<uint64_t*>(static_cast<char*>(out->workspace) + w.scan) };
        hipError_t e = launch_extract_count(a, bd.s);
        if (e == hipSuccess) e = launch_extract_scan(a, ts, bd.s);
        if (e == hipSuccess && nb >= kSideMinBlocks) e = side_stream(ctx, a);
        if (e == hipSuccess) e = launch_extract_emit(a, bd.s);
        if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "extract launch: %s", hipGetErrorString(e));
        if (mtx_status st = mark_done(bd, &out->done)) return st;
        return finish(bd);
    }

    // Allocating form: count, scan, read the totals, allocate exactly, emit.
    if (out->workspace || out->done) mtx_extraction_release(ctx, out);
    std::memset(out, 0, sizeof *out);
    out->block_dim = B;
    out->grid[0] = mm->grid[0]; out->grid[1] = mm->grid[1]; out->grid[2] = mm->grid[2];
    out->max_blocks = (uint32_t)nb;
    out->flags = want;
    out->workspace_bytes = w.bytes;
    hipError_t e = hipMalloc(&out->workspace, w.bytes);
    if (e == hipSuccess) e = hipMalloc(&out->counts, 256);
    if (e == hipSuccess) {
        out->case_indices = static_cast<uint8_t*>(out->workspace) + w.cases;
        out->case_capacity = nb * B * B * B;
    }
    if (e != hipSuccess) {
        mtx_extraction_release(ctx, out);
        return fail(ctx, MTX_ENOMEM, "extraction workspace (%llu B)", (unsigned long long)w.bytes);
    }
    ExtractArgs a = make_extract_args(mm, filt, pc, out, w);
    a.normals = o.normals ? 1 : 0;
    TileSums ts{ reinterpret_cast<uint64_t*>(static_cast<char*>(out->workspace) + w.scan) };
    e = hipMemsetAsync(out->counts, 0, 256, bd.s);
    if (e == hipSuccess) e = launch_extract_count(a, bd.s);
    if (e == hipSuccess) e = launch_extract_scan(a, ts, bd.s);
    uint64_t tot[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    if (e == hipSuccess) e = hipMemcpyAsync(tot, a.totals, 64, hipMemcpyDeviceToHost, bd.s);
    if (e == hipSuccess) e = hipStreamSynchronize(bd.s);
    if (e != hipSuccess) {
        mtx_extraction_release(ctx, out);
        return fail(ctx, MTX_EDEVICE, "extract count/scan: %s", hipGetErrorString(e));
    }
    if (tot[0] > 0xFFFFFFFFull || 3 * tot[1] > 0xFFFFFFFFull || tot[4] > 0xFFFFFFFFull) {
        mtx_extraction_release(ctx, out);
        return fail(ctx, MTX_ECAPACITY, "mesh too large for 32-bit indices (%llu vertices, %llu triangles)",
                    (unsigned long long)tot[0], (unsigned long long)tot[1]);
    }
    out->vertex_capacity = tot[0];
    out->index_capacity = 3 * tot[1];
    out->meshlet_capacity = tot[3];
    out->meshlet_vertex_capacity = tot[4];
    e = hipMalloc(&out->vertices, sizeof(mtx_vertex) * (tot[0] ? tot[0] : 1));
    if (e == hipSuccess) e = hipMalloc(&out->indices, 4 * (tot[1] ? 3 * tot[1] : 1));
    if (e == hipSuccess && o.meshlets) e = hipMalloc(&out->meshlets, sizeof(mtx_meshlet) * (tot[3] ? tot[3] : 1));
    if (e == hipSuccess && o.meshlets) e = hipMalloc(&out->meshlet_vertices, sizeof(mtx_vertex) * (tot[4] ? tot[4] : 1));
    if (e == hipSuccess && o.meshlets) e = hipMalloc(&out->meshlet_vertex_map, 4 * (tot[4] ? tot[4] : 1));
    if (e == hipSuccess && o.meshlets) e = hipMalloc(&out->meshlet_indices, 4 * (tot[1] ? 3 * tot[1] : 1));
    if (e == hipSuccess && o.edge_ids) e = hipMalloc(&out->vertex_edge_ids, 8 * (tot[0] ? tot[0] : 1));
    if (e == hipSuccess && o.edge_ids) e = hipMalloc(&out->vertex_t, 4 * (tot[0] ? tot[0] : 1));
    if (e == hipSuccess && o.voxel_positions) e = hipMalloc(&out->vertex_voxel_pos, 12 * (tot[0] ? tot[0] : 1));
    if (e != hipSuccess) {
        mtx_extraction_release(ctx, out);
        return fail(ctx, MTX_ENOMEM, "extraction outputs (%llu vertices, %llu triangles)",
                    (unsigned long long)tot[0], (unsigned long long)tot[1]);
    }
    a = make_extract_args(mm, filt, pc, out, w);
    a.normals = o.normals ? 1 : 0;
    e = tot[0] >= kSideMinVertices ? side_stream(ctx, a) : hipSuccess;
    if (e == hipSuccess) e = launch_extract_emit(a, bd.s);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "extract emit: %s", hipGetErrorString(e));
    if (mtx_status st = mark_done(bd, &out->done)) return st;
    out->vertexCount = (uint32_t)tot[0];
    out->indexCount = (uint32_t)(3 * tot[1]);
    out->meshletCount = (uint32_t)tot[3];
    out->meshletVertexCount = (uint32_t)tot[4];
    return finish(bd);
}

mtx_status mtx_adjoint(mtx_context* ctx, const mtx_minmax_output* mm, const mtx_filtering_output* filt,
                       const mtx_extraction_output* ext, const mtx_push_constants* pc,
                       const float* dl_dpos, float* dl_dvol, float* dl_diso, void* stream)
{
    if (!ctx) return MTX_EINVAL;
    if (!mm || !filt || !ext || !pc || !dl_dvol || !dl_diso) return fail(ctx, MTX_EINVAL, "null argument");
    if (!ext->workspace || !mm->volume || !filt->block_to_active)
        return fail(ctx, MTX_ESTATE, "adjoint needs the forward pass's outputs");
    if (ext->vertexCount && !dl_dpos) return fail(ctx, MTX_EINVAL, "dl_dpos is null");
    const uint32_t B = mm->block_dim;
    const uint64_t nb = (uint64_t)mm->grid[0] * mm->grid[1] * mm->grid[2];
    if (ext->block_dim != B || ext->max_blocks < nb || ext->grid[0] != mm->grid[0] ||
        ext->grid[1] != mm->grid[1] || ext->grid[2] != mm->grid[2])
        return fail(ctx, MTX_ESTATE, "extraction output does not belong to this min/max output");
    if (!same_grid(mm, filt))
        return fail(ctx, MTX_ESTATE, "filtering output was made for another block grid");
    if (mm->overlap != 1) return fail(ctx, MTX_ESTATE, "adjoint needs overlap-1 block windows");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;
    const WsLayout w = ws_layout(ext->max_blocks, B);
    mtx_extraction_output tmp = *ext;
    AdjointArgs a{};
    a.fwd = make_extract_args(mm, filt, pc, &tmp, w);
    a.dl_dpos = dl_dpos;
    a.dl_dvol = dl_dvol;
    a.dl_diso = dl_diso;
    a.iso_partials = reinterpret_cast<double*>(static_cast<char*>(ext->workspace) + w.adj);
    a.n_wg = adjoint_workgroups(nb);
    hipError_t e = launch_adjoint(a, bd.s);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "adjoint launch: %s", hipGetErrorString(e));
    return finish(bd);
}

mtx_status mtx_read_extraction_counts(mtx_context* ctx, mtx_extraction_output* out)
{
    if (!ctx) return MTX_EINVAL;
    if (!out || !out->counts || !out->workspace) return fail(ctx, MTX_ESTATE, "extraction output is empty");
    std::lock_guard<std::mutex> lk(ctx->mu);
    uint32_t c[4] = { 0, 0, 0, 0 };
    if (mtx_status st = read_after(ctx, out->done, c, out->counts, 16)) return st;
    const WsLayout w = ws_layout(out->max_blocks, out->block_dim);
    uint64_t tot[5] = { 0, 0, 0, 0, 0 };
    if (mtx_status st = read_after(ctx, nullptr, tot, static_cast<char*>(out->workspace) + w.totals, 40))
        return st;
    out->vertexCount = c[0];
    out->indexCount = c[1];
    out->meshletCount = c[2];
    out->meshletVertexCount = c[3];
    if (tot[2] & 2) return fail(ctx, MTX_ESTATE, "active block list does not match the volume");
    if (tot[2] & 4) return fail(ctx, MTX_EDEVICE, "meshlet cut of the triangle pass differs from the count pass");
    if (tot[2] & 1)
        return fail(ctx, MTX_ECAPACITY,
                    "output buffers too small: need %llu vertices, %llu indices, %llu meshlets, %llu meshlet vertices",
                    (unsigned long long)tot[0], (unsigned long long)(3 * tot[1]), (unsigned long long)tot[3],
                    (unsigned long long)tot[4]);
    return MTX_OK;
}

void mtx_extraction_release(mtx_context* ctx, mtx_extraction_output* out)
{
    if (!out) return;
    if (ctx) hipSetDevice(ctx->device);
    // case_indices points into the workspace
    void* ptrs[] = { out->vertices, out->counts, out->indices, out->meshlets, out->vertex_edge_ids,
                     out->vertex_t, out->vertex_voxel_pos, out->workspace, out->meshlet_vertices,
                     out->meshlet_indices, out->meshlet_vertex_map };
    for (void* p : ptrs)
        if (p) hipFree(p);
    destroy_event(&out->done);
    std::memset(out, 0, sizeof *out);
}

// ---- plumbing ---------------------------------------------------------------

mtx_status mtx_device_malloc(void** ptr, uint64_t bytes)
{
    if (!ptr) return MTX_EINVAL;
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    return e == hipSuccess ? MTX_OK : (e == hipErrorOutOfMemory ? MTX_ENOMEM : MTX_EDEVICE);
}

mtx_status mtx_host_malloc(void** ptr, uint64_t bytes)
{
    if (!ptr) return MTX_EINVAL;
    *ptr = nullptr;
    return hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? MTX_OK : MTX_ENOMEM;
}

mtx_status mtx_host_free(void* ptr)
{
    return hipHostFree(ptr) == hipSuccess ? MTX_OK : MTX_EINVAL;
}

mtx_status mtx_device_free(void* ptr)
{
    return hipFree(ptr) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_memcpy_h2d(void* dst, const void* src, uint64_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_memcpy_d2h(void* dst, const void* src, uint64_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_memcpy_d2d(void* dst, const void* src, uint64_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_memcpy_d2d_async(void* dst, const void* src, uint64_t bytes, void* stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)) == hipSuccess
               ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_stream_create(void** stream)
{
    if (!stream) return MTX_EINVAL;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MTX_EDEVICE;
    *stream = s;
    return MTX_OK;
}

mtx_status mtx_stream_destroy(void* stream)
{
    if (!stream) return MTX_EINVAL;
    return hipStreamDestroy(static_cast<hipStream_t>(stream)) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_stream_synchronize(void* stream)
{
    return hipStreamSynchronize(static_cast<hipStream_t>(stream)) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

// ---- profiler -----------------------------------------------------------------

struct mtx_profiler {
    struct Region {
        std::string name;
        hipEvent_t begin = nullptr, end = nullptr;
    };
    std::vector<Region> regions;
    std::vector<size_t> open;
};

mtx_status mtx_profiler_create(mtx_profiler** out)
{
    if (!out) return MTX_EINVAL;
    *out = new (std::nothrow) mtx_profiler;
    return *out ? MTX_OK : MTX_ENOMEM;
}

void mtx_profiler_reset(mtx_profiler* p)
{
    if (!p) return;
    for (auto& r : p->regions) {
        if (r.begin) hipEventDestroy(r.begin);
        if (r.end) hipEventDestroy(r.end);
    }
    p->regions.clear();
    p->open.clear();
}

void mtx_profiler_destroy(mtx_profiler* p)
{
    mtx_profiler_reset(p);
    delete p;
}

mtx_status mtx_profiler_begin(mtx_profiler* p, const char* name, void* stream)
{
    if (!p || !name) return MTX_EINVAL;
    mtx_profiler::Region r;
    r.name = name;
    if (hipEventCreate(&r.begin) != hipSuccess || hipEventCreate(&r.end) != hipSuccess) return MTX_EDEVICE;
    if (hipEventRecord(r.begin, static_cast<hipStream_t>(stream)) != hipSuccess) return MTX_EDEVICE;
    p->regions.push_back(r);
    p->open.push_back(p->regions.size() - 1);
    return MTX_OK;
}

mtx_status mtx_profiler_end(mtx_profiler* p, void* stream)
{
    if (!p || p->open.empty()) return MTX_ESTATE;
    const size_t i = p->open.back();
    p->open.pop_back();
    return hipEventRecord(p->regions[i].end, static_cast<hipStream_t>(stream)) == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

mtx_status mtx_profiler_region(mtx_profiler* p, uint32_t index, char* name, uint32_t cap, float* ms)
{
    if (!p || index >= p->regions.size()) return MTX_EINVAL;
    auto& r = p->regions[index];
    if (hipEventSynchronize(r.end) != hipSuccess) return MTX_EDEVICE;
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.begin, r.end) != hipSuccess) return MTX_EDEVICE;
    if (ms) *ms = t;
    if (name && cap) {
        std::strncpy(name, r.name.c_str(), cap - 1);
        name[cap - 1] = 0;
    }
    return MTX_OK;
}

// ---- OBJ --------------------------------------------------------------------

mtx_status mtx_write_obj(mtx_context* ctx, mtx_extraction_output* ext, const char* path)
{
    if (!ctx) return MTX_EINVAL;
    if (!ext || !path) return fail(ctx, MTX_EINVAL, "null argument");
    if (mtx_status st = mtx_read_extraction_counts(ctx, ext)) return st;
    const uint64_t nv = ext->vertexCount, ni = ext->indexCount;
    std::vector<mtx_vertex> v(nv);
    std::vector<uint32_t> idx(ni);
    if (nv) MTX_HIP(ctx, hipMemcpy(v.data(), ext->vertices, nv * sizeof(mtx_vertex), hipMemcpyDeviceToHost));
    if (ni) MTX_HIP(ctx, hipMemcpy(idx.data(), ext->indices, ni * 4, hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(ctx, MTX_EINVAL, "cannot open %s for writing", path);
    std::fprintf(f, "# isosurface: %llu vertices (one per active edge), %llu triangles\n\n",
                 (unsigned long long)nv, (unsigned long long)(ni / 3));
    std::fprintf(f, "# Vertex positions\n");
    for (const auto& p : v) std::fprintf(f, "v %g %g %g\n", p.position[0], p.position[1], p.position[2]);
    std::fprintf(f, "\n# Vertex normals\n");
    for (const auto& p : v) std::fprintf(f, "vn %g %g %g\n", p.normal[0], p.normal[1], p.normal[2]);
    std::fprintf(f, "\n# Faces\n");
    for (uint64_t i = 0; i + 2 < ni; i += 3) {
        const uint32_t a = idx[i] + 1, b = idx[i + 1] + 1, c = idx[i + 2] + 1;
        std::fprintf(f, "f %u//%u %u//%u %u//%u\n", a, a, b, b, c, c);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? MTX_OK : fail(ctx, MTX_EDEVICE, "write of %s failed", path);
}

uint64_t mtx_struct_size(const char* name)
{
    if (!name) return 0;
#define MTX_SZ(T) \
    if (std::strcmp(name, #T) == 0) return sizeof(T);
    MTX_SZ(mtx_push_constants)
    MTX_SZ(mtx_volume)
    MTX_SZ(mtx_options)
    MTX_SZ(mtx_minmax_output)
    MTX_SZ(mtx_filtering_output)
    MTX_SZ(mtx_vertex)
    MTX_SZ(mtx_meshlet)
    MTX_SZ(mtx_extraction_output)
    MTX_SZ(mtx_batch_volume)
    MTX_SZ(mtx_batch_options)
    MTX_SZ(mtx_batch_range)
    MTX_SZ(mtx_batch_output)
    MTX_SZ(mtx_batch_class_info)
#undef MTX_SZ
    return 0;
}

mtx_status mtx_device_synchronize(void)
{
    return hipDeviceSynchronize() == hipSuccess ? MTX_OK : MTX_EDEVICE;
}

}  // extern "C"
