// Host-side internals shared by the C-ABI translation units: the context,
// error reporting, stream binding.  Not installed; not part of the ABI.
#pragma once

#include "../../include/meshtrex_amd/mtx.h"
#include "mtx_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

struct mtx_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t copy = nullptr;           // host-to-device uploads (created on first use)
    hipStream_t side = nullptr;           // extraction's vertex pass beside its triangle pass
    hipEvent_t side_ev[2] = { nullptr, nullptr };
    std::vector<hipEvent_t> slab_events;  // one per uploaded slab in flight
    void* scratch = nullptr;              // this device's address of the static scratch array
    std::string last_error;
    std::mutex mu;
};

namespace mtx {

inline mtx_status fail(mtx_context* ctx, mtx_status st, const char* fmt, ...)
{
    // a failed runtime call (e.g. an allocation) leaves its error as the
    // thread's last error; clear it so that a later launch check does not
    // report it again
    (void)hipGetLastError();
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return st;
}

#define MTX_HIP(ctx, call)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::mtx::fail((ctx), e_ == hipErrorOutOfMemory ? MTX_ENOMEM : MTX_EDEVICE, \
                               "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

struct Bound {
    mtx_context* ctx;
    hipStream_t s;
    bool sync;
};

// A null stream means the context's own stream, and a synchronous call.
inline mtx_status bind(mtx_context* ctx, void* stream, Bound& b)
{
    b.ctx = ctx;
    b.s = stream ? static_cast<hipStream_t>(stream) : ctx->own;
    b.sync = stream == nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MTX_EDEVICE, "hipSetDevice failed");
    return MTX_OK;
}

inline mtx_status finish(const Bound& b)
{
    if (b.sync) MTX_HIP(b.ctx, hipStreamSynchronize(b.s));
    return MTX_OK;
}

// Records `*ev` (created on first use) on the bound stream: readbacks of an
// output wait on this event, i.e. on the stream that produced the output,
// not on the whole device.
inline mtx_status mark_done(const Bound& b, void** ev)
{
    if (!*ev) {
        hipEvent_t e = nullptr;
        MTX_HIP(b.ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        *ev = e;
    }
    MTX_HIP(b.ctx, hipEventRecord(static_cast<hipEvent_t>(*ev), b.s));
    return MTX_OK;
}

// Blocking copy of `bytes` from device memory once `ev` (the producer's
// completion event) has been reached: the context's own stream waits for the
// event, copies, and is synchronised.  Other streams keep running.
inline mtx_status read_after(mtx_context* ctx, void* ev, void* dst, const void* src, size_t bytes)
{
    MTX_HIP(ctx, hipSetDevice(ctx->device));
    if (ev) MTX_HIP(ctx, hipStreamWaitEvent(ctx->own, static_cast<hipEvent_t>(ev), 0));
    MTX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->own));
    MTX_HIP(ctx, hipStreamSynchronize(ctx->own));
    return MTX_OK;
}

inline void destroy_event(void** ev)
{
    if (*ev) hipEventDestroy(static_cast<hipEvent_t>(*ev));
    *ev = nullptr;
}

inline ActParams make_act(int dtype, float iso, int rule)
{
    ActParams a{};
    a.iso = iso;
    a.dtype = dtype;
    a.rule = rule;
    float r = std::roundf(iso);
    a.iso_r_valid = (r >= 0.0f && r < 4294967296.0f) ? 1 : 0;
    a.iso_r = a.iso_r_valid ? (uint32_t)r : 0u;
    return a;
}

inline uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Extraction workspace layout (offsets in bytes).
struct WsLayout {
    uint64_t counts, vbase, tbase, mbase, sbase, table, cases, starts, adj, totals, scan, bytes;
    uint32_t scan_tiles;
};

// Words of a block's meshlet start bits (one bit per cell).
inline uint64_t start_words(uint32_t B) { return ((uint64_t)B * B * B + 31) / 32; }

inline WsLayout ws_layout(uint64_t max_blocks, uint32_t B)
{
    WsLayout w{};
    uint64_t o = 0;
    w.counts = o; o = align256(o + 16 * max_blocks);   // {vertices, triangles, meshlets, meshlet vertices}
    w.vbase = o;  o = align256(o + 8 * max_blocks);
    w.tbase = o;  o = align256(o + 8 * max_blocks);
    w.mbase = o;  o = align256(o + 8 * max_blocks);
    w.sbase = o;  o = align256(o + 8 * max_blocks);
    w.table = o;  o = align256(o + corner_entry_bytes(B) * (uint64_t)B * B * B * max_blocks);
    w.cases = o;  o = align256(o + (uint64_t)B * B * B * max_blocks);
    w.starts = o; o = align256(o + 4 * start_words(B) * max_blocks);
    w.adj = o;    o = align256(o + 8ull * 4 * adjoint_workgroups(max_blocks));
    w.totals = o; o = align256(o + 8 * 8);
    uint64_t tiles = (max_blocks + scan_tile_items() - 1) / scan_tile_items();
    if (tiles == 0) tiles = 1;
    w.scan_tiles = (uint32_t)tiles;
    w.scan = o;   o = align256(o + 8 * (4 * tiles + 1));
    w.bytes = o;
    return w;
}

}  // namespace mtx
