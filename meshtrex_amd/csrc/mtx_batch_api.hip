// C-ABI of the batched form: planning (autotuner, sizing run, allocation),
// hipGraph capture of the launch sequence, and replay.
#include "mtx_host.h"

#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace mtx;

namespace {

// One dtype class of a batch: its volumes and units, their device descriptors
// and every intermediate buffer of the class's pipeline, in one allocation.
struct ClassPlan {
    int dtype = 0;
    uint32_t B = 0, M = 0;
    std::vector<uint32_t> vols;     // batch volume indices
    std::vector<uint32_t> units;    // batch unit indices
    std::vector<BatchVolume> hv;
    std::vector<uint32_t> hwg0;
    std::vector<BatchUnit> hu;
    std::vector<uint32_t> htile0;
    uint64_t mm_nwg = 0;
    uint32_t ntiles = 0;            // compaction tiles
    uint64_t nb_total = 0;          // blocks over all units: the active list's capacity
    uint64_t voxels = 0;
    WsLayout w{};

    char* slab = nullptr;
    BatchVolume* d_vols = nullptr;
    uint32_t* d_wg0 = nullptr;
    BatchUnit* d_units = nullptr;
    uint32_t* d_tile0 = nullptr;
    uint32_t* d_unit_global = nullptr;
    uint32_t* d_pairs = nullptr;
    uint32_t* d_maps = nullptr;
    uint32_t* d_item_unit = nullptr;
    uint32_t* d_unit_first = nullptr;
    uint32_t* d_count = nullptr;
    uint32_t* d_ids = nullptr;
    char* d_ws = nullptr;           // extraction workspace (ws_layout)
    uint64_t* d_fscan = nullptr;    // compaction tile counts (ntiles uint32)

    // outputs (class-relative capacities); the batch owns the memory
    mtx_vertex* vertices = nullptr;
    uint32_t* indices = nullptr;
    mtx_meshlet* meshlets = nullptr;
    uint64_t* edge_ids = nullptr;
    uint64_t vcap = 0, icap = 0, mcap = 0;
    uint64_t voff = 0, toff = 0, moff = 0, aoff = 0;
    mtx_batch_range* ranges = nullptr;
    int normals = 1;

    // planning-run results
    uint64_t tot[4] = { 0, 0, 0, 0 };
    uint32_t active = 0;
    float tuned_ms[3] = { 0, 0, 0 };
    float tuned_rows_ms[3] = { 0, 0, 0 };
    uint32_t schedule = 1, wg_per_cu = 0;   // 0: the launchers' default
    float tuned_sched_ms[6] = { 0, 0, 0, 0, 0, 0 };

    void release()
    {
        if (slab) (void)hipFree(slab);
        slab = nullptr;
    }
};

struct VolumeIn {
    mtx_volume v;
    std::vector<float> isos;
    uint32_t unit0;
};

}  // namespace

struct mtx_batch {
    mtx_context* ctx = nullptr;
    mtx_batch_options opt{};
    std::vector<VolumeIn> vols;
    uint32_t nunits = 0;
    std::vector<ClassPlan> classes;
    void* out_slab = nullptr;       // vertices, indices, meshlets, edge ids, active ids, ranges
    mtx_batch_output out{};
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> side;  // capture branches of the classes after the first
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Device layout and host descriptors of a class at block edge B, M y-blocks per
// min/max wave.  Allocates the class slab (not the outputs).
mtx_status class_build(mtx_context* ctx, const std::vector<VolumeIn>& vols, const mtx_batch_options& o,
                       ClassPlan& c, uint32_t B, uint32_t M)
{
    c.release();
    c.B = B;
    c.M = M;
    c.hv.clear();
    c.hu.clear();
    c.htile0.clear();
    const uint32_t nv = (uint32_t)c.vols.size();
    std::vector<uint64_t> pair_off(nv);
    uint64_t pairs = 0, nb_total = 0, tiles = 0, voxels = 0;
    for (uint32_t i = 0; i < nv; ++i) {
        const mtx_volume& v = vols[c.vols[i]].v;
        BatchVolume d{};
        d.data = v.data;
        d.X = v.dims[0]; d.Y = v.dims[1]; d.Z = v.dims[2];
        d.gx = (d.X + B - 1) / B; d.gy = (d.Y + B - 1) / B; d.gz = (d.Z + B - 1) / B;
        const uint64_t esz = c.dtype == kF32 ? 4 : 1;
        d.aligned = (d.X % 4u) == 0 && ((uintptr_t)v.data % (4 * esz)) == 0;
        pair_off[i] = pairs;
        const uint64_t nb = (uint64_t)d.gx * d.gy * d.gz;
        pairs += nb;
        voxels += (uint64_t)d.X * d.Y * d.Z;
        c.hv.push_back(d);
    }
    c.mm_nwg = plan_minmax_batch(c.hv.data(), nv, c.dtype, B, M);
    if (c.mm_nwg == 0) return fail(ctx, MTX_EINVAL, "batch min/max plan failed (B %u, M %u)", B, M);
    c.hwg0.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) c.hwg0[i] = c.hv[i].wg0;

    std::vector<uint64_t> map_off;
    for (uint32_t i = 0; i < nv; ++i) {
        const VolumeIn& vin = vols[c.vols[i]];
        const BatchVolume& d = c.hv[i];
        const uint64_t nb = (uint64_t)d.gx * d.gy * d.gz;
        for (float iso : vin.isos) {
            BatchUnit u{};
            u.data = d.data;
            u.X = d.X; u.Y = d.Y; u.Z = d.Z;
            u.gx = d.gx; u.gy = d.gy; u.gz = d.gz;
            u.nb = (uint32_t)nb;
            u.tile0 = (uint32_t)tiles;
            u.act = make_act(c.dtype, iso, (int)o.rule);
            tiles += batch_filter_tiles((uint32_t)nb);
            map_off.push_back(nb_total);
            nb_total += nb;
            c.hu.push_back(u);
            c.htile0.push_back(u.tile0);
        }
    }
    if (nb_total >= 0xFFFFFFFFull || tiles >= 0x7FFFFFFFull)
        return fail(ctx, MTX_EINVAL, "batch class too large (%llu blocks)", (unsigned long long)nb_total);
    c.nb_total = nb_total;
    c.ntiles = (uint32_t)tiles;
    c.voxels = voxels;
    const uint32_t nu = (uint32_t)c.hu.size();
    c.w = ws_layout(nb_total, B);

    // slab layout (nothing in it needs clearing between runs)
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) { uint64_t r = off; off = align256(off + bytes); return r; };
    const uint64_t o_vols = take(sizeof(BatchVolume) * nv);
    const uint64_t o_wg0 = take(4ull * nv);
    const uint64_t o_units = take(sizeof(BatchUnit) * nu);
    const uint64_t o_tile0 = take(4ull * nu);
    const uint64_t o_glob = take(4ull * nu);
    const uint64_t o_pairs = take(8 * pairs);
    const uint64_t o_maps = take(4 * nb_total);
    const uint64_t o_item = take(4 * nb_total);
    const uint64_t o_ids = take(4 * nb_total);
    const uint64_t o_first = take(4ull * (nu + 1));
    const uint64_t o_count = take(8);
    const uint64_t o_ws = take(c.w.bytes);                  // extraction workspace
    const uint64_t o_fscan = take(4 * tiles);               // compaction tile counts
    hipError_t e = hipMalloc(&c.slab, off);
    if (e != hipSuccess) {
        c.slab = nullptr;
        return fail(ctx, MTX_ENOMEM, "batch class buffers (%llu B)", (unsigned long long)off);
    }
    c.d_vols = reinterpret_cast<BatchVolume*>(c.slab + o_vols);
    c.d_wg0 = reinterpret_cast<uint32_t*>(c.slab + o_wg0);
    c.d_units = reinterpret_cast<BatchUnit*>(c.slab + o_units);
    c.d_tile0 = reinterpret_cast<uint32_t*>(c.slab + o_tile0);
    c.d_unit_global = reinterpret_cast<uint32_t*>(c.slab + o_glob);
    c.d_pairs = reinterpret_cast<uint32_t*>(c.slab + o_pairs);
    c.d_maps = reinterpret_cast<uint32_t*>(c.slab + o_maps);
    c.d_item_unit = reinterpret_cast<uint32_t*>(c.slab + o_item);
    c.d_ids = reinterpret_cast<uint32_t*>(c.slab + o_ids);
    c.d_unit_first = reinterpret_cast<uint32_t*>(c.slab + o_first);
    c.d_count = reinterpret_cast<uint32_t*>(c.slab + o_count);
    c.d_ws = c.slab + o_ws;
    c.d_fscan = reinterpret_cast<uint64_t*>(c.slab + o_fscan);

    for (uint32_t i = 0; i < nv; ++i) c.hv[i].pairs = c.d_pairs + 2 * pair_off[i];
    uint32_t k = 0;
    for (uint32_t i = 0; i < nv; ++i) {
        const VolumeIn& vin = vols[c.vols[i]];
        for (size_t j = 0; j < vin.isos.size(); ++j, ++k) {
            c.hu[k].pairs = c.hv[i].pairs;
            c.hu[k].map = c.d_maps + map_off[k];
        }
    }
    e = hipMemcpy(c.d_vols, c.hv.data(), sizeof(BatchVolume) * nv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.d_wg0, c.hwg0.data(), 4ull * nv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.d_units, c.hu.data(), sizeof(BatchUnit) * nu, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.d_tile0, c.htile0.data(), 4ull * nu, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.d_unit_global, c.units.data(), 4ull * nu, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        c.release();
        return fail(ctx, MTX_EDEVICE, "batch descriptors: %s", hipGetErrorString(e));
    }
    return MTX_OK;
}

// Re-plan the min/max launch for M y-blocks per wave (descriptors only).
mtx_status class_replan_rows(mtx_context* ctx, ClassPlan& c, uint32_t M)
{
    const uint32_t nv = (uint32_t)c.hv.size();
    c.mm_nwg = plan_minmax_batch(c.hv.data(), nv, c.dtype, c.B, M);
    if (c.mm_nwg == 0) return fail(ctx, MTX_EINVAL, "batch min/max plan failed (M %u)", M);
    c.M = M;
    for (uint32_t i = 0; i < nv; ++i) c.hwg0[i] = c.hv[i].wg0;
    MTX_HIP(ctx, hipMemcpy(c.d_vols, c.hv.data(), sizeof(BatchVolume) * nv, hipMemcpyHostToDevice));
    MTX_HIP(ctx, hipMemcpy(c.d_wg0, c.hwg0.data(), 4ull * nv, hipMemcpyHostToDevice));
    return MTX_OK;
}

hipError_t class_minmax(const ClassPlan& c, hipStream_t s)
{
    return launch_minmax_batch(c.d_vols, c.d_wg0, (uint32_t)c.hv.size(), c.dtype, c.B, c.M, c.mm_nwg, s);
}

// The class's whole pipeline on `s`: no allocation, no host synchronisation,
// no memset (capturable).
hipError_t class_enqueue(const ClassPlan& c, hipStream_t s)
{
    hipError_t e = class_minmax(c, s);
    if (e == hipSuccess)
        e = launch_filter_batch(c.d_units, c.d_tile0, (uint32_t)c.hu.size(), c.ntiles, c.d_ids, c.d_item_unit,
                                c.d_unit_first, c.d_count, TileSums{ c.d_fscan }, s);
    if (e != hipSuccess) return e;
    ExtractArgs a{};
    a.vol.dtype = c.dtype;
    a.grid.B = c.B;
    a.ids = c.d_ids;
    a.count = c.d_count;
    a.blk_counts = reinterpret_cast<uint32_t*>(c.d_ws + c.w.counts);
    a.blk_vbase = reinterpret_cast<uint64_t*>(c.d_ws + c.w.vbase);
    a.blk_tbase = reinterpret_cast<uint64_t*>(c.d_ws + c.w.tbase);
    a.blk_mbase = reinterpret_cast<uint64_t*>(c.d_ws + c.w.mbase);
    a.blk_sbase = reinterpret_cast<uint64_t*>(c.d_ws + c.w.sbase);
    a.corner_table = c.d_ws + c.w.table;
    a.cases = reinterpret_cast<uint8_t*>(c.d_ws + c.w.cases);
    a.totals = reinterpret_cast<uint64_t*>(c.d_ws + c.w.totals);
    a.vertices = reinterpret_cast<float*>(c.vertices);
    a.indices = c.indices;
    a.edge_ids = c.edge_ids;
    a.meshlets = reinterpret_cast<uint4*>(c.meshlets);
    a.vcap = c.vcap;
    a.icap = c.icap;
    a.mcap = c.mcap;
    a.normals = c.normals;
    a.max_blocks = (uint32_t)c.nb_total;
    a.units = c.d_units;
    a.item_unit = c.d_item_unit;
    a.unit_first = c.d_unit_first;
    a.balanced = c.schedule == 2 ? 1u : 0u;
    a.wg_per_cu = c.wg_per_cu;
    e = launch_extract_count(a, s);
    if (e == hipSuccess) e = launch_extract_scan(a, TileSums{ reinterpret_cast<uint64_t*>(c.d_ws + c.w.scan) }, s);
    if (e == hipSuccess) e = launch_extract_emit(a, s);
    if (e != hipSuccess || !c.ranges) return e;
    BatchRangeArgs r{};
    r.unit_first = c.d_unit_first;
    r.unit_global = c.d_unit_global;
    r.nunits = (uint32_t)c.hu.size();
    r.count = c.d_count;
    r.vbase = a.blk_vbase;
    r.tbase = a.blk_tbase;
    r.mbase = a.blk_mbase;
    r.totals = a.totals;
    r.voff = c.voff; r.toff = c.toff; r.moff = c.moff; r.aoff = c.aoff;
    r.vcap = c.vcap; r.icap = c.icap; r.mcap = c.mcap;
    r.block_dim = c.B;
    r.ranges = c.ranges;
    return launch_batch_ranges(r, s);
}

// Planning run: counts only (zero output capacity), then the totals.
mtx_status class_size(mtx_context* ctx, ClassPlan& c, hipStream_t s)
{
    c.vertices = nullptr; c.indices = nullptr; c.meshlets = nullptr; c.edge_ids = nullptr;
    c.vcap = c.icap = c.mcap = 0;
    c.ranges = nullptr;
    hipError_t e = class_enqueue(c, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c.tot, c.d_ws + c.w.totals, 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&c.active, c.d_count, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "batch planning run: %s", hipGetErrorString(e));
    if (c.tot[2] & 2) return fail(ctx, MTX_ESTATE, "batch planning run: inconsistent active list");
    return MTX_OK;
}

uint64_t with_headroom(uint64_t n, uint32_t pct) { return n + (n * pct + 99) / 100; }

// Outputs owned by the class alone (autotuner candidates).
struct OwnOutputs {
    void* p = nullptr;
    ~OwnOutputs() { if (p) (void)hipFree(p); }
};

mtx_status class_own_outputs(mtx_context* ctx, ClassPlan& c, OwnOutputs& own, bool edge_ids)
{
    const uint64_t V = c.tot[0], T = c.tot[1], Mt = c.tot[3];
    const uint64_t ov = 0, oi = align256(ov + 32 * V), om = align256(oi + 12 * T),
                   oe = align256(om + 16 * Mt), end = align256(oe + (edge_ids ? 8 * V : 0)) + 256;
    if (own.p) (void)hipFree(own.p);
    own.p = nullptr;
    MTX_HIP(ctx, hipMalloc(&own.p, end));
    char* b = static_cast<char*>(own.p);
    c.vertices = reinterpret_cast<mtx_vertex*>(b + ov);
    c.indices = reinterpret_cast<uint32_t*>(b + oi);
    c.meshlets = reinterpret_cast<mtx_meshlet*>(b + om);
    c.edge_ids = edge_ids ? reinterpret_cast<uint64_t*>(b + oe) : nullptr;
    c.vcap = V; c.icap = 3 * T; c.mcap = Mt;
    return MTX_OK;
}

// Median-free timing: the minimum of `iters` runs after one warm-up.
template <typename F>
mtx_status time_min(mtx_context* ctx, hipStream_t s, uint32_t iters, F&& body, float& ms)
{
    hipEvent_t e0, e1;
    MTX_HIP(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, MTX_EDEVICE, "hipEventCreate");
    }
    float best = 1e30f;
    hipError_t e = body();
    for (uint32_t i = 0; e == hipSuccess && i < iters; ++i) {
        e = hipEventRecord(e0, s);
        if (e == hipSuccess) e = body();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess && t < best) best = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "autotuner run: %s", hipGetErrorString(e));
    ms = best;
    return MTX_OK;
}

// Candidate (schedule, workgroups per CU) pairs of the extraction passes.
constexpr uint32_t kSched[6][2] = { { 1, 32 }, { 1, 64 }, { 1, 128 }, { 2, 8 }, { 2, 16 }, { 2, 32 } };

// Autotuner: block edge by the class's whole pipeline time, then the min/max
// wave shape by the min/max time, then the extraction passes' work assignment
// and grid by the whole pipeline again, each the fastest of the timed
// candidates (a fixed option skips its step).
mtx_status class_tune(mtx_context* ctx, const std::vector<VolumeIn>& vols, const mtx_batch_options& o,
                      ClassPlan& c, hipStream_t s)
{
    static const uint32_t Bs[3] = { 4, 8, 16 };
    static const uint32_t Ms[3] = { 1, 2, 4 };
    const uint32_t iters = o.tune_iters ? o.tune_iters : 3;
    uint32_t bestB = o.block_dim;
    if (!bestB) {
        float best = 1e30f;
        for (int k = 0; k < 3; ++k) {
            if (mtx_status st = class_build(ctx, vols, o, c, Bs[k], 2)) return st;
            if (mtx_status st = class_size(ctx, c, s)) return st;
            OwnOutputs own;
            if (mtx_status st = class_own_outputs(ctx, c, own, o.edge_ids != 0)) return st;
            float ms = 0;
            if (mtx_status st = time_min(ctx, s, iters, [&] { return class_enqueue(c, s); }, ms)) return st;
            c.tuned_ms[k] = ms;
            if (ms < best) { best = ms; bestB = Bs[k]; }
            c.vertices = nullptr; c.indices = nullptr; c.meshlets = nullptr; c.edge_ids = nullptr;
        }
    }
    if (mtx_status st = class_build(ctx, vols, o, c, bestB, o.minmax_rows ? o.minmax_rows : 2)) return st;
    if (!o.minmax_rows) {
        float best = 1e30f;
        uint32_t bestM = 2;
        for (int k = 0; k < 3; ++k) {
            if (mtx_status st = class_replan_rows(ctx, c, Ms[k])) return st;
            float ms = 0;
            if (mtx_status st = time_min(ctx, s, iters, [&] { return class_minmax(c, s); }, ms)) return st;
            c.tuned_rows_ms[k] = ms;
            if (ms < best) { best = ms; bestM = Ms[k]; }
        }
        if (mtx_status st = class_replan_rows(ctx, c, bestM)) return st;
    }
    if (mtx_status st = class_size(ctx, c, s)) return st;
    // work assignment and grid of the extraction passes, by the whole pipeline
    if (!o.schedule || !o.wg_per_cu) {
        OwnOutputs own;
        if (mtx_status st = class_own_outputs(ctx, c, own, o.edge_ids != 0)) return st;
        float best = 1e30f;
        uint32_t bs = o.schedule ? o.schedule : 1, bw = o.wg_per_cu;
        for (int k = 0; k < 6; ++k) {
            if (o.schedule && o.schedule != kSched[k][0]) continue;
            if (o.wg_per_cu && o.wg_per_cu != kSched[k][1]) continue;
            c.schedule = kSched[k][0];
            c.wg_per_cu = kSched[k][1];
            float ms = 0;
            if (mtx_status st = time_min(ctx, s, iters, [&] { return class_enqueue(c, s); }, ms)) return st;
            c.tuned_sched_ms[k] = ms;
            if (ms < best) { best = ms; bs = c.schedule; bw = c.wg_per_cu; }
        }
        c.schedule = bs;
        c.wg_per_cu = bw;
        c.vertices = nullptr; c.indices = nullptr; c.meshlets = nullptr; c.edge_ids = nullptr;
    } else {
        c.schedule = o.schedule;
        c.wg_per_cu = o.wg_per_cu;
    }
    return MTX_OK;
}

void batch_free(mtx_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : b->events) (void)hipEventDestroy(ev);
    b->side.clear();
    b->events.clear();
    for (ClassPlan& c : b->classes) c.release();
    if (b->out_slab) (void)hipFree(b->out_slab);
    b->exec = nullptr;
    b->graph = nullptr;
    b->cap = nullptr;
    b->out_slab = nullptr;
}

hipError_t batch_enqueue(mtx_batch* b, hipStream_t s)
{
    for (const ClassPlan& c : b->classes) {
        hipError_t e = class_enqueue(c, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Capture: the classes are independent, so each after the first is recorded
// on a branch forked from the capture stream and joined at the end; the graph
// then lets them run concurrently.
hipError_t batch_capture(mtx_batch* b)
{
    const size_t nc = b->classes.size();
    hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    for (size_t k = 1; e == hipSuccess && k < nc; ++k) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) b->side.push_back(st);
    }
    for (size_t k = 0; e == hipSuccess && k < nc; ++k) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->events.push_back(ev);
    }
    if (e != hipSuccess) return e;
    e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t e2 = hipEventRecord(b->events[0], b->cap);  // fork point
    for (size_t k = 1; e2 == hipSuccess && k < nc; ++k) {
        e2 = hipStreamWaitEvent(b->side[k - 1], b->events[0], 0);
        if (e2 == hipSuccess) e2 = class_enqueue(b->classes[k], b->side[k - 1]);
        if (e2 == hipSuccess) e2 = hipEventRecord(b->events[k], b->side[k - 1]);
    }
    if (e2 == hipSuccess) e2 = class_enqueue(b->classes[0], b->cap);
    for (size_t k = 1; e2 == hipSuccess && k < nc; ++k) e2 = hipStreamWaitEvent(b->cap, b->events[k], 0);
    e = hipStreamEndCapture(b->cap, &b->graph);
    if (e2 != hipSuccess) return e2;
    if (e != hipSuccess) return e;
    return hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
}

}  // namespace

extern "C" {

void mtx_batch_options_default(mtx_batch_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->rule = MTX_RULE_CPU;
    o->normals = 1;
    o->graph = 1;
    o->tune_iters = 3;
}

mtx_status mtx_batch_create(mtx_context* ctx, const mtx_batch_volume* volumes, uint32_t num_volumes,
                            const mtx_batch_options* opt, mtx_batch** out)
{
    if (!ctx) return MTX_EINVAL;
    if (!out) return fail(ctx, MTX_EINVAL, "null argument");
    *out = nullptr;
    if (!volumes || num_volumes == 0) return fail(ctx, MTX_EINVAL, "empty batch");
    mtx_batch_options o;
    if (opt) o = *opt;
    else mtx_batch_options_default(&o);
    if (o.block_dim && !batch_supported_block(o.block_dim))
        return fail(ctx, MTX_EINVAL, "batch block edge must be 4, 8 or 16 (got %u)", o.block_dim);
    if (o.minmax_rows != 0 && o.minmax_rows != 1 && o.minmax_rows != 2 && o.minmax_rows != 4)
        return fail(ctx, MTX_EINVAL, "minmax_rows must be 0, 1, 2 or 4");
    if (o.rule > MTX_RULE_GPU) return fail(ctx, MTX_EINVAL, "unknown rule %u", o.rule);
    if (o.schedule > 2) return fail(ctx, MTX_EINVAL, "schedule must be 0, 1 or 2");
    if (o.wg_per_cu > 1024) return fail(ctx, MTX_EINVAL, "wg_per_cu must be at most 1024");

    mtx_batch* b = new (std::nothrow) mtx_batch;
    if (!b) return fail(ctx, MTX_ENOMEM, "batch");
    b->ctx = ctx;
    b->opt = o;
    uint32_t nunits = 0;
    for (uint32_t i = 0; i < num_volumes; ++i) {
        const mtx_batch_volume& bv = volumes[i];
        const mtx_volume& v = bv.volume;
        const char* bad = nullptr;
        if (!v.data) bad = "null data";
        else if (!v.on_device) bad = "not device-resident";
        else if (v.dtype != MTX_U8 && v.dtype != MTX_F32) bad = "unsupported dtype";
        else if (!v.dims[0] || !v.dims[1] || !v.dims[2]) bad = "empty";
        else if ((uint64_t)v.dims[0] * v.dims[1] > 0xFFFFFFFFull ||
                 (uint64_t)v.dims[0] * v.dims[1] * v.dims[2] > (1ull << 40)) bad = "too large";
        else if (!bv.isovalues || bv.num_isovalues == 0) bad = "no isovalues";
        if (bad) {
            delete b;
            return fail(ctx, MTX_EINVAL, "batch volume %u: %s", i, bad);
        }
        VolumeIn vin;
        vin.v = v;
        vin.isos.assign(bv.isovalues, bv.isovalues + bv.num_isovalues);
        vin.unit0 = nunits;
        nunits += bv.num_isovalues;
        b->vols.push_back(vin);
    }
    b->nunits = nunits;

    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, nullptr, bd)) {
        delete b;
        return st;
    }
    // one class per dtype present, volumes and units in batch order
    for (int dt = 0; dt < 2; ++dt) {
        ClassPlan c;
        c.dtype = dt == 0 ? kU8 : kF32;
        c.normals = o.normals ? 1 : 0;
        for (uint32_t i = 0; i < num_volumes; ++i) {
            if ((int)b->vols[i].v.dtype != c.dtype) continue;
            c.vols.push_back(i);
            for (uint32_t j = 0; j < b->vols[i].isos.size(); ++j) c.units.push_back(b->vols[i].unit0 + j);
        }
        if (!c.vols.empty()) b->classes.push_back(std::move(c));
    }
    auto bail = [&](mtx_status st) {
        batch_free(b);
        delete b;
        return st;
    };
    for (ClassPlan& c : b->classes)
        if (mtx_status st = class_tune(ctx, b->vols, o, c, bd.s)) return bail(st);

    // batch outputs: every class's planning-run counts plus headroom
    uint64_t V = 0, T = 0, Mt = 0, A = 0;
    for (ClassPlan& c : b->classes) {
        c.voff = V; c.toff = T; c.moff = Mt; c.aoff = A;
        c.vcap = with_headroom(c.tot[0], o.headroom_pct);
        c.icap = 3 * with_headroom(c.tot[1], o.headroom_pct);
        c.mcap = with_headroom(c.tot[3], o.headroom_pct);
        V += c.vcap; T += c.icap / 3; Mt += c.mcap; A += c.nb_total;
    }
    if (V > 0xFFFFFFFFull || 3 * T > 0xFFFFFFFFull || A > 0xFFFFFFFFull) {
        std::string d;
        for (const ClassPlan& c : b->classes) {
            char buf[160];
            snprintf(buf, sizeof buf, " [dtype %d B %u: %u active blocks, %llu vertices, %llu triangles]", c.dtype,
                     c.B, c.active, (unsigned long long)c.tot[0], (unsigned long long)c.tot[1]);
            d += buf;
        }
        return bail(fail(ctx, MTX_ECAPACITY, "batch outputs exceed 32-bit offsets:%s", d.c_str()));
    }
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) { uint64_t r = off; off = align256(off + (bytes ? bytes : 1)); return r; };
    const uint64_t o_v = take(32 * V), o_i = take(12 * T), o_m = take(16 * Mt);
    const uint64_t o_e = o.edge_ids ? take(8 * V) : 0, o_a = take(4 * A), o_r = take(sizeof(mtx_batch_range) * nunits);
    if (hipMalloc(&b->out_slab, off) != hipSuccess) {
        b->out_slab = nullptr;
        return bail(fail(ctx, MTX_ENOMEM, "batch outputs (%llu B)", (unsigned long long)off));
    }
    char* base = static_cast<char*>(b->out_slab);
    b->out.vertices = reinterpret_cast<mtx_vertex*>(base + o_v);
    b->out.indices = reinterpret_cast<uint32_t*>(base + o_i);
    b->out.meshlets = reinterpret_cast<mtx_meshlet*>(base + o_m);
    b->out.vertex_edge_ids = o.edge_ids ? reinterpret_cast<uint64_t*>(base + o_e) : nullptr;
    b->out.active_block_ids = reinterpret_cast<uint32_t*>(base + o_a);
    b->out.ranges = reinterpret_cast<mtx_batch_range*>(base + o_r);
    b->out.vertex_capacity = V;
    b->out.index_capacity = 3 * T;
    b->out.meshlet_capacity = Mt;
    b->out.num_units = nunits;
    b->out.num_classes = (uint32_t)b->classes.size();
    for (ClassPlan& c : b->classes) {
        c.vertices = b->out.vertices + c.voff;
        c.indices = b->out.indices + 3 * c.toff;
        c.meshlets = b->out.meshlets + c.moff;
        c.edge_ids = b->out.vertex_edge_ids ? b->out.vertex_edge_ids + c.voff : nullptr;
        c.ranges = b->out.ranges;
        // the class's active list goes to the batch's id buffer from now on
        c.d_ids = b->out.active_block_ids + c.aoff;
    }
    if (o.graph) {
        hipError_t e = batch_capture(b);
        if (e != hipSuccess) return bail(fail(ctx, MTX_EDEVICE, "batch graph: %s", hipGetErrorString(e)));
    }
    *out = b;
    return MTX_OK;
}

mtx_status mtx_batch_run(mtx_batch* b, void* stream)
{
    if (!b) return MTX_EINVAL;
    mtx_context* ctx = b->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, stream, bd)) return st;
    hipError_t e = b->exec ? hipGraphLaunch(b->exec, bd.s) : batch_enqueue(b, bd.s);
    if (e != hipSuccess) return fail(ctx, MTX_EDEVICE, "batch run: %s", hipGetErrorString(e));
    return finish(bd);
}

mtx_status mtx_batch_output_get(const mtx_batch* b, mtx_batch_output* out)
{
    if (!b || !out) return MTX_EINVAL;
    *out = b->out;
    return MTX_OK;
}

mtx_status mtx_batch_read_ranges(mtx_batch* b, mtx_batch_range* host, uint32_t n)
{
    if (!b) return MTX_EINVAL;
    mtx_context* ctx = b->ctx;
    if (!host || n > b->nunits) return fail(ctx, MTX_EINVAL, "ranges: bad buffer or count %u", n);
    std::lock_guard<std::mutex> lk(ctx->mu);
    Bound bd;
    if (mtx_status st = bind(ctx, nullptr, bd)) return st;
    MTX_HIP(ctx, hipStreamSynchronize(bd.s));
    MTX_HIP(ctx, hipMemcpy(host, b->out.ranges, sizeof(mtx_batch_range) * n, hipMemcpyDeviceToHost));
    uint32_t flags = 0;
    for (uint32_t i = 0; i < n; ++i) flags |= host[i].flags;
    if (flags & 2) return fail(ctx, MTX_ESTATE, "batch: inconsistent active list");
    if (flags & 1) return fail(ctx, MTX_ECAPACITY, "batch: outputs exceeded the planned capacity");
    return MTX_OK;
}

mtx_status mtx_batch_class_info_get(const mtx_batch* b, uint32_t index, mtx_batch_class_info* out)
{
    if (!b || !out || index >= b->classes.size()) return MTX_EINVAL;
    const ClassPlan& c = b->classes[index];
    std::memset(out, 0, sizeof *out);
    out->dtype = c.dtype == kF32 ? MTX_F32 : MTX_U8;
    out->block_dim = c.B;
    out->minmax_rows = c.M;
    out->num_volumes = (uint32_t)c.vols.size();
    out->num_units = (uint32_t)c.hu.size();
    out->active_blocks = c.active;
    out->blocks = c.nb_total;
    out->voxels = c.voxels;
    out->vertices = c.tot[0];
    out->triangles = c.tot[1];
    for (int k = 0; k < 3; ++k) {
        out->tuned_ms[k] = c.tuned_ms[k];
        out->tuned_rows_ms[k] = c.tuned_rows_ms[k];
    }
    out->schedule = c.schedule;
    out->wg_per_cu = c.wg_per_cu;
    for (int k = 0; k < 6; ++k) out->tuned_sched_ms[k] = c.tuned_sched_ms[k];
    return MTX_OK;
}

void mtx_batch_destroy(mtx_batch* b)
{
    if (!b) return;
    mtx_context* ctx = b->ctx;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->own);
        (void)hipDeviceSynchronize();
        batch_free(b);
    }
    delete b;
}

}  // extern "C"
