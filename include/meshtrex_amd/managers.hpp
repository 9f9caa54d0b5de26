// meshtrex_amd C++ host API: the reference's manager functions and output
// structs, header-only over the C ABI in mtx.h.
//
//   computeMinMaxMip           src/minMaxManager.h:20-21
//   filterActiveBlocks         src/filteringManager.h:9-10
//   readActiveBlockCount       src/filteringManager.h:13
//   extractMeshletDescriptors  src/extractionManager.h:10-11
//   loadVolume                 src/volume.h:14-42
//   writeGPUExtractionToOBJ    src/extractionTestUtils.h:14-17
//
// Same names, argument order and output field names.  What stands in for the
// Vulkan types: MtxContext for VulkanContext, a stream (hipStream_t, passed as
// void*) for VkCommandBuffer externalCmd, MtxProfiler for GPUProfiler, device
// pointers for VkImage/VkBuffer.  A null stream means the call synchronises
// before returning (the reference's own-command-buffer case).  Outputs are
// move-only and free their device memory in their destructors; cleanup() is
// kept for source compatibility.  Errors throw std::runtime_error, as the
// reference's managers do (src/extractionManager.cpp:276,313-338).
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <regex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mtx.h"

namespace meshtrex {

using Stream = void*;  // hipStream_t
using VertexData = mtx_vertex;
using MeshletDescriptor = mtx_meshlet;

// PushConstants of src/common.h:29-34 (64 bytes, same layout).
struct PushConstants : mtx_push_constants {
    PushConstants() { std::memset(static_cast<mtx_push_constants*>(this), 0, sizeof(mtx_push_constants)); }
    PushConstants(const uint32_t dims[3], uint32_t blockDim, float iso) : PushConstants()
    {
        mtx_push_constants_init(this, dims, blockDim, iso);
    }
};
static_assert(sizeof(PushConstants) == 64, "PushConstants must stay 64 bytes");

inline void check(mtx_context* ctx, mtx_status st)
{
    if (st != MTX_OK)
        throw std::runtime_error(std::string(mtx_status_string(st)) + ": " + mtx_context_last_error(ctx));
}

class MtxContext {
public:
    explicit MtxContext(int device = 0)
    {
        mtx_status st = mtx_context_create(device, &h_);
        if (st != MTX_OK) throw std::runtime_error(std::string("mtx_context_create: ") + mtx_status_string(st));
    }
    ~MtxContext() { mtx_context_destroy(h_); }
    MtxContext(const MtxContext&) = delete;
    MtxContext& operator=(const MtxContext&) = delete;
    mtx_context* get() const { return h_; }

private:
    mtx_context* h_ = nullptr;
};

class MtxProfiler {
public:
    MtxProfiler()
    {
        if (mtx_profiler_create(&h_) != MTX_OK) throw std::runtime_error("mtx_profiler_create");
    }
    ~MtxProfiler() { mtx_profiler_destroy(h_); }
    MtxProfiler(const MtxProfiler&) = delete;
    MtxProfiler& operator=(const MtxProfiler&) = delete;
    void beginProfileRegion(Stream s, const char* name) { mtx_profiler_begin(h_, name, s); }
    void endProfileRegion(Stream s) { mtx_profiler_end(h_, s); }
    // (name, milliseconds) of every region, in begin order; blocks on the GPU.
    std::vector<std::pair<std::string, float>> results() const
    {
        std::vector<std::pair<std::string, float>> r;
        char name[128];
        float ms = 0.f;
        for (uint32_t i = 0; mtx_profiler_region(h_, i, name, sizeof name, &ms) == MTX_OK; ++i) r.emplace_back(name, ms);
        return r;
    }
    void reset() { mtx_profiler_reset(h_); }

private:
    mtx_profiler* h_ = nullptr;
};

// Volume of src/volume.h:8-12: dims, type name ("uint8" or "float32") and the
// voxels, x-fastest.  `device_data`, when set, is a device pointer used in
// place of `volume_data` (no upload).
struct Volume {
    uint32_t volume_dims[3] = { 0, 0, 0 };
    std::string volume_type = "uint8";
    std::vector<uint8_t> volume_data;
    const void* device_data = nullptr;
};

// Nothing is deferred: device memory is released in stream order.  Kept so
// that code written against the reference's outputs still compiles.
struct TemporaryResources {
    void cleanup() {}
};

// Buffer of src/buffer.h:3-8: `buffer` is the device pointer (the VkBuffer's
// place), `size` its size in bytes; `data` (a host mapping) is always null.
struct Buffer {
    void* buffer = nullptr;
    void* data = nullptr;
    size_t size = 0;
};

// Image of src/image.h:3-7: `image` and `imageView` are both the device
// pointer of the texels (a linear x-fastest array), `size` its bytes.
struct Image {
    const void* image = nullptr;
    const void* imageView = nullptr;
    size_t size = 0;
};

// MinMaxOutput of src/minMaxOutput.h:10-64.  volumeImage: the device volume;
// minMaxImage: the block min/max pyramid (uint32 key pairs, level 0 first);
// minMaxMipViews: each level's first pair (the mip views' place).
struct MinMaxOutput {
    Image volumeImage{};
    Image minMaxImage{};
    std::vector<const uint32_t*> minMaxMipViews;
    TemporaryResources tempResources;
    mtx_minmax_output c{};
    mtx_context* ctx = nullptr;

    MinMaxOutput() = default;
    MinMaxOutput(MinMaxOutput&& o) noexcept : c(o.c), ctx(o.ctx)
    {
        o.c = mtx_minmax_output{};
        o.ctx = nullptr;
        sync();
        o.sync();
    }
    MinMaxOutput& operator=(MinMaxOutput&& o) noexcept
    {
        if (this != &o) {
            cleanup();
            c = o.c;
            ctx = o.ctx;
            o.c = mtx_minmax_output{};
            o.ctx = nullptr;
            sync();
            o.sync();
        }
        return *this;
    }
    MinMaxOutput(const MinMaxOutput&) = delete;
    MinMaxOutput& operator=(const MinMaxOutput&) = delete;
    ~MinMaxOutput() { cleanup(); }
    void cleanup()
    {
        if (ctx) mtx_minmax_release(ctx, &c);
        ctx = nullptr;
        c = mtx_minmax_output{};
        sync();
    }
    // cleanup(device) of src/minMaxOutput.h:57 (the device argument is not needed)
    template <typename Device> void cleanup(Device) { cleanup(); }
    // the reference-named members from the C output
    void sync()
    {
        const size_t esz = c.dtype == MTX_F32 ? 4 : 1;
        volumeImage.image = volumeImage.imageView = c.volume;
        volumeImage.size = c.volume ? (size_t)c.dims[0] * c.dims[1] * c.dims[2] * esz : 0;
        minMaxImage.image = minMaxImage.imageView = c.minmax;
        minMaxImage.size = c.minmax ? 8 * (size_t)c.level_offset[c.num_levels] : 0;
        minMaxMipViews.clear();
        for (uint32_t l = 0; c.minmax && l < c.num_levels; ++l) minMaxMipViews.push_back(c.minmax + 2 * c.level_offset[l]);
    }
};

// FilteringOutput of src/filteringOutput.h:11-74.
struct FilteringOutput {
    Buffer compactedBlockIdBuffer{};   // uint32 active block IDs, ascending
    Buffer activeBlockCountBuffer{};   // one uint32, the count (stays on the device)
    uint32_t activeBlockCount = 0;     // host copy, filled by readActiveBlockCount
    TemporaryResources tempResources;
    mtx_filtering_output c{};
    mtx_context* ctx = nullptr;

    FilteringOutput() = default;
    FilteringOutput(FilteringOutput&& o) noexcept : activeBlockCount(o.activeBlockCount), c(o.c), ctx(o.ctx)
    {
        o.c = mtx_filtering_output{};
        o.ctx = nullptr;
        o.activeBlockCount = 0;
        sync();
        o.sync();
    }
    FilteringOutput& operator=(FilteringOutput&& o) noexcept
    {
        if (this != &o) {
            cleanup();
            c = o.c;
            ctx = o.ctx;
            activeBlockCount = o.activeBlockCount;
            o.c = mtx_filtering_output{};
            o.ctx = nullptr;
            o.activeBlockCount = 0;
            sync();
            o.sync();
        }
        return *this;
    }
    FilteringOutput(const FilteringOutput&) = delete;
    FilteringOutput& operator=(const FilteringOutput&) = delete;
    ~FilteringOutput() { cleanup(); }
    void cleanup()
    {
        if (ctx) mtx_filtering_release(ctx, &c);
        ctx = nullptr;
        c = mtx_filtering_output{};
        sync();
    }
    // cleanup(device) of src/filteringOutput.h:69
    template <typename Device> void cleanup(Device) { cleanup(); }
    void sync()
    {
        compactedBlockIdBuffer.buffer = c.compacted_block_ids;
        compactedBlockIdBuffer.size = c.compacted_block_ids ? 4 * (size_t)c.capacity : 0;
        activeBlockCountBuffer.buffer = c.active_block_count;
        activeBlockCountBuffer.size = c.active_block_count ? 4 : 0;
    }
};

// ExtractionOutput of src/extractionOutput.h:10-122.  vertexBuffer /
// indexBuffer: the shared mesh (each vertex once, PMB ownership);
// meshletDescriptorBuffer: meshlets over the renderer-layout mesh
// meshletVertexBuffer / meshletIndexBuffer (see mtx_extraction_output); the
// three count buffers are uint32 counters on the device.
struct ExtractionOutput {
    Buffer vertexBuffer{};
    Buffer vertexCountBuffer{};
    Buffer indexBuffer{};
    Buffer indexCountBuffer{};
    Buffer meshletDescriptorBuffer{};
    Buffer meshletDescriptorCountBuffer{};
    Buffer meshletVertexBuffer{};
    Buffer meshletIndexBuffer{};
    uint32_t meshletCount = 0;
    uint32_t vertexCount = 0;
    uint32_t indexCount = 0;
    TemporaryResources tempResources;
    mtx_extraction_output c{};
    mtx_context* ctx = nullptr;

    ExtractionOutput() = default;
    ExtractionOutput(ExtractionOutput&& o) noexcept
        : meshletCount(o.meshletCount), vertexCount(o.vertexCount), indexCount(o.indexCount), c(o.c), ctx(o.ctx)
    {
        o.c = mtx_extraction_output{};
        o.ctx = nullptr;
        o.meshletCount = o.vertexCount = o.indexCount = 0;
        sync();
        o.sync();
    }
    ExtractionOutput& operator=(ExtractionOutput&& o) noexcept
    {
        if (this != &o) {
            cleanup();
            c = o.c;
            ctx = o.ctx;
            meshletCount = o.meshletCount;
            vertexCount = o.vertexCount;
            indexCount = o.indexCount;
            o.c = mtx_extraction_output{};
            o.ctx = nullptr;
            o.meshletCount = o.vertexCount = o.indexCount = 0;
            sync();
            o.sync();
        }
        return *this;
    }
    ExtractionOutput(const ExtractionOutput&) = delete;
    ExtractionOutput& operator=(const ExtractionOutput&) = delete;
    ~ExtractionOutput() { cleanup(); }
    void cleanup()
    {
        if (ctx) mtx_extraction_release(ctx, &c);
        ctx = nullptr;
        c = mtx_extraction_output{};
        meshletCount = vertexCount = indexCount = 0;
        sync();
    }
    void sync()
    {
        auto set = [](Buffer& b, void* p, size_t bytes) {
            b.buffer = p;
            b.size = p ? bytes : 0;
        };
        char* counts = reinterpret_cast<char*>(c.counts);
        set(vertexBuffer, c.vertices, sizeof(mtx_vertex) * (size_t)c.vertex_capacity);
        set(indexBuffer, c.indices, 4 * (size_t)c.index_capacity);
        set(meshletDescriptorBuffer, c.meshlets, sizeof(mtx_meshlet) * (size_t)c.meshlet_capacity);
        set(meshletVertexBuffer, c.meshlet_vertices, sizeof(mtx_vertex) * (size_t)c.meshlet_vertex_capacity);
        set(meshletIndexBuffer, c.meshlet_indices, 4 * (size_t)c.index_capacity);
        set(vertexCountBuffer, counts, 4);
        set(indexCountBuffer, counts ? counts + 4 : nullptr, 4);
        set(meshletDescriptorCountBuffer, counts ? counts + 8 : nullptr, 4);
    }
};

inline MinMaxOutput computeMinMaxMip(MtxContext& context, const Volume& volume, PushConstants& pushConstants,
                                     Stream externalStream = nullptr, MtxProfiler* profiler = nullptr,
                                     const mtx_options* options = nullptr)
{
    mtx_volume v{};
    v.dims[0] = volume.volume_dims[0];
    v.dims[1] = volume.volume_dims[1];
    v.dims[2] = volume.volume_dims[2];
    if (volume.volume_type == "uint8") v.dtype = MTX_U8;
    else if (volume.volume_type == "float32") v.dtype = MTX_F32;
    else throw std::runtime_error("unsupported volume type " + volume.volume_type);
    v.data = volume.device_data ? volume.device_data : static_cast<const void*>(volume.volume_data.data());
    v.on_device = volume.device_data ? 1u : 0u;
    mtx_options o{};
    if (options) o = *options;
    else mtx_options_default(&o);
    const uint32_t build = o.build_pyramid;
    o.build_pyramid = 0;
    MinMaxOutput out;
    out.ctx = context.get();
    if (profiler) profiler->beginProfileRegion(externalStream, "MinMax_Leaf_Pass");
    mtx_status st = mtx_compute_min_max(context.get(), &v, &pushConstants, &o, externalStream, &out.c);
    if (profiler) profiler->endProfileRegion(externalStream);
    out.sync();
    check(context.get(), st);
    if (build && out.c.num_levels > 1) {
        if (profiler) profiler->beginProfileRegion(externalStream, "MinMax_Octree_Reduction");
        st = mtx_build_pyramid(context.get(), &out.c, externalStream);
        if (profiler) profiler->endProfileRegion(externalStream);
        check(context.get(), st);
    }
    return out;
}

inline FilteringOutput filterActiveBlocks(MtxContext& context, MinMaxOutput& minMaxOutput, PushConstants& pushConstants,
                                          Stream externalStream = nullptr, MtxProfiler* profiler = nullptr,
                                          const mtx_options* options = nullptr)
{
    FilteringOutput out;
    out.ctx = context.get();
    if (profiler) profiler->beginProfileRegion(externalStream, "Active_Block_Filtering");
    mtx_status st = mtx_filter_active_blocks(context.get(), &minMaxOutput.c, &pushConstants, options, externalStream,
                                             &out.c);
    if (profiler) profiler->endProfileRegion(externalStream);
    out.sync();
    check(context.get(), st);
    return out;
}

inline void readActiveBlockCount(MtxContext& context, FilteringOutput& filteringOutput)
{
    check(context.get(), mtx_read_active_count(context.get(), &filteringOutput.c));
    filteringOutput.activeBlockCount = filteringOutput.c.activeBlockCount;
}

// Unlike the reference (src/extractionManager.cpp:253-256), no host-side
// active count is needed: the count stays on the device.
inline ExtractionOutput extractMeshletDescriptors(MtxContext& context, MinMaxOutput& minMaxOutput,
                                                  FilteringOutput& filterOutput, PushConstants& pushConstants,
                                                  Stream externalStream = nullptr, MtxProfiler* profiler = nullptr,
                                                  const mtx_options* options = nullptr)
{
    ExtractionOutput out;
    out.ctx = context.get();
    if (profiler) profiler->beginProfileRegion(externalStream, "Mesh_Extraction");
    mtx_status st = mtx_extract(context.get(), &minMaxOutput.c, &filterOutput.c, &pushConstants, options,
                                externalStream, &out.c);
    if (profiler) profiler->endProfileRegion(externalStream);
    out.sync();
    check(context.get(), st);
    out.meshletCount = out.c.meshletCount;
    out.vertexCount = out.c.vertexCount;
    out.indexCount = out.c.indexCount;
    return out;
}

// src/volume.h:14-42: '<name>_<X>x<Y>x<Z>_<type>.raw'; uint8 and float32.
inline Volume loadVolume(const char* path)
{
    const std::string file = path;
    const std::regex re(R"((\w+)_(\d+)x(\d+)x(\d+)_(.+)\.raw)");
    std::smatch m;
    if (!std::regex_search(file, m, re) || m.size() != 6)
        throw std::runtime_error("Invalaid raw file naming scheme: expected '<name>_<X>x<Y>x<Z>_<data type>.raw' but '" +
                                 file + "' did not match");
    Volume v;
    v.volume_dims[0] = (uint32_t)std::stoul(m[2]);
    v.volume_dims[1] = (uint32_t)std::stoul(m[3]);
    v.volume_dims[2] = (uint32_t)std::stoul(m[4]);
    v.volume_type = m[5];
    size_t esz = 0;
    if (v.volume_type == "uint8") esz = 1;
    else if (v.volume_type == "float32") esz = 4;
    else throw std::runtime_error("unsupported raw volume type " + v.volume_type);
    const size_t bytes = (size_t)v.volume_dims[0] * v.volume_dims[1] * v.volume_dims[2] * esz;
    v.volume_data.resize(bytes);
    std::ifstream fin(file, std::ios::binary);
    if (!fin || !fin.read(reinterpret_cast<char*>(v.volume_data.data()), (std::streamsize)bytes))
        throw std::runtime_error("Failed to read volume data from " + file);
    return v;
}

inline void writeGPUExtractionToOBJ(MtxContext& context, ExtractionOutput& extractionResult, const char* filePath)
{
    check(context.get(), mtx_write_obj(context.get(), &extractionResult.c, filePath));
    extractionResult.vertexCount = extractionResult.c.vertexCount;
    extractionResult.indexCount = extractionResult.c.indexCount;
    extractionResult.meshletCount = extractionResult.c.meshletCount;
}

}  // namespace meshtrex
