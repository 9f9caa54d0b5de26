// Drives the C++ host API the way the reference's profiled main path does
// (src/meshtrex.cpp:96-156): load a raw volume, min-max and filter on an
// external stream (the reference's external command buffer), read the active
// count, extract, write an OBJ.  It reads the output structs' members by the
// reference's names (src/minMaxOutput.h:11-13, src/filteringOutput.h:12-15,
// src/extractionOutput.h:11-22) and prints the counts, the counters read back
// through those members and the profiler regions as one JSON line.
//   managers_demo <name_XxYxZ_type.raw> <isovalue> <block edge> <out.obj>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "meshtrex_amd/managers.hpp"

namespace {

uint32_t read_u32(const void* device_ptr)
{
    uint32_t v = 0xFFFFFFFFu;
    if (device_ptr && mtx_memcpy_d2h(&v, device_ptr, 4) != MTX_OK) throw std::runtime_error("readback failed");
    return v;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s volume.raw isovalue block_edge out.obj\n", argv[0]);
        return 2;
    }
    try {
        using namespace meshtrex;
        MtxContext context(0);
        Volume volume = loadVolume(argv[1]);
        PushConstants pushConstants(volume.volume_dims, (uint32_t)std::atoi(argv[3]), (float)std::atof(argv[2]));
        MtxProfiler profiler;
        Stream cmd = nullptr;
        if (mtx_stream_create(&cmd) != MTX_OK) throw std::runtime_error("stream");
        MinMaxOutput minMaxOutput = computeMinMaxMip(context, volume, pushConstants, cmd, &profiler);
        FilteringOutput filteringResult = filterActiveBlocks(context, minMaxOutput, pushConstants, cmd, &profiler);
        mtx_stream_synchronize(cmd);
        readActiveBlockCount(context, filteringResult);
        minMaxOutput.tempResources.cleanup();
        filteringResult.tempResources.cleanup();
        ExtractionOutput extraction =
            extractMeshletDescriptors(context, minMaxOutput, filteringResult, pushConstants, cmd, &profiler);
        mtx_stream_synchronize(cmd);
        extraction.tempResources.cleanup();
        writeGPUExtractionToOBJ(context, extraction, argv[4]);
        // the reference-named members, as the reference's code uses them
        const uint32_t countViaBuffer = read_u32(filteringResult.activeBlockCountBuffer.buffer);
        const uint32_t firstId = filteringResult.activeBlockCount ? read_u32(filteringResult.compactedBlockIdBuffer.buffer)
                                                                  : 0xFFFFFFFFu;
        // the renderer layout, under the renderer's rule (src/shaders/render.task.glsl:61-63,125-128):
        // triangles with an index outside their meshlet's vertex range would be dropped
        std::vector<mtx_meshlet> desc(extraction.meshletCount);
        std::vector<uint32_t> mind(extraction.indexCount);
        if (!desc.empty() && mtx_memcpy_d2h(desc.data(), extraction.meshletDescriptorBuffer.buffer,
                                            desc.size() * sizeof(mtx_meshlet)) != MTX_OK)
            throw std::runtime_error("readback failed");
        if (!mind.empty() && mtx_memcpy_d2h(mind.data(), extraction.meshletIndexBuffer.buffer, 4 * mind.size()) != MTX_OK)
            throw std::runtime_error("readback failed");
        uint64_t dropped = 0, covered = 0, overfull = 0;
        for (const mtx_meshlet& m : desc) {
            if (m.vertexCount > 64 || m.primitiveCount > 126) ++overfull;
            for (uint32_t t = 0; t < m.primitiveCount; ++t) {
                bool ok = true;
                for (uint32_t k = 0; k < 3; ++k) {
                    const uint32_t i = mind[m.indexOffset + 3 * t + k];
                    ok = ok && i >= m.vertexOffset && i < m.vertexOffset + m.vertexCount;
                }
                dropped += ok ? 0 : 1;
                ++covered;
            }
        }
        std::printf("{\"activeBlockCount\": %u, \"vertexCount\": %u, \"indexCount\": %u, \"meshletCount\": %u, "
                    "\"levels\": %zu, \"activeBlockCountBuffer\": %u, \"firstActiveBlockId\": %u, "
                    "\"compactedBlockIdBuffer_size\": %zu, \"volumeImage_size\": %zu, \"minMaxImage_size\": %zu, "
                    "\"vertexCountBuffer\": %u, \"indexCountBuffer\": %u, \"meshletDescriptorCountBuffer\": %u, "
                    "\"vertexBuffer_size\": %zu, \"indexBuffer_size\": %zu, \"meshletDescriptorBuffer_size\": %zu, "
                    "\"meshletVertexBuffer_size\": %zu, \"meshletIndexBuffer_size\": %zu, "
                    "\"rendererTriangles\": %llu, \"rendererDropped\": %llu, \"overfullMeshlets\": %llu, "
                    "\"regions\": {",
                    filteringResult.activeBlockCount, extraction.vertexCount, extraction.indexCount,
                    extraction.meshletCount, minMaxOutput.minMaxMipViews.size(), countViaBuffer, firstId,
                    filteringResult.compactedBlockIdBuffer.size, minMaxOutput.volumeImage.size,
                    minMaxOutput.minMaxImage.size, read_u32(extraction.vertexCountBuffer.buffer),
                    read_u32(extraction.indexCountBuffer.buffer),
                    read_u32(extraction.meshletDescriptorCountBuffer.buffer), extraction.vertexBuffer.size,
                    extraction.indexBuffer.size, extraction.meshletDescriptorBuffer.size,
                    extraction.meshletVertexBuffer.size, extraction.meshletIndexBuffer.size,
                    (unsigned long long)covered, (unsigned long long)dropped, (unsigned long long)overfull);
        bool first = true;
        for (const auto& r : profiler.results()) {
            std::printf("%s\"%s\": %.4f", first ? "" : ", ", r.first.c_str(), r.second);
            first = false;
        }
        std::printf("}}\n");
        extraction.cleanup();
        minMaxOutput.cleanup(nullptr);
        filteringResult.cleanup(nullptr);
        mtx_stream_destroy(cmd);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
