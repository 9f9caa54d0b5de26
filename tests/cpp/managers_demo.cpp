// Drives the C++ host API the way the reference's profiled main path does
// (src/meshtrex.cpp:96-156): load a raw volume, min-max, filter, read the
// active count, extract, write an OBJ; prints the counts and profiler regions
// as one JSON line.
//   managers_demo <name_XxYxZ_type.raw> <isovalue> <block edge> <out.obj>
#include <cstdio>
#include <cstdlib>

#include "meshtrex_amd/managers.hpp"

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
        MinMaxOutput minMaxOutput = computeMinMaxMip(context, volume, pushConstants, nullptr, &profiler);
        FilteringOutput filteringResult = filterActiveBlocks(context, minMaxOutput, pushConstants, nullptr, &profiler);
        readActiveBlockCount(context, filteringResult);
        ExtractionOutput extraction =
            extractMeshletDescriptors(context, minMaxOutput, filteringResult, pushConstants, nullptr, &profiler);
        writeGPUExtractionToOBJ(context, extraction, argv[4]);
        std::printf("{\"activeBlockCount\": %u, \"vertexCount\": %u, \"indexCount\": %u, \"meshletCount\": %u, "
                    "\"levels\": %u, \"regions\": {",
                    filteringResult.activeBlockCount, extraction.vertexCount, extraction.indexCount,
                    extraction.meshletCount, minMaxOutput.c.num_levels);
        bool first = true;
        for (const auto& r : profiler.results()) {
            std::printf("%s\"%s\": %.4f", first ? "" : ", ", r.first.c_str(), r.second);
            first = false;
        }
        std::printf("}}\n");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
