// Driver for the reference's own CPU oracle headers (TEST INFRASTRUCTURE).
// It is compiled against the headers where they lie in the reference tree
// (see oracle/Makefile) and writes their outputs as raw files:
//   ref_driver minmax <volume.raw> <pairs.out>   computeMinMaxFromFile (src/testMinMax.h:28-135)
//   ref_driver count  <pairs.in> <iso>           computeActiveBlockCountCPU (src/testMinMax.h:220-242)
//   ref_driver tables <tables.out>               MarchingCubes::triTable/edgeTable/numVerticesTable (src/mc_tables.h)
//   ref_driver time   <volume.raw> <iso> <reps>  median wall time of computeMinMaxFromFile followed by
//                                                computeActiveBlockCountCPU (the reference's CPU path as a
//                                                whole, file read included; its console output discarded)
#include <chrono>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "testMinMax.h"
#include "mc_tables.h"

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    std::string mode = argv[1];
    if (mode == "minmax" && argc == 4) {
        std::vector<MinMaxResult> r = computeMinMaxFromFile(argv[2]);
        FILE* f = std::fopen(argv[3], "wb");
        if (!f) return 1;
        std::fwrite(r.data(), sizeof(MinMaxResult), r.size(), f);
        std::fclose(f);
        return 0;
    }
    if (mode == "count" && argc == 4) {
        FILE* f = std::fopen(argv[2], "rb");
        if (!f) return 1;
        std::vector<MinMaxResult> r;
        MinMaxResult m;
        while (std::fread(&m, sizeof(m), 1, f) == 1) r.push_back(m);
        std::fclose(f);
        uint32_t n = computeActiveBlockCountCPU(r, std::strtof(argv[3], nullptr));
        std::printf("ACTIVE %u\n", n);
        return 0;
    }
    if (mode == "time" && argc == 5) {
        const float iso = std::strtof(argv[3], nullptr);
        const int reps = std::atoi(argv[4]);
        if (reps < 1) return 2;
        std::vector<double> secs;
        uint32_t n = 0;
        std::streambuf* out = std::cout.rdbuf();
        std::cout.rdbuf(nullptr);
        for (int k = 0; k < reps; ++k) {
            auto t0 = std::chrono::steady_clock::now();
            std::vector<MinMaxResult> r = computeMinMaxFromFile(argv[2]);
            n = computeActiveBlockCountCPU(r, iso);
            secs.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        std::cout.clear();
        std::cout.rdbuf(out);
        std::sort(secs.begin(), secs.end());
        std::printf("SECONDS %.6f ACTIVE %u REPS %d\n", secs[secs.size() / 2], n, reps);
        return 0;
    }
    if (mode == "tables") {
        FILE* f = std::fopen(argv[2], "wb");
        if (!f) return 1;
        std::fwrite(MarchingCubes::triTable, sizeof(MarchingCubes::triTable), 1, f);
        std::fwrite(MarchingCubes::edgeTable, sizeof(MarchingCubes::edgeTable), 1, f);
        std::fwrite(MarchingCubes::numVerticesTable, sizeof(MarchingCubes::numVerticesTable), 1, f);
        std::fclose(f);
        return 0;
    }
    return 2;
}
