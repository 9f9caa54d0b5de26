// Driver for the reference's own CPU oracle headers (TEST INFRASTRUCTURE).
// It is compiled against the headers where they lie in the reference tree
// (see oracle/Makefile) and writes their outputs as raw files:
//   ref_driver minmax <volume.raw> <pairs.out>   computeMinMaxFromFile (src/testMinMax.h:28-135)
//   ref_driver count  <pairs.in> <iso>           computeActiveBlockCountCPU (src/testMinMax.h:220-242)
//   ref_driver tables <tables.out>               MarchingCubes::triTable/edgeTable/numVerticesTable (src/mc_tables.h)
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
