// Internal interface between the C-ABI host layer (mtx_api.hip) and the
// kernel launchers (mtx_kernels.hip).  Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mtx {

enum Dtype : int { kU8 = 0, kF32 = 1 };
enum Rule : int { kRuleCpu = 0, kRuleGpu = 1 };

// Activity test parameters, prepared once on the host.
struct ActParams {
    float iso;        // isovalue
    uint32_t iso_r;   // round(iso) for the u8 CPU rule
    int iso_r_valid;  // round(iso) representable as uint32
    int dtype;
    int rule;
};

struct VolDesc {
    const void* data;
    uint32_t X, Y, Z;
    int dtype;
};

struct Grid {
    uint32_t B;
    uint32_t gx, gy, gz;
    uint64_t nb() const { return (uint64_t)gx * gy * gz; }
};

// Scratch of the two-launch scans: per-tile sums written by the first launch
// and read by the second (no clearing needed).
struct TileSums {
    uint64_t* sums;
};

// ---- launchers (all asynchronous on `s`) ----------------------------------

// Block min/max, B=8, overlap 1, coalesced wave64 kernel.  Optionally writes
// the activity bit mask (one 32-bit word per 32 x-blocks) for `act`.
// rows: y-blocks per wavefront (1, 2, 4) or 0 to choose from the shape.
// Covers z-blocks [bzb, bze); their voxel planes [8 bzb, min(8 bze, Z - 1)]
// must be resident.
hipError_t launch_minmax_b8(const VolDesc& v, const Grid& g, uint32_t bzb, uint32_t bze, uint32_t* pairs,
                            uint32_t* mask, const ActParams* act, uint32_t rows, hipStream_t s);

// Block min/max for any B, overlap 0 or 1 (one wave per block).
hipError_t launch_minmax_generic(const VolDesc& v, const Grid& g, int overlap, uint32_t* pairs,
                                 hipStream_t s);

// Pyramid levels 1..L-1 from level 0.  ext/off describe every level.
hipError_t launch_pyramid(uint32_t* pairs, uint32_t levels, const uint32_t (*ext)[3],
                          const uint64_t* off, hipStream_t s);

// Tiles of a compaction of nb blocks (one uint32 of TileSums each).
uint32_t filter_tiles(uint64_t nb, bool from_mask, uint32_t words_per_row, const Grid& g);

// Ordered compaction of active block IDs from the fused bit mask.  Also writes
// the dense map block ID -> active index (0xFFFFFFFF for inactive blocks).
hipError_t launch_compact_mask(const uint32_t* mask, const Grid& g, uint32_t words_per_row,
                               uint32_t* ids, uint32_t* count, uint32_t* map, TileSums ts,
                               hipStream_t s);

// Ordered compaction of active block IDs from leaf min/max pairs (+ dense map).
// coarse: pyramid level 2 (4^3-block ancestors) of the same pairs, with its
// extent, to skip blocks under inactive ancestors; or nullptr.
hipError_t launch_filter_pairs(const uint32_t* pairs, const Grid& g, const ActParams& act,
                               uint32_t* ids, uint32_t* count, uint32_t* map, TileSums ts,
                               const uint32_t* coarse, const uint32_t (*coarse_ext)[3], hipStream_t s);


// ---- batches -----------------------------------------------------------------

// One (volume, isovalue) unit of a batch, as the batched kernels read it.  The
// active lists of all units of a batch class are concatenated in unit order.
struct BatchUnit {
    const void* data;           // the unit's volume (x fastest)
    const uint32_t* pairs;      // the volume's leaf min/max key pairs, 2 words per block
    uint32_t* map;              // this unit's dense block ID -> active index map
    uint32_t X, Y, Z;
    uint32_t gx, gy, gz;
    uint32_t tile0;             // first compaction tile of this unit
    uint32_t nb;                // gx * gy * gz
    ActParams act;              // isovalue and activity rule
    uint32_t _pad[3];
};

// One volume of a batch class for the batched min/max (B = 4, 8 or 16, overlap 1).
struct BatchVolume {
    const void* data;
    uint32_t* pairs;            // 2 words per block
    uint32_t X, Y, Z;
    uint32_t gx, gy, gz;
    uint32_t aligned;           // X % 4 == 0 and data 16-byte aligned
    uint32_t ntx, wx, ntxg, nygrp, zc;   // launch plan of this volume
    uint32_t wg0;               // first workgroup of this volume
};

// Fill the launch plan of every volume (block edge B, M y-blocks per wave)
// and return the workgroup count (0: not supported).  wg0 of each volume must
// then be copied to a separate device array for the workgroup -> volume search.
uint64_t plan_minmax_batch(BatchVolume* vols, uint32_t nvol, int dtype, uint32_t B, uint32_t M);
hipError_t launch_minmax_batch(const BatchVolume* vols, const uint32_t* wg0, uint32_t nvol, int dtype,
                               uint32_t B, uint32_t M, uint64_t nwg, hipStream_t s);

// Ordered compaction over every unit of a batch class: active block IDs (local
// to their unit), each active block's unit, the first active index of every
// unit (unit_first[nunits] = total) and the total count; writes each unit's map.
uint32_t batch_filter_tiles(uint32_t nb);   // compaction tiles of a unit of nb blocks
hipError_t launch_filter_batch(const BatchUnit* units, const uint32_t* tile0, uint32_t nunits,
                               uint32_t ntiles, uint32_t* ids, uint32_t* item_unit,
                               uint32_t* unit_first, uint32_t* count, TileSums ts, hipStream_t s);

// Per-unit output ranges after the extraction scan (see mtx_batch_range).
struct BatchRangeArgs {
    const uint32_t* unit_first;
    const uint32_t* unit_global;   // class-local unit -> batch unit index
    uint32_t nunits;
    const uint32_t* count;
    const uint64_t* vbase;
    const uint64_t* tbase;
    const uint64_t* mbase;
    const uint64_t* totals;
    uint64_t voff, toff, moff, aoff;   // the class's base in the batch buffers
    uint64_t vcap, icap, mcap;         // the class's capacities (relative to its base)
    uint32_t block_dim;
    void* ranges;                      // mtx_batch_range[batch units]
};
hipError_t launch_batch_ranges(const BatchRangeArgs& a, hipStream_t s);

// ---- extraction --------------------------------------------------------------

struct ExtractArgs {
    VolDesc vol;
    Grid grid;
    float iso;
    const uint32_t* ids;        // active block IDs (ascending)
    const uint32_t* count;      // device active count
    const uint32_t* blk_map;    // dense block ID -> active index (0xFFFFFFFF: inactive)
    uint32_t* blk_counts;       // per active block: {vertices, triangles, meshlets, meshlet vertices}
    void* corner_table;         // per active block: B^3 entries (prefix<<3 | flags)
    uint64_t* blk_vbase;        // per active block vertex base
    uint64_t* blk_tbase;        // per active block triangle base
    uint64_t* blk_mbase;        // per active block meshlet base
    uint64_t* blk_sbase;        // per active block meshlet-vertex base
    uint32_t* starts;           // per active block: meshlet start bits, one per cell ((B^3+31)/32
                                // words); nullptr: no meshlets
    uint64_t* totals;           // [0]=vertices [1]=triangles [2]=error flags [3]=meshlets
                                // [4]=meshlet vertices
    uint8_t* cases;             // per active block: B^3 case bytes (0 for cells outside the volume)
    // outputs
    float* vertices;            // 8 floats per vertex (position xyzw, normal xyzw)
    uint32_t* indices;          // 3 per triangle
    uint64_t* edge_ids;         // optional
    float* vox_pos;             // optional: 3 floats per vertex, voxel units
    float* edge_t;              // optional: t along the owning edge
    uint4* meshlets;            // with starts: {vertexOffset, indexOffset, vertexCount, primitiveCount}
    float* mverts;              // with starts: meshlet vertices, 8 floats each
    uint32_t* mindices;         // with starts: 3 per triangle, into mverts
    uint32_t* mvmap;            // with starts: per meshlet vertex, its vertex index
    uint64_t vcap, icap, mcap;  // capacities (vertices, indices, meshlets)
    uint64_t mvcap;             // capacity of mverts / mvmap
    uint32_t* out_counts;       // optional device uint32[3]: vertexCount, indexCount, meshletCount
    int normals;
    uint32_t max_blocks;        // capacity of the per-block arrays
    // batched form (units != nullptr): vol, iso and blk_map come per active
    // block from units[item_unit[i]]; indices and meshlet offsets are relative
    // to the unit's first vertex / index (unit_first: first active index per unit)
    const BatchUnit* units;
    const uint32_t* item_unit;
    const uint32_t* unit_first;
    // work assignment of the vertex and triangle passes: 0 strided over the
    // list, 1 contiguous ranges of equal cost (vertices / triangles + a fixed
    // cost per block); persistent workgroups per CU (0: the measured default)
    uint32_t balanced;
    uint32_t wg_per_cu;
    // single volume: a second stream and two events of the context; the
    // vertex pass runs on it beside the triangle pass (nullptr: one stream)
    hipStream_t side;
    hipEvent_t side_ev[2];
};

bool extract_supported_block(uint32_t B);
bool batch_supported_block(uint32_t B);  // batched form: B = 4, 8, 16
size_t corner_entry_bytes(uint32_t B);
uint32_t scan_tile_items();

hipError_t launch_extract_count(const ExtractArgs& a, hipStream_t s);
// Also clears the emit pass's error flags (totals[2]); TileSums: 3 words per scan tile.
hipError_t launch_extract_scan(const ExtractArgs& a, TileSums ts, hipStream_t s);
hipError_t launch_extract_emit(const ExtractArgs& a, hipStream_t s);

// Adjoint: dL/dvoxel (dense, overwritten) and dL/diso (one float, overwritten)
// from dL/d(vertex position in voxel units), topology of the forward pass.
struct AdjointArgs {
    ExtractArgs fwd;            // forward topology (map, corner tables, vertex bases)
    const float* dl_dpos;       // 3 per vertex
    float* dl_dvol;             // one per voxel
    float* dl_diso;             // one float
    double* iso_partials;       // one double per adjoint workgroup
    uint32_t n_wg;
};
uint32_t adjoint_workgroups(uint64_t nb);
hipError_t launch_adjoint(const AdjointArgs& a, hipStream_t s);

}  // namespace mtx
