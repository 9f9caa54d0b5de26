/*
 * meshtrex_amd C ABI: isosurface extraction (block min/max -> active-block
 * filter -> PMB marching cubes) on one AMD MI355X (gfx950) GPU.
 *
 * Plain C: POD structs, device pointers as `void*`/typed pointers, sizes as
 * fixed-width integers, streams as `void*` (a hipStream_t).  Every entry point
 * returns an mtx_status and never asserts; mtx_context_last_error() gives the
 * message of the last failure on a context.
 *
 * The three stages replace the reference's host managers:
 *   mtx_compute_min_max      <- computeMinMaxMip          (src/minMaxManager.h:20-21,
 *                                                          src/minMaxManager.cpp:63-225)
 *   mtx_filter_active_blocks <- filterActiveBlocks        (src/filteringManager.h:9-10,
 *                                                          src/filteringManager.cpp:20-141)
 *   mtx_read_active_count    <- readActiveBlockCount      (src/filteringManager.h:13,
 *                                                          src/filteringManager.cpp:144-164)
 *   mtx_extract              <- extractMeshletDescriptors (src/extractionManager.h:10-11,
 *                                                          src/extractionManager.cpp:250-717)
 * and the output structs mirror MinMaxOutput (src/minMaxOutput.h:10-64),
 * FilteringOutput (src/filteringOutput.h:11-74) and ExtractionOutput
 * (src/extractionOutput.h:10-134).  The C++ wrapper with the reference's
 * names and RAII outputs is include/meshtrex_amd/managers.hpp.
 *
 * Streams: a NULL stream means "use the context's own stream and synchronise
 * before returning" (the reference's externalCmd == VK_NULL_HANDLE case);
 * with a caller stream every call is asynchronous and stream-ordered.  A
 * context's scratch memory is shared by its calls: use one context per
 * concurrently running stream.
 */
#ifndef MESHTREX_AMD_MTX_H
#define MESHTREX_AMD_MTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTX_ABI_VERSION 1
#define MTX_MAX_LEVELS 24

typedef enum mtx_status {
    MTX_OK = 0,
    MTX_EINVAL = 1,     /* bad argument (dims, block edge, dtype, null pointer) */
    MTX_ENOMEM = 2,     /* device allocation failed */
    MTX_EDEVICE = 3,    /* HIP runtime error */
    MTX_ECAPACITY = 4,  /* output buffers too small (reuse mode) */
    MTX_ESTATE = 5      /* inputs do not belong together (e.g. filter of another min-max) */
} mtx_status;

typedef enum mtx_dtype { MTX_U8 = 0, MTX_F32 = 1 } mtx_dtype;

/* Block activity rule.
 * MTX_RULE_CPU: min != max and round(iso) in [min, max] (u8, integer compare)
 *               or iso in [min, max] (f32): src/blockFilteringTestUtils.h:456-468.
 * MTX_RULE_GPU: iso >= min and iso <= max in fp32: the reference's live shader
 *               src/shaders/occupiedBlockPrefixSum.comp.glsl:84. */
typedef enum mtx_rule { MTX_RULE_CPU = 0, MTX_RULE_GPU = 1 } mtx_rule;

/* Byte-compatible with the reference's PushConstants (src/common.h:29-34):
 * uvec4 volumeDim, uvec4 blockDim, uvec4 blockGridDim, float isovalue; 64 B.
 * Only .x of blockDim is used (cubic blocks, edge 1..16 cells). */
typedef struct mtx_push_constants {
    uint32_t volumeDim[4];
    uint32_t blockDim[4];
    uint32_t blockGridDim[4];
    float isovalue;
    uint32_t _pad[3];
} mtx_push_constants;

/* The volume: x-fastest, linear index x + X*(y + Y*z) (src/testMinMax.h:97-101).
 * on_device = 1: `data` is device memory and stays owned by the caller.
 * on_device = 0: `data` is host memory; mtx_compute_min_max uploads it into a
 * device buffer owned by the min-max output (the reference's volumeImage). */
typedef struct mtx_volume {
    const void* data;
    uint32_t dims[3];
    uint32_t dtype;     /* mtx_dtype */
    uint32_t on_device;
    uint32_t _pad;
} mtx_volume;

/* Options; pass NULL for the defaults shown.  Output structs passed to the
 * entry points must be zero-initialised, or come from an earlier call. */
typedef struct mtx_options {
    uint32_t overlap;        /* 1: block window [B*b, B*b+B] (live GPU shader, default);
                                0: [B*b, B*b+B-1] (reference CPU oracle; min-max only) */
    uint32_t rule;           /* mtx_rule, default MTX_RULE_CPU */
    uint32_t build_pyramid;  /* 1 (default): all mip levels of the min/max pyramid */
    uint32_t fuse_filter;    /* 1 (default): min-max also writes the activity mask of
                                pc->isovalue, so the filter does not re-read the pairs */
    uint32_t normals;        /* 1 (default): vertex normals; 0: normal = (0,0,0,0) */
    uint32_t edge_ids;       /* 1: also write each vertex's owning edge ID (default 0) */
    uint32_t case_indices;   /* reserved (the case indices are always kept, see
                                mtx_extraction_output.case_indices) */
    uint32_t voxel_positions;/* 1: also write float3 positions in voxel units (default 0) */
    uint32_t reuse;          /* 1: reuse the output's existing buffers when they fit the
                                problem: no allocation, no host synchronisation.  For
                                mtx_extract an overflow is then reported by
                                mtx_read_extraction_counts (default 0) */
    uint32_t minmax_rows;    /* B = 8 min-max: y-blocks per wavefront (1, 2 or 4); 0 = chosen
                                from the volume's shape (default) */
    uint32_t coarse_reject;  /* filter from min/max pairs: skip a block whose pyramid-level-2
                                ancestor (4^3 blocks) is inactive, without reading its own pair
                                (src/shaders/occupiedBlockPrefixSum.comp.glsl:47-63); needs the
                                pyramid; same result (default 0) */
    uint32_t meshlets;       /* 1 (default): meshlet descriptors and the renderer-layout mesh
                                (see mtx_extraction_output); 0: the shared mesh only */
    uint32_t _reserved[4];
} mtx_options;

/* Min/max of every block (level 0) and the min/max pyramid, as uint32 key
 * pairs: u8 values as-is, f32 through the order-preserving map
 * key = bits ^ (sign ? 0xFFFFFFFF : 0x80000000) (mtx_key_f32). */
typedef struct mtx_minmax_output {
    uint32_t* minmax;                          /* device, 2 * level_offset[num_levels] */
    uint64_t level_offset[MTX_MAX_LEVELS + 1]; /* in pairs */
    uint32_t level_extent[MTX_MAX_LEVELS][3];
    uint32_t num_levels;                       /* 1 + floor(log2(max grid extent)) */
    uint32_t levels_built;                     /* levels holding data (1 = leaf only) */
    uint32_t dtype;
    uint32_t dims[3];
    uint32_t block_dim;
    uint32_t grid[3];
    uint32_t overlap;
    const void* volume;                        /* device volume the pairs describe */
    void* owned_volume;                        /* set when uploaded from host memory */
    uint32_t* active_mask;                     /* fused activity bits, or NULL */
    uint32_t mask_words_per_row;
    float mask_isovalue;
    uint32_t mask_rule;
    uint32_t mask_valid;
} mtx_minmax_output;

typedef struct mtx_filtering_output {
    uint32_t* compacted_block_ids;  /* device, ascending, capacity = number of blocks */
    uint32_t* active_block_count;   /* device, one uint32 */
    uint32_t* block_to_active;      /* device, one per block: index in the list, or 0xFFFFFFFF */
    uint32_t activeBlockCount;      /* host mirror, filled by mtx_read_active_count */
    uint32_t capacity;
    float isovalue;
    uint32_t rule;
    uint32_t block_dim;             /* block grid the list was made for: extraction and the */
    uint32_t grid[3];               /* adjoint reject a min/max output of another grid */
    void* done;                     /* hipEvent_t recorded after the filter's last launch */
} mtx_filtering_output;

/* VertexData (src/extractionOutput.h:131-134): NDC position (p/dim)*2-1, w = 1,
 * and unit normal, w = 0. */
typedef struct mtx_vertex {
    float position[4];
    float normal[4];
} mtx_vertex;

/* MeshletDescriptor (src/extractionOutput.h:124-129). */
typedef struct mtx_meshlet {
    uint32_t vertexOffset;
    uint32_t indexOffset;
    uint32_t vertexCount;
    uint32_t primitiveCount;
} mtx_meshlet;

/* Meshlet limits of the reference (src/shaders/marching_cubes_pmb.task.glsl:24-25). */
#define MTX_MESHLET_MAX_VERTICES 64
#define MTX_MESHLET_MAX_PRIMITIVES 126

/* Two meshes are written.  The shared mesh: `vertices` holds each vertex once
 * (one per active grid edge, PMB ownership) and `indices` indexes it.  The
 * renderer layout (option meshlets, default on): `meshlets` are the
 * reference's meshlet descriptors -- a greedy partition of each active block's
 * occupied cells in cell order into runs of at most 64 vertices and 126
 * triangles (src/shaders/marching_cubes_pmb.task.glsl:182-215) -- over
 * `meshlet_vertices` / `meshlet_indices`: each meshlet's vertices are
 * contiguous, every index of its triangles lies in
 * [vertexOffset, vertexOffset + vertexCount) (the renderer's acceptance rule,
 * src/shaders/render.task.glsl:61-63,125-128), and a vertex used by several
 * meshlets has a copy in each.  Triangle k is the same triangle in both:
 * meshlet_vertex_map[meshlet_indices[i]] == indices[i] for every i, and
 * meshlet_vertices[j] == vertices[meshlet_vertex_map[j]]. */
typedef struct mtx_extraction_output {
    mtx_vertex* vertices;        /* device */
    uint32_t* counts;            /* device uint32[4]: vertexCount, indexCount, meshletCount,
                                    meshletVertexCount */
    uint32_t* indices;           /* device, 3 per triangle, global vertex indices */
    mtx_meshlet* meshlets;       /* device, renderer layout (see above), or NULL (meshlets = 0) */
    uint64_t* vertex_edge_ids;   /* optional (edge_ids): 3*linear(owner corner) + axis */
    float* vertex_t;             /* optional (edge_ids): t along the owning edge, in [0,1] */
    float* vertex_voxel_pos;     /* optional: float3 per vertex, voxel units */
    uint8_t* case_indices;       /* case index of every cell of every active block, B^3 bytes per
                                    block in list order, x-fastest; 0 outside the volume (a view
                                    into the workspace) */
    void* workspace;             /* per-block topology kept for the adjoint */
    uint64_t workspace_bytes;
    uint64_t vertex_capacity;    /* in vertices */
    uint64_t index_capacity;     /* in uint32 indices */
    uint64_t meshlet_capacity;   /* in descriptors */
    uint64_t case_capacity;      /* in bytes */
    uint32_t meshletCount;       /* host mirrors (mtx_read_extraction_counts) */
    uint32_t vertexCount;
    uint32_t indexCount;
    uint32_t block_dim;
    uint32_t max_blocks;
    uint32_t flags;              /* options the buffers were made for */
    /* renderer layout (option meshlets) */
    mtx_vertex* meshlet_vertices;   /* device: meshlet vertices, each meshlet's contiguous */
    uint32_t* meshlet_indices;      /* device: 3 per triangle, into meshlet_vertices */
    uint32_t* meshlet_vertex_map;   /* device: per meshlet vertex, its index in `vertices` */
    uint64_t meshlet_vertex_capacity;
    uint32_t meshletVertexCount;    /* host mirror (mtx_read_extraction_counts) */
    uint32_t grid[3];               /* block grid of the extraction */
    void* done;                     /* hipEvent_t recorded after the last launch */
    uint32_t _pad;
} mtx_extraction_output;

/* ---- context --------------------------------------------------------------- */

typedef struct mtx_context mtx_context;

int mtx_abi_version(void);
const char* mtx_status_string(mtx_status s);
mtx_status mtx_context_create(int device, mtx_context** out);
void mtx_context_destroy(mtx_context* ctx);
const char* mtx_context_last_error(const mtx_context* ctx);
void mtx_options_default(mtx_options* o);

/* Fill a push-constant block for dims / block edge / isovalue the way the
 * reference's main does (src/meshtrex.cpp:76-84): blockGridDim = ceil(dims/B). */
void mtx_push_constants_init(mtx_push_constants* pc, const uint32_t dims[3], uint32_t block_dim,
                             float isovalue);

/* ---- the path ------------------------------------------------------------- */

mtx_status mtx_compute_min_max(mtx_context* ctx, const mtx_volume* volume,
                               const mtx_push_constants* pc, const mtx_options* opt, void* stream,
                               mtx_minmax_output* out);
void mtx_minmax_release(mtx_context* ctx, mtx_minmax_output* out);
/* Levels 1.. of the min/max pyramid (src/minMaxManager.cpp:145-179), for an
 * output computed with build_pyramid = 0. */
mtx_status mtx_build_pyramid(mtx_context* ctx, mtx_minmax_output* mm, void* stream);

mtx_status mtx_filter_active_blocks(mtx_context* ctx, const mtx_minmax_output* mm,
                                    const mtx_push_constants* pc, const mtx_options* opt,
                                    void* stream, mtx_filtering_output* out);
/* mtx_filter_active_blocks into caller-provided buffers, allocating nothing:
 * no hipMalloc, hipFree or allocating library call.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same
 * results as mtx_filter_active_blocks: `compacted_block_ids` (room for one
 * uint32 per block), `active_block_count` (one uint32), `block_to_active`
 * (one uint32 per block).  The compaction's per-tile counts live in a static
 * device array of 2^18 tiles (1 MiB on each GPU the library runs on, held
 * until the process ends); a grid needing more tiles (over 2^29 blocks
 * filtered from min/max pairs, or over 2^24 words of the fused activity mask)
 * is rejected with MTX_EINVAL.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU race; nothing in it has to start at zero. */
mtx_status mtx_filter_active_blocks_noalloc(mtx_context* ctx, const mtx_minmax_output* mm,
                                            const mtx_push_constants* pc, const mtx_options* opt,
                                            void* stream, uint32_t* compacted_block_ids,
                                            uint32_t* active_block_count, uint32_t* block_to_active);
/* Blocking 4-byte readback of the active count into out->activeBlockCount. */
mtx_status mtx_read_active_count(mtx_context* ctx, mtx_filtering_output* out);
void mtx_filtering_release(mtx_context* ctx, mtx_filtering_output* out);

/* Marching cubes over the active blocks.  Needs overlap-1 min-max.  Does not
 * need a host-side active count (the count stays on the device). */
mtx_status mtx_extract(mtx_context* ctx, const mtx_minmax_output* mm,
                       const mtx_filtering_output* filt, const mtx_push_constants* pc,
                       const mtx_options* opt, void* stream, mtx_extraction_output* out);
/* Blocking readback of the three counters into the host mirrors; returns
 * MTX_ECAPACITY if the last reuse-mode extraction overflowed its buffers and
 * MTX_ESTATE if the active list did not match the volume. */
mtx_status mtx_read_extraction_counts(mtx_context* ctx, mtx_extraction_output* out);
void mtx_extraction_release(mtx_context* ctx, mtx_extraction_output* out);

/* ---- differentiation ------------------------------------------------------ */

/* Adjoint of the extraction with its topology frozen: from dL/d(vertex
 * position) -- `dl_dpos`, 3 floats per vertex in voxel units, in the vertex
 * order of `ext` -- to dL/d(voxel value) (`dl_dvol`, one float per voxel,
 * overwritten everywhere, zero where no vertex depends on the voxel) and
 * dL/d(isovalue) (`dl_diso`, one float).  Positions are p = owner + t * e_axis
 * with t = (iso - v0) / (v1 - v0) (src/shaders/marching_cubes_pmb.mesh.glsl:122-141);
 * t is recomputed from the volume `mm` describes, which must be the one `ext`
 * was extracted from.  No reference counterpart: the reference has no
 * differentiation.  Deterministic (no atomics). */
mtx_status mtx_adjoint(mtx_context* ctx, const mtx_minmax_output* mm, const mtx_filtering_output* filt,
                       const mtx_extraction_output* ext, const mtx_push_constants* pc,
                       const float* dl_dpos, float* dl_dvol, float* dl_diso, void* stream);

/* ---- batches ----------------------------------------------------------------
 *
 * Many volumes of any dimensions, each with one or more isovalues, extracted by
 * one launch sequence: every (volume, isovalue) pair is a "unit".  Units are
 * numbered volume by volume, isovalues in the order given.  mtx_batch_create
 * plans the batch once -- it picks block edge and min/max wave shape per dtype
 * (autotuned by timing the candidates on the batch itself unless fixed in the
 * options), sizes the outputs from a planning run and, by default, captures the
 * whole sequence (min/max, compaction, count, scan, vertices, triangles,
 * per-unit ranges for every dtype class) in one hipGraph.  mtx_batch_run then
 * replays it: no allocation, no host synchronisation.  Each unit's slice of
 * the outputs equals what mtx_extract gives for that volume and isovalue at
 * the unit's block edge, with indices and meshlet offsets relative to the
 * unit's first vertex and index.  Volumes must stay device-resident at the
 * same addresses; their contents may change between runs (a run whose
 * outputs exceed the planned capacity flags the units concerned).  No
 * reference counterpart: the reference extracts one volume at one isovalue
 * per call (src/meshtrex.cpp:107-132). */

typedef struct mtx_batch mtx_batch;

typedef struct mtx_batch_volume {
    mtx_volume volume;           /* device-resident: volume.on_device must be 1 */
    const float* isovalues;      /* host array of num_isovalues, read by mtx_batch_create */
    uint32_t num_isovalues;
    uint32_t _pad;
} mtx_batch_volume;

typedef struct mtx_batch_options {
    uint32_t block_dim;          /* 4, 8 or 16; 0 = autotuned per dtype (default) */
    uint32_t minmax_rows;        /* y-blocks per min/max wavefront 1, 2, 4; 0 = autotuned (default) */
    uint32_t rule;               /* activity rule (default MTX_RULE_CPU) */
    uint32_t normals;            /* default 1 */
    uint32_t edge_ids;           /* vertex_edge_ids output (default 0) */
    uint32_t graph;              /* 1: replay one captured hipGraph (default); 0: launch directly */
    uint32_t tune_iters;         /* timed runs per autotuner candidate (default 3) */
    uint32_t headroom_pct;       /* output capacity above the planning run's counts (default 0) */
    uint32_t schedule;           /* work assignment of the vertex and triangle passes: 0 = autotuned
                                    (default), 1 = strided over the active list, 2 = contiguous
                                    ranges of equal cost (vertices / triangles per block) */
    uint32_t wg_per_cu;          /* persistent workgroups per CU of the extraction passes; 0 =
                                    autotuned with the schedule (default) */
    uint32_t _reserved[6];
} mtx_batch_options;

/* One unit's slice of the batch outputs, written on the device by every run. */
typedef struct mtx_batch_range {
    uint64_t vertex_offset;      /* into vertices (and vertex_edge_ids) */
    uint64_t vertex_count;
    uint64_t index_offset;       /* into indices; index values are relative to vertex_offset */
    uint64_t index_count;
    uint64_t meshlet_offset;     /* into meshlets; their offsets are relative to the unit */
    uint64_t meshlet_count;
    uint32_t active_offset;      /* into active_block_ids (block IDs within the unit's volume) */
    uint32_t active_count;
    uint32_t block_dim;          /* the unit's block edge */
    uint32_t flags;              /* bit 0: capacity exceeded (output incomplete); bit 1: inconsistent list */
} mtx_batch_range;

typedef struct mtx_batch_output {
    mtx_vertex* vertices;        /* device */
    uint32_t* indices;           /* device */
    mtx_meshlet* meshlets;       /* device */
    uint64_t* vertex_edge_ids;   /* device, or NULL without the edge_ids option */
    uint32_t* active_block_ids;  /* device */
    mtx_batch_range* ranges;     /* device, one per unit */
    uint64_t vertex_capacity;
    uint64_t index_capacity;
    uint64_t meshlet_capacity;
    uint32_t num_units;
    uint32_t num_classes;        /* (dtype, block edge) classes the batch runs */
} mtx_batch_output;

/* What the planner chose for one dtype class. */
typedef struct mtx_batch_class_info {
    uint32_t dtype;
    uint32_t block_dim;
    uint32_t minmax_rows;
    uint32_t num_volumes;
    uint32_t num_units;
    uint32_t active_blocks;      /* planning run */
    uint64_t blocks;             /* sum of blocks over the class's units */
    uint64_t voxels;             /* sum of voxels over the class's volumes */
    uint64_t vertices;           /* planning run */
    uint64_t triangles;          /* planning run */
    float tuned_ms[3];           /* autotuner: class pipeline at B = 4, 8, 16 (0 = not timed) */
    float tuned_rows_ms[3];      /* autotuner: min/max at 1, 2, 4 y-blocks per wave (0 = not timed) */
    uint32_t schedule;           /* chosen work assignment (1 strided, 2 cost-balanced) */
    uint32_t wg_per_cu;          /* chosen persistent workgroups per CU */
    float tuned_sched_ms[6];     /* autotuner: class pipeline with (schedule, workgroups per CU) =
                                    (1, 32), (1, 64), (1, 128), (2, 8), (2, 16), (2, 32)
                                    (0 = not timed) */
} mtx_batch_class_info;

void mtx_batch_options_default(mtx_batch_options* o);
mtx_status mtx_batch_create(mtx_context* ctx, const mtx_batch_volume* volumes, uint32_t num_volumes,
                            const mtx_batch_options* opt, mtx_batch** out);
/* Enqueue one run on `stream` (NULL: the context's stream, and wait for it). */
mtx_status mtx_batch_run(mtx_batch* b, void* stream);
mtx_status mtx_batch_output_get(const mtx_batch* b, mtx_batch_output* out);
/* Blocking copy of the first n unit ranges of the last run to host memory;
 * returns MTX_ECAPACITY if any unit's flags has bit 0 set, MTX_ESTATE for bit 1. */
mtx_status mtx_batch_read_ranges(mtx_batch* b, mtx_batch_range* host, uint32_t n);
mtx_status mtx_batch_class_info_get(const mtx_batch* b, uint32_t index, mtx_batch_class_info* out);
void mtx_batch_destroy(mtx_batch* b);

/* ---- profiling (the reference's GPUProfiler regions, src/gpuProfiler.h:44-107) ---- */

/* Named GPU-time regions recorded with HIP events on the stream a stage runs
 * on; the managers use the reference's region names MinMax_Leaf_Pass,
 * MinMax_Octree_Reduction, Active_Block_Filtering and Mesh_Extraction
 * (src/minMaxManager.cpp:130,141, src/filteringManager.cpp:85,
 * src/extractionManager.cpp:635). */
typedef struct mtx_profiler mtx_profiler;
mtx_status mtx_profiler_create(mtx_profiler** out);
void mtx_profiler_destroy(mtx_profiler* p);
mtx_status mtx_profiler_begin(mtx_profiler* p, const char* name, void* stream);
mtx_status mtx_profiler_end(mtx_profiler* p, void* stream);
/* Blocking: milliseconds of region `index` (in begin order); name copied into
 * `name` (capacity `cap`).  Returns MTX_EINVAL past the last region. */
mtx_status mtx_profiler_region(mtx_profiler* p, uint32_t index, char* name, uint32_t cap, float* ms);
void mtx_profiler_reset(mtx_profiler* p);

/* ---- on-disk output -------------------------------------------------------- */

/* Wavefront OBJ of an extraction, as the reference's writeGPUExtractionToOBJ
 * writes it (src/extractionTestUtils.cpp:328-366): "v x y z" per vertex (the
 * NDC position), "vn x y z" per vertex, "f a//a b//b c//c" per triangle,
 * 1-based.  Blocking. */
mtx_status mtx_write_obj(mtx_context* ctx, mtx_extraction_output* ext, const char* path);

/* ---- device memory plumbing (for bindings without a device allocator) --- */

mtx_status mtx_device_malloc(void** ptr, uint64_t bytes);
mtx_status mtx_device_free(void* ptr);
/* Pinned (page-locked) host memory.  A volume read into it uploads by
 * asynchronous DMA: mtx_compute_min_max copies a host volume in slabs of whole
 * block planes on a copy stream and runs the min/max of each slab as soon as
 * it is resident (volume ingest, src/volume.h:14-42 and
 * src/minMaxManager.cpp:18-58,115-125). */
mtx_status mtx_host_malloc(void** ptr, uint64_t bytes);
mtx_status mtx_host_free(void* ptr);
mtx_status mtx_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
mtx_status mtx_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
mtx_status mtx_memcpy_d2d(void* dst, const void* src, uint64_t bytes);
mtx_status mtx_memcpy_d2d_async(void* dst, const void* src, uint64_t bytes, void* stream);
/* A non-blocking HIP stream at the default priority (the external stream a
 * caller passes where the reference passes its external command buffer). */
mtx_status mtx_stream_create(void** stream);
mtx_status mtx_stream_destroy(void* stream);
mtx_status mtx_stream_synchronize(void* stream);
mtx_status mtx_device_synchronize(void);
uint32_t mtx_key_f32(float f);
float mtx_unkey_f32(uint32_t k);
/* sizeof of a public struct by name ("mtx_volume", ...), 0 if unknown: lets a
 * binding check its mirror of the layouts. */
uint64_t mtx_struct_size(const char* name);

#ifdef __cplusplus
}
#endif

#endif /* MESHTREX_AMD_MTX_H */
