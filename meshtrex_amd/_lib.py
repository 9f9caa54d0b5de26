"""ctypes view of the C ABI in include/meshtrex_amd/mtx.h.

Loads the in-tree libmeshtrex_amd.so built by meshtrex_amd.build; there is no
fallback: if the library is missing, importing a function raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmeshtrex_amd.so")
SYNTH_PATH = os.path.join(HERE, "libmtx_synth.so")

u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
f32 = ctypes.c_float
vp = ctypes.c_void_p

MTX_MAX_LEVELS = 24
MTX_OK, MTX_EINVAL, MTX_ENOMEM, MTX_EDEVICE, MTX_ECAPACITY, MTX_ESTATE = range(6)
MTX_U8, MTX_F32 = 0, 1
MTX_RULE_CPU, MTX_RULE_GPU = 0, 1


class PushConstants(ctypes.Structure):
    """PushConstants of src/common.h:29-34 (64 bytes)."""
    _fields_ = [("volumeDim", u32 * 4), ("blockDim", u32 * 4), ("blockGridDim", u32 * 4),
                ("isovalue", f32), ("_pad", u32 * 3)]


class CVolume(ctypes.Structure):
    _fields_ = [("data", vp), ("dims", u32 * 3), ("dtype", u32), ("on_device", u32), ("_pad", u32)]


class Options(ctypes.Structure):
    _fields_ = [("overlap", u32), ("rule", u32), ("build_pyramid", u32), ("fuse_filter", u32),
                ("normals", u32), ("edge_ids", u32), ("case_indices", u32),
                ("voxel_positions", u32), ("reuse", u32), ("minmax_rows", u32), ("coarse_reject", u32),
                ("meshlets", u32), ("_reserved", u32 * 4)]


class CMinMaxOutput(ctypes.Structure):
    _fields_ = [("minmax", vp), ("level_offset", u64 * (MTX_MAX_LEVELS + 1)),
                ("level_extent", (u32 * 3) * MTX_MAX_LEVELS), ("num_levels", u32),
                ("levels_built", u32), ("dtype", u32),
                ("dims", u32 * 3), ("block_dim", u32), ("grid", u32 * 3), ("overlap", u32),
                ("volume", vp), ("owned_volume", vp), ("active_mask", vp),
                ("mask_words_per_row", u32), ("mask_isovalue", f32), ("mask_rule", u32),
                ("mask_valid", u32)]


class CFilteringOutput(ctypes.Structure):
    _fields_ = [("compacted_block_ids", vp), ("active_block_count", vp), ("block_to_active", vp),
                ("activeBlockCount", u32), ("capacity", u32), ("isovalue", f32), ("rule", u32),
                ("block_dim", u32), ("grid", u32 * 3), ("done", vp)]


class CExtractionOutput(ctypes.Structure):
    _fields_ = [("vertices", vp), ("counts", vp), ("indices", vp), ("meshlets", vp),
                ("vertex_edge_ids", vp), ("vertex_t", vp), ("vertex_voxel_pos", vp), ("case_indices", vp),
                ("workspace", vp), ("workspace_bytes", u64), ("vertex_capacity", u64),
                ("index_capacity", u64), ("meshlet_capacity", u64), ("case_capacity", u64),
                ("meshletCount", u32),
                ("vertexCount", u32), ("indexCount", u32), ("block_dim", u32),
                ("max_blocks", u32), ("flags", u32),
                ("meshlet_vertices", vp), ("meshlet_indices", vp), ("meshlet_vertex_map", vp),
                ("meshlet_vertex_capacity", u64), ("meshletVertexCount", u32), ("grid", u32 * 3),
                ("done", vp), ("_pad", u32)]


class CBatchVolume(ctypes.Structure):
    _fields_ = [("volume", CVolume), ("isovalues", ctypes.POINTER(f32)), ("num_isovalues", u32),
                ("_pad", u32)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("block_dim", u32), ("minmax_rows", u32), ("rule", u32), ("normals", u32),
                ("edge_ids", u32), ("graph", u32), ("tune_iters", u32), ("headroom_pct", u32),
                ("schedule", u32), ("wg_per_cu", u32), ("_reserved", u32 * 6)]


class BatchRange(ctypes.Structure):
    _fields_ = [("vertex_offset", u64), ("vertex_count", u64), ("index_offset", u64),
                ("index_count", u64), ("meshlet_offset", u64), ("meshlet_count", u64),
                ("active_offset", u32), ("active_count", u32), ("block_dim", u32), ("flags", u32)]


class CBatchOutput(ctypes.Structure):
    _fields_ = [("vertices", vp), ("indices", vp), ("meshlets", vp), ("vertex_edge_ids", vp),
                ("active_block_ids", vp), ("ranges", vp), ("vertex_capacity", u64),
                ("index_capacity", u64), ("meshlet_capacity", u64), ("num_units", u32),
                ("num_classes", u32)]


class BatchClassInfo(ctypes.Structure):
    _fields_ = [("dtype", u32), ("block_dim", u32), ("minmax_rows", u32), ("num_volumes", u32),
                ("num_units", u32), ("active_blocks", u32), ("blocks", u64), ("voxels", u64),
                ("vertices", u64), ("triangles", u64), ("tuned_ms", f32 * 3), ("tuned_rows_ms", f32 * 3),
                ("schedule", u32), ("wg_per_cu", u32), ("tuned_sched_ms", f32 * 6)]


# name -> (restype, argtypes)
_SIGS = {
    "mtx_abi_version": (ctypes.c_int, []),
    "mtx_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "mtx_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
    "mtx_context_destroy": (None, [vp]),
    "mtx_context_last_error": (ctypes.c_char_p, [vp]),
    "mtx_options_default": (None, [ctypes.POINTER(Options)]),
    "mtx_push_constants_init": (None, [ctypes.POINTER(PushConstants), ctypes.POINTER(u32), u32, f32]),
    "mtx_compute_min_max": (ctypes.c_int, [vp, ctypes.POINTER(CVolume), ctypes.POINTER(PushConstants),
                                           ctypes.POINTER(Options), vp, ctypes.POINTER(CMinMaxOutput)]),
    "mtx_minmax_release": (None, [vp, ctypes.POINTER(CMinMaxOutput)]),
    "mtx_build_pyramid": (ctypes.c_int, [vp, ctypes.POINTER(CMinMaxOutput), vp]),
    "mtx_filter_active_blocks": (ctypes.c_int, [vp, ctypes.POINTER(CMinMaxOutput),
                                                ctypes.POINTER(PushConstants), ctypes.POINTER(Options),
                                                vp, ctypes.POINTER(CFilteringOutput)]),
    "mtx_filter_active_blocks_noalloc": (ctypes.c_int, [vp, ctypes.POINTER(CMinMaxOutput),
                                                        ctypes.POINTER(PushConstants), ctypes.POINTER(Options),
                                                        vp, vp, vp, vp]),
    "mtx_read_active_count": (ctypes.c_int, [vp, ctypes.POINTER(CFilteringOutput)]),
    "mtx_filtering_release": (None, [vp, ctypes.POINTER(CFilteringOutput)]),
    "mtx_extract": (ctypes.c_int, [vp, ctypes.POINTER(CMinMaxOutput), ctypes.POINTER(CFilteringOutput),
                                   ctypes.POINTER(PushConstants), ctypes.POINTER(Options), vp,
                                   ctypes.POINTER(CExtractionOutput)]),
    "mtx_read_extraction_counts": (ctypes.c_int, [vp, ctypes.POINTER(CExtractionOutput)]),
    "mtx_extraction_release": (None, [vp, ctypes.POINTER(CExtractionOutput)]),
    "mtx_device_malloc": (ctypes.c_int, [ctypes.POINTER(vp), u64]),
    "mtx_device_free": (ctypes.c_int, [vp]),
    "mtx_host_malloc": (ctypes.c_int, [ctypes.POINTER(vp), u64]),
    "mtx_host_free": (ctypes.c_int, [vp]),
    "mtx_memcpy_h2d": (ctypes.c_int, [vp, vp, u64]),
    "mtx_memcpy_d2h": (ctypes.c_int, [vp, vp, u64]),
    "mtx_memcpy_d2d": (ctypes.c_int, [vp, vp, u64]),
    "mtx_memcpy_d2d_async": (ctypes.c_int, [vp, vp, u64, vp]),
    "mtx_stream_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "mtx_stream_destroy": (ctypes.c_int, [vp]),
    "mtx_stream_synchronize": (ctypes.c_int, [vp]),
    "mtx_device_synchronize": (ctypes.c_int, []),
    "mtx_key_f32": (u32, [f32]),
    "mtx_unkey_f32": (f32, [u32]),
    "mtx_struct_size": (u64, [ctypes.c_char_p]),
    "mtx_profiler_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "mtx_profiler_destroy": (None, [vp]),
    "mtx_profiler_begin": (ctypes.c_int, [vp, ctypes.c_char_p, vp]),
    "mtx_profiler_end": (ctypes.c_int, [vp, vp]),
    "mtx_profiler_region": (ctypes.c_int, [vp, u32, ctypes.c_char_p, u32, ctypes.POINTER(f32)]),
    "mtx_profiler_reset": (None, [vp]),
    "mtx_write_obj": (ctypes.c_int, [vp, ctypes.POINTER(CExtractionOutput), ctypes.c_char_p]),
    "mtx_adjoint": (ctypes.c_int, [vp, ctypes.POINTER(CMinMaxOutput), ctypes.POINTER(CFilteringOutput),
                                   ctypes.POINTER(CExtractionOutput), ctypes.POINTER(PushConstants),
                                   vp, vp, vp, vp]),
    "mtx_batch_options_default": (None, [ctypes.POINTER(BatchOptions)]),
    "mtx_batch_create": (ctypes.c_int, [vp, ctypes.POINTER(CBatchVolume), u32, ctypes.POINTER(BatchOptions),
                                        ctypes.POINTER(vp)]),
    "mtx_batch_run": (ctypes.c_int, [vp, vp]),
    "mtx_batch_output_get": (ctypes.c_int, [vp, ctypes.POINTER(CBatchOutput)]),
    "mtx_batch_read_ranges": (ctypes.c_int, [vp, ctypes.POINTER(BatchRange), u32]),
    "mtx_batch_class_info_get": (ctypes.c_int, [vp, u32, ctypes.POINTER(BatchClassInfo)]),
    "mtx_batch_destroy": (None, [vp]),
}

SYNTH_SIGS = {
    "mtx_synth_blobs_u8": (ctypes.c_int, [vp, ctypes.POINTER(u32), u64, ctypes.c_int, vp]),
    "mtx_synth_blobs_f32": (ctypes.c_int, [vp, ctypes.POINTER(u32), u64, ctypes.c_int, vp]),
    "mtx_synth_vessels_u8": (ctypes.c_int, [vp, ctypes.POINTER(u32), u64, ctypes.c_int, ctypes.c_int, vp]),
    "mtx_synth_sdf_f32": (ctypes.c_int, [vp, ctypes.POINTER(u32), u64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, vp]),
    "mtx_bench_copy": (ctypes.c_int, [vp, vp, u64, vp]),
    "mtx_bench_read": (ctypes.c_int, [vp, u64, vp, vp]),
}

_lib = None
_synth = None


def _one_runtime():
    """torch ships its own HIP and HSA runtimes, and its libraries ask for them
    as libamdhip64.so / libhsa-runtime64.so; this library links the system
    ones by their versioned names.  Two runtimes in one process cannot both
    open the GPU, so the runtime is loaded under the names torch asks for
    before this library: a later `import torch` then reuses it, and after an
    earlier `import torch` these names resolve to torch's copies, which this
    library then uses."""
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        try:
            ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            pass


def _load(path, sigs):
    if not os.path.exists(path):
        raise ImportError("%s is not built; run `python -m meshtrex_amd.build`" % path)
    _one_runtime()
    lib = ctypes.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The extraction library (raises ImportError if it is not built)."""
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH, _SIGS)
    return _lib


def synth():
    global _synth
    if _synth is None:
        _synth = _load(SYNTH_PATH, SYNTH_SIGS)
    return _synth


def exported_symbols():
    return sorted(_SIGS)
