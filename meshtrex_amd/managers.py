"""Python host API with the reference's names, over the C ABI.

    computeMinMaxMip(context, volume, pushConstants, ...)      src/minMaxManager.h:20-21
    filterActiveBlocks(context, minMaxOutput, pushConstants)   src/filteringManager.h:9-10
    readActiveBlockCount(context, filteringOutput)             src/filteringManager.h:13
    extractMeshletDescriptors(context, mm, filt, pushConstants) src/extractionManager.h:10-11

Outputs own device memory (freed by cleanup() or on garbage collection) and
expose the reference's field names (activeBlockCount, vertexCount,
indexCount, meshletCount).  Every non-OK status raises RuntimeError, as the
reference's managers throw std::runtime_error (src/extractionManager.cpp:276).
Readback helpers return numpy arrays.
"""
import ctypes
import re

import numpy as np

from . import _lib as L

_DT = {"uint8": L.MTX_U8, "float32": L.MTX_F32}


class MtxError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__("%s: %s" % (L.lib().mtx_status_string(status).decode(), message))


def _check(ctx, st):
    if st != L.MTX_OK:
        msg = L.lib().mtx_context_last_error(ctx.handle if ctx else None)
        raise MtxError(st, msg.decode() if msg else "")


class Context:
    """One GPU (the reference's VulkanContext, src/vulkan_context.h:7-45)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        st = L.lib().mtx_context_create(device, ctypes.byref(h))
        if st != L.MTX_OK:
            raise MtxError(st, "cannot create a context on device %d" % device)
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            L.lib().mtx_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceBuffer:
    """Raw device allocation with host copies (no torch needed)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        st = L.lib().mtx_device_malloc(ctypes.byref(p), nbytes)
        if st != L.MTX_OK:
            raise MtxError(st, "device allocation of %d bytes" % nbytes)
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _ok(L.lib().mtx_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_numpy(self, dtype, count=None, offset_bytes=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset_bytes) // dtype.itemsize
        return read_device(self.ptr + offset_bytes, dtype, count)

    def free(self):
        if self.ptr:
            L.lib().mtx_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ok(st):
    if st != L.MTX_OK:
        raise MtxError(st, "device copy")


def read_device(ptr, dtype, count):
    out = np.empty(int(count), dtype=dtype)
    if count and ptr:
        _ok(L.lib().mtx_memcpy_d2h(out.ctypes.data, ptr, out.nbytes))
    return out


class Volume:
    """Volume of src/volume.h:8-12: dims, type name and data (host numpy or DeviceBuffer)."""

    def __init__(self, volume_dims, volume_type, volume_data):
        self.volume_dims = tuple(int(d) for d in volume_dims)
        if volume_type not in _DT:
            raise ValueError("unsupported volume type %r (uint8 or float32)" % volume_type)
        self.volume_type = volume_type
        self.volume_data = volume_data

    def _c(self):
        v = L.CVolume()
        v.dims[:] = self.volume_dims
        v.dtype = _DT[self.volume_type]
        if isinstance(self.volume_data, DeviceBuffer):
            v.data = self.volume_data.ptr
            v.on_device = 1
        elif isinstance(self.volume_data, int):
            v.data = self.volume_data
            v.on_device = 1
        else:
            a = np.ascontiguousarray(self.volume_data)
            want = np.dtype(self.volume_type)
            if a.dtype != want:
                raise ValueError("volume data is %s, type says %s" % (a.dtype, want))
            if a.size != int(np.prod(self.volume_dims, dtype=np.int64)):
                raise ValueError("volume data has %d elements for dims %s" % (a.size, self.volume_dims))
            self._host = a
            v.data = a.ctypes.data
            v.on_device = 0
        return v


_NAME_RE = re.compile(r"(\w+)_(\d+)x(\d+)x(\d+)_(.+)\.raw")


class PinnedBuffer:
    """Page-locked host memory (mtx_host_malloc) with a numpy view: a volume in
    it uploads by asynchronous DMA, overlapped with the min/max of the slabs
    already resident."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        st = L.lib().mtx_host_malloc(ctypes.byref(p), nbytes)
        if st != L.MTX_OK:
            raise MtxError(st, "pinned host allocation of %d bytes" % nbytes)
        self.ptr = p.value
        self.nbytes = nbytes

    def view(self, dtype, count):
        buf = (ctypes.c_char * self.nbytes).from_address(self.ptr)
        return np.frombuffer(buf, dtype=dtype, count=count)

    def free(self):
        if self.ptr:
            L.lib().mtx_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def pinned_array(shape_or_n, dtype):
    """A numpy array in pinned host memory (the buffer lives as long as the array)."""
    n = int(np.prod(shape_or_n))
    dtype = np.dtype(dtype)
    buf = PinnedBuffer(n * dtype.itemsize)
    a = buf.view(dtype, n)
    # keep the allocation alive with the array
    holder = np.ndarray.__new__(_PinnedArray, a.shape, dtype, buffer=a)
    holder._pinned = buf
    return holder


class _PinnedArray(np.ndarray):
    pass


def loadVolume(path, pinned=False):
    """Raw-file loader with the reference's naming contract (src/volume.h:14-42):
    '<name>_<X>x<Y>x<Z>_<type>.raw'.  Types uint8 and float32 are read; the
    reference reads 2 bytes per voxel for any non-uint8 type (src/volume.h:30),
    which has no consumer on its path (its upload is R8 only), so other types
    are rejected here.  pinned=True reads the file into pinned host memory."""
    m = _NAME_RE.search(str(path))
    if not m:
        raise RuntimeError("Invalaid raw file naming scheme: expected "
                           "'<name>_<X>x<Y>x<Z>_<data type>.raw' but %r did not match" % str(path))
    dims = (int(m.group(2)), int(m.group(3)), int(m.group(4)))
    vtype = m.group(5)
    if vtype not in _DT:
        raise RuntimeError("unsupported raw volume type %r" % vtype)
    n = dims[0] * dims[1] * dims[2]
    if pinned:
        data = pinned_array(n, vtype)
        with open(path, "rb") as f:
            got = f.readinto(memoryview(data).cast("B"))
        if got != data.nbytes:
            raise RuntimeError("Failed to read volume data: %d of %d bytes" % (got, data.nbytes))
    else:
        data = np.fromfile(path, dtype=np.dtype(vtype), count=n)
        if data.size != n:
            raise RuntimeError("Failed to read volume data: %d of %d voxels" % (data.size, n))
    return Volume(dims, vtype, data)


def makePushConstants(dims, block_dim, isovalue):
    """PushConstants as src/meshtrex.cpp:76-84 builds them (grid = ceil(dims/B))."""
    pc = L.PushConstants()
    d = (L.u32 * 3)(*dims)
    L.lib().mtx_push_constants_init(ctypes.byref(pc), d, block_dim, isovalue)
    return pc


def makeOptions(**kw):
    o = L.Options()
    L.lib().mtx_options_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k) or k.startswith("_"):
            raise TypeError("unknown option %r" % k)
        setattr(o, k, int(v))
    return o


def _stream(s):
    if s is None:
        return None
    return int(s)


class Buffer:
    """Buffer of src/buffer.h:3-8: `buffer` is the device pointer (int, 0 when
    absent), `size` its size in bytes; `data` (a host mapping) is None."""

    __slots__ = ("buffer", "size", "data")

    def __init__(self, ptr, size):
        self.buffer = int(ptr or 0)
        self.size = int(size) if ptr else 0
        self.data = None

    def __repr__(self):
        return "Buffer(buffer=0x%x, size=%d)" % (self.buffer, self.size)


class Image:
    """Image of src/image.h:3-7: `image` and `imageView` are the device pointer
    of the texels (a linear x-fastest array), `size` its bytes."""

    __slots__ = ("image", "imageView", "size")

    def __init__(self, ptr, size):
        self.image = self.imageView = int(ptr or 0)
        self.size = int(size) if ptr else 0


class MinMaxOutput:
    """MinMaxOutput (src/minMaxOutput.h:10-64): device volume + min/max pyramid."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.c = L.CMinMaxOutput()

    @property
    def num_levels(self):
        return self.c.num_levels

    @property
    def grid(self):
        return tuple(self.c.grid)

    def level_extent(self, level):
        return tuple(self.c.level_extent[level])

    # the reference's members (src/minMaxOutput.h:11-13)
    @property
    def volumeImage(self):
        esz = 4 if self.c.dtype == L.MTX_F32 else 1
        return Image(self.c.volume, esz * self.c.dims[0] * self.c.dims[1] * self.c.dims[2])

    @property
    def minMaxImage(self):
        return Image(self.c.minmax, 8 * self.c.level_offset[self.c.num_levels])

    @property
    def minMaxMipViews(self):
        """Device pointer of each pyramid level's first key pair."""
        if not self.c.minmax:
            return []
        return [self.c.minmax + 8 * self.c.level_offset[l] for l in range(self.c.num_levels)]

    def read_pairs(self, level=0):
        """(n, 2) uint32 key pairs of one level, x-fastest."""
        o0, o1 = self.c.level_offset[level], self.c.level_offset[level + 1]
        return read_device(self.c.minmax + 8 * o0, np.uint32, 2 * (o1 - o0)).reshape(-1, 2)

    def cleanup(self, device=None):
        if self.ctx is not None and self.c.minmax:
            L.lib().mtx_minmax_release(self.ctx.handle, ctypes.byref(self.c))

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass


class FilteringOutput:
    """FilteringOutput (src/filteringOutput.h:11-74)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.c = L.CFilteringOutput()

    @property
    def activeBlockCount(self):
        return self.c.activeBlockCount

    # the reference's members (src/filteringOutput.h:12-13)
    @property
    def compactedBlockIdBuffer(self):
        return Buffer(self.c.compacted_block_ids, 4 * self.c.capacity)

    @property
    def activeBlockCountBuffer(self):
        return Buffer(self.c.active_block_count, 4)

    def read_ids(self):
        n = read_device(self.c.active_block_count, np.uint32, 1)[0]
        return read_device(self.c.compacted_block_ids, np.uint32, n)

    def cleanup(self, device=None):
        if self.ctx is not None and self.c.compacted_block_ids:
            L.lib().mtx_filtering_release(self.ctx.handle, ctypes.byref(self.c))

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass


class ExtractionOutput:
    """ExtractionOutput (src/extractionOutput.h:10-122) with RAII cleanup."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.c = L.CExtractionOutput()

    @property
    def vertexCount(self):
        return self.c.vertexCount

    @property
    def indexCount(self):
        return self.c.indexCount

    @property
    def meshletCount(self):
        return self.c.meshletCount

    @property
    def meshletVertexCount(self):
        return self.c.meshletVertexCount

    # the reference's members (src/extractionOutput.h:11-16); the renderer
    # layout the meshlet descriptors refer to is meshletVertexBuffer /
    # meshletIndexBuffer
    @property
    def vertexBuffer(self):
        return Buffer(self.c.vertices, 32 * self.c.vertex_capacity)

    @property
    def indexBuffer(self):
        return Buffer(self.c.indices, 4 * self.c.index_capacity)

    @property
    def meshletDescriptorBuffer(self):
        return Buffer(self.c.meshlets, 16 * self.c.meshlet_capacity)

    @property
    def meshletVertexBuffer(self):
        return Buffer(self.c.meshlet_vertices, 32 * self.c.meshlet_vertex_capacity)

    @property
    def meshletIndexBuffer(self):
        return Buffer(self.c.meshlet_indices, 4 * self.c.index_capacity)

    @property
    def vertexCountBuffer(self):
        return Buffer(self.c.counts, 4)

    @property
    def indexCountBuffer(self):
        return Buffer(self.c.counts + 4 if self.c.counts else 0, 4)

    @property
    def meshletDescriptorCountBuffer(self):
        return Buffer(self.c.counts + 8 if self.c.counts else 0, 4)

    def read_counts(self):
        _check(self.ctx, L.lib().mtx_read_extraction_counts(self.ctx.handle, ctypes.byref(self.c)))
        return self.c.vertexCount, self.c.indexCount, self.c.meshletCount

    def read_vertices(self):
        """(n, 8) float32: position xyzw (NDC, w=1), normal xyzw (w=0)."""
        return read_device(self.c.vertices, np.float32, 8 * self.c.vertexCount).reshape(-1, 8)

    def read_indices(self):
        return read_device(self.c.indices, np.uint32, self.c.indexCount)

    def read_edge_ids(self):
        return read_device(self.c.vertex_edge_ids, np.uint64, self.c.vertexCount)

    def read_t(self):
        """t of each vertex along its owning edge (written with the edge IDs)."""
        return read_device(self.c.vertex_t, np.float32, self.c.vertexCount)

    def read_voxel_positions(self):
        return read_device(self.c.vertex_voxel_pos, np.float32, 3 * self.c.vertexCount).reshape(-1, 3)

    def read_meshlets(self):
        """(m, 4) uint32 meshlet descriptors: vertexOffset, indexOffset, vertexCount, primitiveCount."""
        return read_device(self.c.meshlets, np.uint32, 4 * self.c.meshletCount).reshape(-1, 4)

    def read_meshlet_vertices(self):
        """(s, 8) float32: the renderer-layout vertices the meshlets refer to."""
        return read_device(self.c.meshlet_vertices, np.float32, 8 * self.c.meshletVertexCount).reshape(-1, 8)

    def read_meshlet_vertex_map(self):
        """Per meshlet vertex, its index in the shared vertex buffer."""
        return read_device(self.c.meshlet_vertex_map, np.uint32, self.c.meshletVertexCount)

    def read_meshlet_indices(self):
        """3 per triangle, into the meshlet vertices (same triangle order as read_indices)."""
        return read_device(self.c.meshlet_indices, np.uint32, self.c.indexCount)

    def read_case_indices(self, n_active, block_dim):
        return read_device(self.c.case_indices, np.uint8, n_active * block_dim ** 3)

    def cleanup(self):
        if self.ctx is not None and self.c.workspace:
            L.lib().mtx_extraction_release(self.ctx.handle, ctypes.byref(self.c))

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass


class MtxProfiler:
    """GPU-time regions (the reference's GPUProfiler, src/gpuProfiler.h:44-107),
    recorded with HIP events on the stream each stage runs on."""

    def __init__(self):
        h = ctypes.c_void_p()
        if L.lib().mtx_profiler_create(ctypes.byref(h)) != L.MTX_OK:
            raise MtxError(L.MTX_ENOMEM, "profiler")
        self.handle = h

    def begin(self, name, stream=None):
        L.lib().mtx_profiler_begin(self.handle, name.encode(), _stream(stream))

    def end(self, stream=None):
        L.lib().mtx_profiler_end(self.handle, _stream(stream))

    def results(self):
        """[(region name, milliseconds)] in begin order (blocks on the GPU)."""
        out, i = [], 0
        buf = ctypes.create_string_buffer(128)
        ms = ctypes.c_float()
        while L.lib().mtx_profiler_region(self.handle, i, buf, 128, ctypes.byref(ms)) == L.MTX_OK:
            out.append((buf.value.decode(), ms.value))
            i += 1
        return out

    def reset(self):
        L.lib().mtx_profiler_reset(self.handle)

    def __del__(self):
        try:
            L.lib().mtx_profiler_destroy(self.handle)
        except Exception:
            pass


def computeMinMaxMip(context, volume, pushConstants, stream=None, profiler=None, options=None):
    out = MinMaxOutput(context)
    v = volume._c()
    o = options if options is not None else makeOptions()
    build = o.build_pyramid
    leaf = L.Options.from_buffer_copy(o)
    leaf.build_pyramid = 0
    if profiler:
        profiler.begin("MinMax_Leaf_Pass", stream)
    st = L.lib().mtx_compute_min_max(context.handle, ctypes.byref(v), ctypes.byref(pushConstants),
                                     ctypes.byref(leaf), _stream(stream), ctypes.byref(out.c))
    if profiler:
        profiler.end(stream)
    _check(context, st)
    if build and out.c.num_levels > 1:
        if profiler:
            profiler.begin("MinMax_Octree_Reduction", stream)
        st = L.lib().mtx_build_pyramid(context.handle, ctypes.byref(out.c), _stream(stream))
        if profiler:
            profiler.end(stream)
        _check(context, st)
    out._volume_ref = volume
    return out


def filterActiveBlocks(context, minMaxOutput, pushConstants, stream=None, profiler=None,
                       options=None, out=None):
    out = out if out is not None else FilteringOutput(context)
    if profiler:
        profiler.begin("Active_Block_Filtering", stream)
    st = L.lib().mtx_filter_active_blocks(context.handle, ctypes.byref(minMaxOutput.c),
                                          ctypes.byref(pushConstants),
                                          ctypes.byref(options) if options is not None else None,
                                          _stream(stream), ctypes.byref(out.c))
    if profiler:
        profiler.end(stream)
    _check(context, st)
    return out


def readActiveBlockCount(context, filteringOutput):
    _check(context, L.lib().mtx_read_active_count(context.handle, ctypes.byref(filteringOutput.c)))


def extractMeshletDescriptors(context, minMaxOutput, filterOutput, pushConstants, stream=None,
                              profiler=None, options=None, out=None):
    out = out if out is not None else ExtractionOutput(context)
    if profiler:
        profiler.begin("Mesh_Extraction", stream)
    st = L.lib().mtx_extract(context.handle, ctypes.byref(minMaxOutput.c), ctypes.byref(filterOutput.c),
                             ctypes.byref(pushConstants),
                             ctypes.byref(options) if options is not None else None,
                             _stream(stream), ctypes.byref(out.c))
    if profiler:
        profiler.end(stream)
    _check(context, st)
    return out


def writeGPUExtractionToOBJ(context, extractionResult, filePath):
    """OBJ as src/extractionTestUtils.cpp:328-366 writes it (v, vn, f a//a, 1-based)."""
    _check(context, L.lib().mtx_write_obj(context.handle, ctypes.byref(extractionResult.c),
                                          str(filePath).encode()))


def key_f32(x):
    return L.lib().mtx_key_f32(x)


def unkey_f32(k):
    return L.lib().mtx_unkey_f32(k)
