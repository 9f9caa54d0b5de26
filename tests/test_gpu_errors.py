"""Failure paths of the single-volume C ABI, on the GPU.

Each call below is rejected on the host before any launch, with the status
that replaces the reference's throw or assert at that point
(src/extractionManager.cpp:276,313-338, src/common.h:9-14): MTX_EINVAL for a
bad argument, MTX_ESTATE for inputs that do not belong together, MTX_ENOMEM
for a request the device cannot hold, MTX_ECAPACITY for reuse-mode buffers
that are too small.  Also here: the fp32 interpolation rule for tiny gaps
(t = 0.5 unless |v1 - v0| > 1e-5, src/shaders/marching_cubes_pmb.mesh.glsl:129-134),
subnormal values included, against the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle
from meshtrex_amd import _lib as L
from meshtrex_amd import (DeviceBuffer, ExtractionOutput, FilteringOutput, MinMaxOutput, Volume,
                          computeMinMaxMip, extractMeshletDescriptors, filterActiveBlocks, makeOptions,
                          makePushConstants)
from meshtrex_amd.managers import MtxError

pytestmark = pytest.mark.gpu


def blobs(dims, seed=11):
    return oracle.gen_blobs_u8(dims, oracle.SEED_BASE + seed, 8)


def cvol(buf, dims, dtype=L.MTX_U8, on_device=1):
    v = L.CVolume()
    v.data = buf.ptr if buf is not None else None
    v.dims[:] = dims
    v.dtype = dtype
    v.on_device = on_device
    return v


def mm_call(ctx, v, pc, opt=None, out=None):
    out = out if out is not None else MinMaxOutput(ctx)
    st = L.lib().mtx_compute_min_max(ctx.handle, ctypes.byref(v) if v is not None else None,
                                     ctypes.byref(pc) if pc is not None else None,
                                     ctypes.byref(opt) if opt is not None else None, None,
                                     ctypes.byref(out.c) if out is not None else None)
    return st, out


@pytest.fixture(scope="module")
def vol_a(ctx):
    dims = (40, 36, 33)
    data = blobs(dims)
    return dims, data, DeviceBuffer.from_numpy(data)


# ---- MTX_EINVAL -------------------------------------------------------------------

@pytest.mark.parametrize("B", [0, 17])
def test_block_edge_out_of_range(ctx, vol_a, B):
    dims, _, buf = vol_a
    pc = makePushConstants(dims, B, 128.0)
    st, _ = mm_call(ctx, cvol(buf, dims), pc)
    assert st == L.MTX_EINVAL


def test_non_cubic_blocks(ctx, vol_a):
    dims, _, buf = vol_a
    pc = makePushConstants(dims, 8, 128.0)
    pc.blockDim[1] = 4
    st, _ = mm_call(ctx, cvol(buf, dims), pc)
    assert st == L.MTX_EINVAL
    assert b"cubic" in L.lib().mtx_context_last_error(ctx.handle)


def test_invalid_dims(ctx, vol_a):
    dims, _, buf = vol_a
    bad = (0, dims[1], dims[2])
    st, _ = mm_call(ctx, cvol(buf, bad), makePushConstants(bad, 8, 128.0))
    assert st == L.MTX_EINVAL
    huge = (1 << 17, 1 << 17, 8)    # X * Y beyond 32 bits
    st, _ = mm_call(ctx, cvol(buf, huge), makePushConstants(huge, 8, 128.0))
    assert st == L.MTX_EINVAL


def test_push_constants_that_do_not_match(ctx, vol_a):
    dims, _, buf = vol_a
    pc = makePushConstants((dims[0] + 1, dims[1], dims[2]), 8, 128.0)
    st, _ = mm_call(ctx, cvol(buf, dims), pc)
    assert st == L.MTX_EINVAL
    pc = makePushConstants(dims, 8, 128.0)
    pc.blockGridDim[2] += 1
    st, _ = mm_call(ctx, cvol(buf, dims), pc)
    assert st == L.MTX_EINVAL


def test_null_pointers(ctx, vol_a):
    dims, _, buf = vol_a
    pc = makePushConstants(dims, 8, 128.0)
    assert mm_call(ctx, None, pc)[0] == L.MTX_EINVAL
    assert mm_call(ctx, cvol(buf, dims), None)[0] == L.MTX_EINVAL
    v = cvol(buf, dims)
    v.data = None
    assert mm_call(ctx, v, pc)[0] == L.MTX_EINVAL
    assert L.lib().mtx_compute_min_max(None, ctypes.byref(cvol(buf, dims)), ctypes.byref(pc), None, None,
                                       None) == L.MTX_EINVAL
    mm = computeMinMaxMip(ctx, Volume(dims, "uint8", buf), pc)
    f = filterActiveBlocks(ctx, mm, pc)
    assert L.lib().mtx_extract(ctx.handle, ctypes.byref(mm.c), ctypes.byref(f.c), ctypes.byref(pc), None, None,
                               None) == L.MTX_EINVAL
    assert L.lib().mtx_filter_active_blocks(ctx.handle, ctypes.byref(mm.c), None, None, None,
                                            ctypes.byref(f.c)) == L.MTX_EINVAL


def test_bad_options(ctx, vol_a):
    dims, _, buf = vol_a
    pc = makePushConstants(dims, 8, 128.0)
    assert mm_call(ctx, cvol(buf, dims), pc, makeOptions(overlap=2))[0] == L.MTX_EINVAL
    assert mm_call(ctx, cvol(buf, dims), pc, makeOptions(minmax_rows=3))[0] == L.MTX_EINVAL
    v = cvol(buf, dims)
    v.dtype = 7
    assert mm_call(ctx, v, pc)[0] == L.MTX_EINVAL
    mm = computeMinMaxMip(ctx, Volume(dims, "uint8", buf), pc)
    f = FilteringOutput(ctx)
    assert L.lib().mtx_filter_active_blocks(ctx.handle, ctypes.byref(mm.c), ctypes.byref(pc),
                                            ctypes.byref(makeOptions(rule=5)), None,
                                            ctypes.byref(f.c)) == L.MTX_EINVAL


# ---- MTX_ESTATE -------------------------------------------------------------------

def test_extract_needs_overlap_one(ctx, vol_a):
    dims, _, buf = vol_a
    pc = makePushConstants(dims, 8, 128.0)
    mm = computeMinMaxMip(ctx, Volume(dims, "uint8", buf), pc, options=makeOptions(overlap=0))
    f = filterActiveBlocks(ctx, mm, pc)
    with pytest.raises(MtxError) as e:
        extractMeshletDescriptors(ctx, mm, f, pc)
    assert e.value.status == L.MTX_ESTATE


def test_filter_of_another_grid_is_rejected(ctx, vol_a):
    """A filter made for another block grid: rejected by extraction and the
    adjoint before any launch (its list and map do not describe these blocks)."""
    dims, _, buf = vol_a
    pc = makePushConstants(dims, 8, 128.0)
    mm = computeMinMaxMip(ctx, Volume(dims, "uint8", buf), pc)
    f = filterActiveBlocks(ctx, mm, pc)
    ex = extractMeshletDescriptors(ctx, mm, f, pc)
    dims_b = (72, 64, 50)
    data_b = blobs(dims_b, 12)
    pc_b = makePushConstants(dims_b, 8, 128.0)
    mm_b = computeMinMaxMip(ctx, Volume(dims_b, "uint8", DeviceBuffer.from_numpy(data_b)), pc_b)
    f_b = filterActiveBlocks(ctx, mm_b, pc_b)
    for mmx, fx, pcx in ((mm, f_b, pc), (mm_b, f, pc_b)):
        with pytest.raises(MtxError) as e:
            extractMeshletDescriptors(ctx, mmx, fx, pcx)
        assert e.value.status == L.MTX_ESTATE
        assert "block grid" in str(e.value)
    # same grid, other isovalue
    pc_iso = makePushConstants(dims, 8, 100.0)
    with pytest.raises(MtxError) as e:
        extractMeshletDescriptors(ctx, mm, f, pc_iso)
    assert e.value.status == L.MTX_ESTATE
    # the adjoint: a filter or an extraction of another grid
    dvol = DeviceBuffer(4 * int(np.prod(dims)))
    diso = DeviceBuffer(4)
    dpos = DeviceBuffer(12 * max(ex.vertexCount, 1))
    for mmx, fx, exx, pcx in ((mm, f_b, ex, pc), (mm_b, f_b, ex, pc_b)):
        st = L.lib().mtx_adjoint(ctx.handle, ctypes.byref(mmx.c), ctypes.byref(fx.c), ctypes.byref(exx.c),
                                 ctypes.byref(pcx), dpos.ptr, dvol.ptr, diso.ptr, None)
        assert st == L.MTX_ESTATE


def test_empty_outputs_are_rejected(ctx, vol_a):
    dims, _, buf = vol_a
    pc = makePushConstants(dims, 8, 128.0)
    mm, f, ex = MinMaxOutput(ctx), FilteringOutput(ctx), ExtractionOutput(ctx)
    assert L.lib().mtx_filter_active_blocks(ctx.handle, ctypes.byref(mm.c), ctypes.byref(pc), None, None,
                                            ctypes.byref(f.c)) == L.MTX_ESTATE
    assert L.lib().mtx_read_active_count(ctx.handle, ctypes.byref(f.c)) == L.MTX_ESTATE
    assert L.lib().mtx_read_extraction_counts(ctx.handle, ctypes.byref(ex.c)) == L.MTX_ESTATE
    assert L.lib().mtx_build_pyramid(ctx.handle, ctypes.byref(mm.c), None) == L.MTX_ESTATE


# ---- MTX_ENOMEM, MTX_ECAPACITY ----------------------------------------------------

def test_oversize_extraction_is_enomem(ctx):
    """An extraction whose workspace no device can hold (2^31 blocks of 16^3
    cells, ~44 TB): the allocation fails and MTX_ENOMEM comes back before any
    launch.  The min/max and filter outputs are descriptions only here (their
    pointers are never dereferenced on this path)."""
    B, G = 16, (2048, 2048, 512)
    dims = tuple(g * B for g in G)
    mm = L.CMinMaxOutput()
    mm.minmax = 0x1000
    mm.volume = 0x1000
    mm.dtype = L.MTX_U8
    mm.dims[:] = dims
    mm.block_dim = B
    mm.grid[:] = G
    mm.overlap = 1
    f = L.CFilteringOutput()
    f.compacted_block_ids = 0x1000
    f.active_block_count = 0x1000
    f.block_to_active = 0x1000
    f.capacity = 0xFFFFFFFF
    f.block_dim = B
    f.grid[:] = G
    f.isovalue = 1.0
    pc = L.PushConstants()
    pc.volumeDim[:3] = dims
    pc.blockDim[:3] = (B, B, B)
    pc.blockGridDim[:3] = G
    pc.isovalue = 1.0
    ex = L.CExtractionOutput()
    st = L.lib().mtx_extract(ctx.handle, ctypes.byref(mm), ctypes.byref(f), ctypes.byref(pc), None, None,
                             ctypes.byref(ex))
    assert st == L.MTX_ENOMEM
    assert not ex.workspace and not ex.vertices
    # the context stays usable: a real extraction still works afterwards
    dims = (24, 20, 18)
    data = blobs(dims, 13)
    pcs = makePushConstants(dims, 8, 128.0)
    m2 = computeMinMaxMip(ctx, Volume(dims, "uint8", DeviceBuffer.from_numpy(data)), pcs)
    f2 = filterActiveBlocks(ctx, m2, pcs)
    e2 = extractMeshletDescriptors(ctx, m2, f2, pcs)
    assert e2.read_counts()[0] == len(oracle.extract(data, dims, 8, 128.0, normals=False)["edge_ids"])


def test_reuse_capacity_overflow_is_reported(ctx):
    """Reuse-mode buffers sized for a sparse volume, then a denser one: the
    overflow is flagged on the device and reported as MTX_ECAPACITY."""
    dims = (48, 48, 48)
    sparse = np.zeros(int(np.prod(dims)), np.uint8)
    sparse[24 + 48 * (24 + 48 * 24)] = 255
    dense = blobs(dims, 14)
    pc = makePushConstants(dims, 8, 128.0)
    vb = DeviceBuffer.from_numpy(sparse)
    v = Volume(dims, "uint8", vb)
    mm = computeMinMaxMip(ctx, v, pc)
    f = filterActiveBlocks(ctx, mm, pc)
    ex = extractMeshletDescriptors(ctx, mm, f, pc)
    assert ex.read_counts()[0] == 6
    L.lib().mtx_memcpy_h2d(vb.ptr, dense.ctypes.data, dense.nbytes)
    mm = computeMinMaxMip(ctx, v, pc)
    f = filterActiveBlocks(ctx, mm, pc)
    extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(reuse=1), out=ex)
    with pytest.raises(MtxError) as e:
        ex.read_counts()
    assert e.value.status == L.MTX_ECAPACITY


# ---- t on tiny gaps ---------------------------------------------------------------

def test_tiny_and_subnormal_gaps_follow_the_reference_rule(ctx):
    """Active edges whose values differ by at most 1e-5 (subnormal values
    among them) get t = 0.5, as the reference's shader gives them; wider gaps
    the quotient.  Owning edge IDs and t against the oracle, which applies the
    same rule in fp32 for the bound and FP64 for the quotient."""
    dims = (12, 10, 9)
    rng = np.random.default_rng(5)
    vol = rng.uniform(0.5, 1.0, int(np.prod(dims))).astype(np.float32)  # all above iso 0
    pairs = [(-4e-6, 4e-6), (-1e-40, 1e-40), (-1e-45, 2e-45), (-6e-6, 5e-6), (-2e-5, 1e-5),
             (-1e-3, 2e-3), (0.0, 3e-6), (-3e-6, 1.2e-5)]
    for k, (a, b) in enumerate(pairs):
        x, y, z = 1 + k, 2 + (k % 3), 3 + (k % 4)
        i = x + dims[0] * (y + dims[1] * z)
        vol[i], vol[i + 1] = np.float32(a), np.float32(b)
    pc = makePushConstants(dims, 8, 0.0)
    mm = computeMinMaxMip(ctx, Volume(dims, "float32", DeviceBuffer.from_numpy(vol)), pc)
    f = filterActiveBlocks(ctx, mm, pc)
    ex = extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(edge_ids=1))
    ex.read_counts()
    ref = oracle.extract(vol, dims, 8, 0.0, normals=False)
    np.testing.assert_array_equal(ex.read_edge_ids(), ref["edge_ids"])
    t = ex.read_t().astype(np.float64)
    assert np.all(np.abs(t - ref["t"]) <= 1e-6)
    # the guarded edges: exactly 0.5
    e = ref["edge_ids"].astype(np.int64)
    vox, ax = e // 3, e % 3
    st = np.array([1, dims[0], dims[0] * dims[1]])[ax]
    d32 = (vol[vox + st] - vol[vox]).astype(np.float32)
    guarded = np.abs(d32) <= np.float32(1e-5)
    assert guarded.sum() == 4     # the four planted pairs with |v1 - v0| <= 1e-5
    assert np.all(t[guarded] == 0.5)
    assert np.all(ref["t"][guarded] == 0.5)
