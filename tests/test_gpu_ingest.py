"""Volume ingest: a host volume (pageable or pinned) uploads in two slabs of
whole block planes (a head, and a tail of 1/50 of the block planes, at least
one; volumes of fewer than 4 block planes in one), each slab's min/max
running as soon as it is resident -- with blocking copies for a synchronous
call, on the context's copy stream for a stream-ordered one.  The result must
equal the min/max of the same volume already on the device, and the oracle's
pairs and active list, bit for bit, for volumes with odd dimensions."""
import numpy as np
import pytest

import oracle
from meshtrex_amd import hiprt
from meshtrex_amd import (DeviceBuffer, Volume, computeMinMaxMip, extractMeshletDescriptors,
                          filterActiveBlocks, loadVolume, makeOptions, makePushConstants, pinned_array,
                          read_device)

pytestmark = pytest.mark.gpu


def minmax_of(ctx, vol, dims, iso, stream=None):
    pc = makePushConstants(dims, 8, iso)
    mm = computeMinMaxMip(ctx, vol, pc, stream=stream)
    f = filterActiveBlocks(ctx, mm, pc, stream=stream)
    if stream is not None:
        hiprt.stream_sync(stream)
    pairs = mm.read_pairs(0)
    mask = read_device(mm.c.active_mask, np.uint32, mm.c.mask_words_per_row * mm.c.grid[1] * mm.c.grid[2])
    ids = f.read_ids()
    f.cleanup()
    mm.cleanup()
    return pairs, mask, ids


@pytest.mark.parametrize("dims,dtype", [((421, 403, 399), np.float32), ((701, 651, 653), np.uint8),
                                        ((77, 53, 20), np.float32), ((45, 38, 16), np.uint8)])
@pytest.mark.parametrize("pinned,on_stream", [(False, False), (True, False), (True, True)])
def test_slab_upload_equals_device_resident(ctx, dims, dtype, pinned, on_stream):
    n = int(np.prod(dims))
    rng = np.random.default_rng(11)
    if dtype == np.uint8:
        src = rng.integers(0, 256, n, dtype=np.uint8)
        iso = 128.0
    else:
        src = rng.standard_normal(n).astype(np.float32)
        iso = 1.5
    tname = "uint8" if dtype == np.uint8 else "float32"
    host = src
    if pinned:
        host = pinned_array(n, dtype)
        host[:] = src
    dev = DeviceBuffer.from_numpy(src)
    ref = minmax_of(ctx, Volume(dims, tname, dev), dims, iso)
    stream = hiprt.stream_create() if on_stream else None
    got = minmax_of(ctx, Volume(dims, tname, host), dims, iso, stream)
    if stream is not None:
        hiprt.stream_destroy(stream)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    dev.free()
    # and the oracle's pairs and active list of the same volume
    pairs = oracle.minmax(src, dims, 8)
    np.testing.assert_array_equal(got[0], pairs)
    np.testing.assert_array_equal(got[2], oracle.filter_ids(pairs, oracle.U8 if dtype == np.uint8 else oracle.F32,
                                                            iso))


def test_pinned_load_volume_extracts_the_same(ctx, tmp_path):
    dims = (45, 38, 29)
    x, y, z = np.meshgrid(np.arange(dims[0]), np.arange(dims[1]), np.arange(dims[2]), indexing="ij")
    r = np.sqrt((x - 22.0) ** 2 + (y - 19.0) ** 2 + (z - 14.0) ** 2)
    data = np.clip(255 - 12 * r, 0, 255).astype(np.uint8).transpose(2, 1, 0).copy()
    path = tmp_path / ("sphere_%dx%dx%d_uint8.raw" % dims)
    data.tofile(path)
    out = []
    for pinned in (False, True):
        vol = loadVolume(path, pinned=pinned)
        pc = makePushConstants(dims, 8, 128.0)
        mm = computeMinMaxMip(ctx, vol, pc)
        f = filterActiveBlocks(ctx, mm, pc)
        ex = extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(edge_ids=1))
        ex.read_counts()
        out.append((ex.read_edge_ids(), ex.read_indices()))
        ex.cleanup()
        f.cleanup()
        mm.cleanup()
    assert out[0][0].size > 0
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
