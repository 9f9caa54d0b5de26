"""mtx_filter_active_blocks_noalloc: the filter into caller buffers, with its
scratch in the library's static device array, against
mtx_filter_active_blocks (ids, count and map bit-exact) on two calls in a row,
the second with a smaller volume, on both compaction paths (the fused
activity mask and the min/max pairs), synchronous and on a caller stream; a
grid beyond the scratch is rejected before any launch."""
import ctypes

import numpy as np
import pytest

import oracle
from meshtrex_amd import _lib as L
from meshtrex_amd import (DeviceBuffer, Volume, computeMinMaxMip, filterActiveBlocks, makeOptions,
                          makePushConstants, readActiveBlockCount)
from meshtrex_amd.managers import read_device

pytestmark = pytest.mark.gpu


def noalloc(ctx, mm, pc, opt, stream, ids, count, bmap):
    return L.lib().mtx_filter_active_blocks_noalloc(ctx.handle, ctypes.byref(mm.c), ctypes.byref(pc),
                                                    ctypes.byref(opt) if opt is not None else None, stream,
                                                    ids.ptr, count.ptr, bmap.ptr)


@pytest.mark.parametrize("use_mask", [True, False])
@pytest.mark.parametrize("on_stream", [False, True])
def test_noalloc_filter_matches_filter(ctx, use_mask, on_stream):
    big, small = (150, 131, 97), (41, 33, 29)
    nb_big = int(np.prod([(d + 7) // 8 for d in big]))
    ids = DeviceBuffer(4 * nb_big)
    count = DeviceBuffer(4)
    bmap = DeviceBuffer(4 * nb_big)
    stream = None
    if on_stream:
        s = ctypes.c_void_p()
        assert L.lib().mtx_stream_create(ctypes.byref(s)) == L.MTX_OK
        stream = s.value
    try:
        for k, dims in enumerate((big, small)):
            data = oracle.gen_blobs_u8(dims, oracle.SEED_BASE + 40 + k, 16)
            pc = makePushConstants(dims, 8, 128.0)
            mm = computeMinMaxMip(ctx, Volume(dims, "uint8", DeviceBuffer.from_numpy(data)), pc)
            # the pairs path: a filter whose isovalue or rule is not the fused mask's
            opt = makeOptions() if use_mask else makeOptions(rule=L.MTX_RULE_GPU)
            f = filterActiveBlocks(ctx, mm, pc, options=opt)
            readActiveBlockCount(ctx, f)
            # leftovers of the previous call must not matter
            L.lib().mtx_memcpy_h2d(count.ptr, np.array([0xDEADBEEF], np.uint32).ctypes.data, 4)
            assert noalloc(ctx, mm, pc, opt, stream, ids, count, bmap) == L.MTX_OK
            if stream:
                assert L.lib().mtx_stream_synchronize(stream) == L.MTX_OK
            nb = int(np.prod([(d + 7) // 8 for d in dims]))
            n = int(read_device(count.ptr, np.uint32, 1)[0])
            assert n == f.activeBlockCount
            np.testing.assert_array_equal(read_device(ids.ptr, np.uint32, n), f.read_ids())
            np.testing.assert_array_equal(read_device(bmap.ptr, np.uint32, nb),
                                          read_device(f.c.block_to_active, np.uint32, nb))
            rule = L.MTX_RULE_CPU if use_mask else L.MTX_RULE_GPU
            want = oracle.filter_ids(oracle.minmax(data, dims, 8), oracle.U8, 128.0, rule=rule)
            np.testing.assert_array_equal(f.read_ids(), want)
            f.cleanup()
            mm.cleanup()
    finally:
        if stream:
            L.lib().mtx_stream_destroy(stream)


def test_noalloc_filter_rejects_grid_beyond_scratch(ctx):
    """2^30 blocks from pairs need 2^19 compaction tiles, twice the static
    scratch: MTX_EINVAL before any launch (the output pointers are never used)."""
    mm = L.CMinMaxOutput()
    mm.minmax = 0x1000
    mm.volume = 0x1000
    mm.dims[:] = (1024, 1024, 1024)
    mm.block_dim = 1
    mm.grid[:] = (1024, 1024, 1024)
    mm.overlap = 1
    pc = makePushConstants((1024, 1024, 1024), 1, 0.5)

    class Fake:
        c = mm
    ids = count = bmap = DeviceBuffer(4)
    assert noalloc(ctx, Fake, pc, None, None, ids, count, bmap) == L.MTX_EINVAL
    assert b"scratch" in L.lib().mtx_context_last_error(ctx.handle)
    assert L.lib().mtx_filter_active_blocks_noalloc(ctx.handle, ctypes.byref(mm), ctypes.byref(pc), None, None,
                                                    None, count.ptr, bmap.ptr) == L.MTX_EINVAL
