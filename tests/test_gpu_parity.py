"""GPU parity: the HIP path through the C ABI vs the CPU oracle, same inputs.

Bit-exact: block min/max (every level), active-block lists, case indices,
vertex/triangle counts, owning edge IDs, index buffers.  FP32 vertex data vs
the oracle's FP64: t within 1e-6 (edge-local voxel units), NDC position
within 2e-6, normals within 1e-4 rad.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from meshtrex_amd import (computeMinMaxMip, extractMeshletDescriptors, filterActiveBlocks, makeOptions,
                          makePushConstants, readActiveBlockCount)
from parity_checks import check_extraction, check_meshlets, extract_both, rand_volume, run_minmax, vol_of

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


# ---- block min/max -------------------------------------------------------------

# x extents of every residue mod 4: unaligned rows are staged through LDS and
# the lane at the volume's last row end takes a shift of 1, 2 or 3 voxels
FAST_DIMS = [(64, 48, 40), (61, 33, 17), (300, 20, 26), (256, 9, 8), (8, 8, 8), (1, 1, 1),
             (1030, 3, 5), (2, 2, 2), (63, 17, 9), (7, 5, 3)]


@pytest.mark.parametrize("dims", FAST_DIMS)
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_minmax_b8_vs_oracle(ctx, dims, dtype):
    # rows 0: the planner's choice (small volumes: the plane-split kernel);
    # 1, 2: the block-column kernel with that many y-blocks per wave
    data = rand_volume(dims, dtype, 1)
    want = oracle.minmax(data, dims, 8, overlap=1)
    for rows in (0, 1, 2):
        mm, _ = run_minmax(ctx, data, dims, 8, minmax_rows=rows)
        np.testing.assert_array_equal(mm.read_pairs(0), want, err_msg="rows %d" % rows)


@pytest.mark.parametrize("B", [1, 2, 3, 4, 5, 7, 16])
@pytest.mark.parametrize("overlap", [0, 1])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_minmax_generic_vs_oracle(ctx, B, overlap, dtype):
    dims = (37, 29, 23)
    data = rand_volume(dims, dtype, 2 + B)
    mm, _ = run_minmax(ctx, data, dims, B, overlap=overlap)
    np.testing.assert_array_equal(mm.read_pairs(0), oracle.minmax(data, dims, B, overlap=overlap))


def test_minmax_legacy_matches_reference_cpu_golden(ctx):
    """GPU min/max in the reference CPU oracle's mode (B=4, no overlap) on the
    256^3 blob volume reproduces computeMinMaxFromFile's output byte for byte."""
    g = json.load(open(os.path.join(HERE, "golden", "reference_cpu_golden.json")))
    dims = (256, 256, 256)
    data = oracle.gen_blobs_u8(dims, g["volume"]["seed"], 16)
    assert hashlib.sha256(data.tobytes()).hexdigest() == g["volume"]["sha256"]
    mm, _ = run_minmax(ctx, data, dims, 4, overlap=0)
    pairs = mm.read_pairs(0)
    assert hashlib.sha256(pairs.tobytes()).hexdigest() == g["minmax_b4_overlap0"]["sha256"]


@pytest.mark.parametrize("dims,B", [((64, 48, 40), 8), ((200, 136, 72), 8), ((37, 29, 23), 3),
                                    ((130, 17, 9), 1)])
def test_pyramid_vs_oracle(ctx, dims, B):
    data = rand_volume(dims, np.float32, 5)
    mm, _ = run_minmax(ctx, data, dims, B)
    leaf = oracle.minmax(data, dims, B)
    allp, off, ext = oracle.pyramid(leaf, oracle.grid(dims, B))
    assert mm.num_levels == len(ext)
    for lev in range(mm.num_levels):
        assert mm.level_extent(lev) == ext[lev]
        np.testing.assert_array_equal(mm.read_pairs(lev), allp[off[lev]:off[lev + 1]])


# ---- filter --------------------------------------------------------------------

ISOS_U8 = [-3.0, 0.0, 80.0, 80.4, 80.6, 127.5, 128.0, 255.0, 300.0]
ISOS_F32 = [-1e30, -0.5, -0.0, 0.0, 0.25, 1.5, 1e30]


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("fuse", [0, 1])
def test_filter_u8_vs_oracle(ctx, rule, fuse):
    dims = (96, 70, 66)
    data = oracle.gen_blobs_u8(dims, 11, 6)
    for iso in ISOS_U8:
        mm, pc = run_minmax(ctx, data, dims, 8, iso=iso, rule=rule, fuse_filter=fuse)
        f = filterActiveBlocks(ctx, mm, pc, options=makeOptions(rule=rule))
        readActiveBlockCount(ctx, f)
        want = oracle.filter_ids(oracle.minmax(data, dims, 8), oracle.U8, iso, rule)
        assert f.activeBlockCount == want.size, iso
        np.testing.assert_array_equal(f.read_ids(), want)


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("fuse", [0, 1])
def test_filter_f32_vs_oracle(ctx, rule, fuse):
    dims = (130, 41, 35)
    data = oracle.gen_sdf_f32(dims, 21, 6, 3, 0.2)
    for iso in ISOS_F32:
        want = oracle.filter_ids(oracle.minmax(data, dims, 8), oracle.F32, iso, rule)
        for rows in (0, 1):   # the fused mask from both min/max kernels
            mm, pc = run_minmax(ctx, data, dims, 8, iso=iso, rule=rule, fuse_filter=fuse, minmax_rows=rows)
            f = filterActiveBlocks(ctx, mm, pc, options=makeOptions(rule=rule))
            readActiveBlockCount(ctx, f)
            assert f.activeBlockCount == want.size, (iso, rows)
            np.testing.assert_array_equal(f.read_ids(), want)


def test_filter_generic_block_and_large_grid(ctx):
    # many filter tiles (look-back across workgroups): 1x1x1 blocks
    dims = (96, 96, 64)
    data = rand_volume(dims, np.uint8, 9)
    for rule in (0, 1):
        mm, pc = run_minmax(ctx, data, dims, 1, iso=100.0)
        f = filterActiveBlocks(ctx, mm, pc, options=makeOptions(rule=rule))
        readActiveBlockCount(ctx, f)
        want = oracle.filter_ids(oracle.minmax(data, dims, 1), oracle.U8, 100.0, rule)
        np.testing.assert_array_equal(f.read_ids(), want)


# ---- extraction ---------------------------------------------------------------

@pytest.mark.parametrize("B", [8, 4, 3, 16])
def test_extract_blobs_u8(ctx, B):
    dims = (72, 64, 57)
    data = oracle.gen_blobs_u8(dims, 31, 8)
    for iso in (128.0, 80.5):
        check_extraction(ctx, data, dims, B, iso)


@pytest.mark.parametrize("dims", [(64, 64, 64), (65, 33, 17), (41, 70, 9)])
def test_extract_sdf_f32(ctx, dims):
    data = oracle.gen_sdf_f32(dims, 41, 6, 3, 0.15)
    check_extraction(ctx, data, dims, 8, 0.0)


def test_extract_noise_all_cases(ctx):
    # random u8 noise hits (almost) every case and many neighbour-block owners
    dims = (33, 34, 35)
    data = rand_volume(dims, np.uint8, 77)
    check_extraction(ctx, data, dims, 8, 127.5)
    check_extraction(ctx, data, dims, 5, 100.0)


def test_extract_gpu_rule_same_geometry(ctx):
    dims = (48, 40, 44)
    data = oracle.gen_blobs_u8(dims, 51, 5)
    check_extraction(ctx, data, dims, 8, 100.0, rule=1)


def test_extract_empty_and_degenerate(ctx):
    dims = (40, 40, 40)
    data = oracle.gen_blobs_u8(dims, 61, 4)
    for iso in (-1.0, 255.0, 1000.0):
        _, f, ex, ref = extract_both(ctx, data, dims, 8, iso)
        assert ex.vertexCount == 0 and ex.indexCount == 0 and len(ref["tris"]) == 0
    for d in [(1, 20, 20), (20, 1, 1), (2, 2, 2)]:
        check_extraction(ctx, rand_volume(d, np.uint8, 3), d, 8, 127.5)


def test_extract_single_hot_voxel(ctx):
    """Known answer: one interior voxel 255 in a field of 0, iso 128: the 8
    cells around it give one triangle each, on 6 vertices at t = 127/255 from
    the hot voxel."""
    dims = (17, 17, 17)
    data = np.zeros(dims[::-1], np.uint8)
    data[9, 8, 7] = 255  # z, y, x
    mm, f, ex, ref = extract_both(ctx, data.reshape(-1), dims, 8, 128.0)
    assert ex.vertexCount == 6 and ex.indexCount == 24
    vp = ex.read_voxel_positions().astype(np.float64)
    d = np.linalg.norm(vp - np.array([7.0, 8.0, 9.0]), axis=1)
    np.testing.assert_allclose(d, 127.0 / 255.0, atol=1e-6)


def test_extract_host_volume_upload(ctx):
    dims = (40, 36, 30)
    data = oracle.gen_blobs_u8(dims, 71, 5)
    pc = makePushConstants(dims, 8, 128.0)
    mm = computeMinMaxMip(ctx, vol_of(data, dims, device=False), pc)
    f = filterActiveBlocks(ctx, mm, pc)
    ex = extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(edge_ids=1))
    ex.read_counts()
    ref = oracle.extract(data, dims, 8, 128.0, normals=False)
    np.testing.assert_array_equal(ex.read_edge_ids(), ref["edge_ids"])


def test_extract_reuse_mode_and_capacity(ctx):
    dims = (64, 64, 64)
    data = oracle.gen_blobs_u8(dims, 81, 8)
    mm, pc = run_minmax(ctx, data, dims, 8, iso=128.0)
    f = filterActiveBlocks(ctx, mm, pc)
    ex = extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(edge_ids=1))
    ex.read_counts()
    first = ex.read_indices().copy()
    # second run into the same buffers: identical output, no allocation
    extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(edge_ids=1, reuse=1), out=ex)
    ex.read_counts()
    np.testing.assert_array_equal(ex.read_indices(), first)
    # a lower isovalue needs more room than the buffers have: reported, not written past
    pc2 = makePushConstants(dims, 8, 40.0)
    mm2 = computeMinMaxMip(ctx, vol_of(data, dims), pc2)
    f2 = filterActiveBlocks(ctx, mm2, pc2)
    ref = oracle.extract_count(data, dims, 8, 40.0)
    if ref[0] > ex.c.vertex_capacity or 3 * ref[1] > ex.c.index_capacity:
        extractMeshletDescriptors(ctx, mm2, f2, pc2, options=makeOptions(edge_ids=1, reuse=1), out=ex)
        from meshtrex_amd import MtxError
        with pytest.raises(MtxError, match="MTX_ECAPACITY"):
            ex.read_counts()


def test_block_map(ctx):
    """Dense block -> active-index map, from the fused mask and from the pairs."""
    dims = (90, 77, 61)
    data = oracle.gen_blobs_u8(dims, 91, 10)
    B, iso = 8, 128.0
    for fuse in (1, 0):
        mm, pc = run_minmax(ctx, data, dims, B, iso=iso, fuse_filter=fuse)
        f = filterActiveBlocks(ctx, mm, pc)
        ids = f.read_ids()
        G = oracle.grid(dims, B)
        nb = G[0] * G[1] * G[2]
        from meshtrex_amd.managers import read_device
        m = read_device(f.c.block_to_active, np.uint32, nb)
        want = np.full(nb, 0xFFFFFFFF, np.uint32)
        want[ids] = np.arange(ids.size, dtype=np.uint32)
        np.testing.assert_array_equal(m, want)


def test_meshlets_off_and_reuse(ctx):
    """meshlets=0 writes the shared mesh only (no meshlet buffers); reuse mode
    rewrites the same meshlets into the same buffers."""
    dims = (70, 61, 53)
    data = oracle.gen_blobs_u8(dims, 92, 10)
    mm, pc = run_minmax(ctx, data, dims, 8, iso=128.0)
    f = filterActiveBlocks(ctx, mm, pc)
    ex0 = extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(meshlets=0))
    ex0.read_counts()
    assert not ex0.c.meshlets and not ex0.c.meshlet_vertices and ex0.meshletCount == 0
    ex = extractMeshletDescriptors(ctx, mm, f, pc)
    ex.read_counts()
    first = (ex.read_meshlets().copy(), ex.read_meshlet_indices().copy())
    np.testing.assert_array_equal(ex0.read_indices(), ex.read_indices())
    extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(reuse=1), out=ex)
    ex.read_counts()
    np.testing.assert_array_equal(ex.read_meshlets(), first[0])
    np.testing.assert_array_equal(ex.read_meshlet_indices(), first[1])
    check_meshlets(ex, data, dims, 8, 128.0)


@pytest.mark.parametrize("B,rule", [(8, 0), (8, 1), (4, 0), (5, 1)])
def test_coarse_reject_gives_the_same_list(ctx, B, rule):
    # sparse vessels: most level-2 ancestors inactive
    dims = (150, 131, 97)
    data = oracle.gen_vessels_u8(dims, oracle.SEED_BASE + 1, 4, 8)
    iso = 80.0
    pc = makePushConstants(dims, B, iso)
    vol = vol_of(data, dims)
    # fused mask made for another isovalue: the filter reads the pairs
    mm = computeMinMaxMip(ctx, vol, makePushConstants(dims, B, 7.0), options=makeOptions(rule=rule))
    assert mm.c.levels_built > 2
    from meshtrex_amd.managers import read_device
    out = []
    for coarse in (0, 1):
        f = filterActiveBlocks(ctx, mm, pc, options=makeOptions(rule=rule, coarse_reject=coarse))
        readActiveBlockCount(ctx, f)
        nb = int(np.prod(mm.grid))
        out.append((f.read_ids(), read_device(f.c.block_to_active, np.uint32, nb)))
        f.cleanup()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    ref = oracle.filter_ids(oracle.minmax(data, dims, B, 1), oracle.U8, iso, rule)
    np.testing.assert_array_equal(out[1][0], ref)
    assert 0 < ref.size < nb // 2
