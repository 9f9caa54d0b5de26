"""Batched form (mtx_batch_*) vs the single-volume path and the oracle.

Every unit of a batch -- one (volume, isovalue) pair -- must give exactly what
the single-volume path gives for that volume and isovalue at the same block
edge: the same active block IDs, vertex records (bit-identical floats), index
buffer and owning edge IDs, with indices relative to the unit's first
vertex.  The batched form writes the shared mesh only: no meshlets.  One unit per dtype is
also compared with the oracle directly.  Replays of the captured graph must
repeat, follow new volume contents, and flag outputs past the planned
capacity.
"""
import numpy as np
import pytest

import oracle
from meshtrex_amd import DeviceBuffer, MtxError, Volume
from meshtrex_amd.batch import Batch
from parity_checks import check_units_against_plain

pytestmark = pytest.mark.gpu

# u8 blobs and f32 SDFs; x extents of every residue mod 4 (rows staged through
# LDS), one multiple of 4 (aligned rows) and one narrower than 4 voxels
SPECS = [
    ((37, 45, 29), "u8", (90.5, 128.0, 200.0)),
    ((50, 33, 41), "f32", (0.0, 3.5)),
    ((66, 19, 23), "u8", (128.0,)),
    ((23, 61, 35), "f32", (-1.0, 0.0, 2.0)),
    ((5, 7, 3), "u8", (100.0,)),
    ((71, 39, 17), "f32", (1.5,)),
    ((64, 21, 18), "f32", (0.0,)),
    ((3, 11, 9), "u8", (128.0,)),
]


def make_data(dims, kind, seed):
    if kind == "u8":
        return oracle.gen_blobs_u8(dims, seed, 6)
    return oracle.gen_sdf_f32(dims, seed, 6, 3, min(dims) / 1024.0 * 4.0, 0.05)


@pytest.fixture(scope="module")
def volumes():
    out = []
    for k, (dims, kind, isos) in enumerate(SPECS):
        data = make_data(dims, kind, 0x6D747278 + 100 + k)
        buf = DeviceBuffer.from_numpy(data)
        out.append((data, dims, Volume(dims, "uint8" if kind == "u8" else "float32", buf), isos, buf))
    return out


@pytest.mark.parametrize("B,schedule,wg", [(4, 1, 64), (8, 1, 32), (16, 1, 64), (4, 2, 8), (8, 2, 16),
                                            (16, 2, 8)])
def test_batch_units_equal_single_volume_path(ctx, volumes, B, schedule, wg):
    # strided and cost-balanced work assignment give the same outputs
    with Batch(ctx, [(v[2], v[3]) for v in volumes], block_dim=B, minmax_rows=2, edge_ids=True,
               schedule=schedule, wg_per_cu=wg) as b:
        b.run()
        ranges = check_units_against_plain(ctx, b, volumes)
        assert sum(r["vertex_count"] for r in ranges) > 0
        # slices are disjoint and ordered within each dtype class
        for info in b.class_info():
            assert info["block_dim"] == B


def test_batch_unit_matches_oracle(ctx, volumes):
    with Batch(ctx, [(v[2], v[3]) for v in volumes], block_dim=8, edge_ids=True) as b:
        b.run()
        ranges = b.read_ranges()
        # unit 1: first u8 volume at its second isovalue; unit 3: first f32 volume, first isovalue
        for u, (k, j) in ((1, (0, 1)), (3, (1, 0))):
            data, dims, _, isos, _ = volumes[k]
            ref = oracle.extract(data, dims, 8, isos[j], normals=True)
            eid = b.unit_edge_ids(u, ranges)
            v, i = b.unit_mesh(u, ranges)
            np.testing.assert_array_equal(eid, ref["edge_ids"])
            np.testing.assert_array_equal(i.reshape(-1, 3), ref["tris"])
            assert np.max(np.abs(v[:, :3] - ref["ndc"])) <= 2e-6


def test_batch_autotuner_and_graph_replays(ctx, volumes):
    with Batch(ctx, [(v[2], v[3]) for v in volumes], edge_ids=True, tune_iters=1) as b:
        infos = b.class_info()
        assert {i["dtype"] for i in infos} == {"uint8", "float32"}
        for i in infos:
            assert i["block_dim"] in (4, 8, 16) and i["minmax_rows"] in (1, 2, 4)
            assert all(t > 0 for t in i["tuned_ms"]) and all(t > 0 for t in i["tuned_rows_ms"])
            assert i["tuned_ms"][(4, 8, 16).index(i["block_dim"])] == min(i["tuned_ms"])
            # the work assignment and grid: the fastest of the six timed candidates
            cands = [(1, 32), (1, 64), (1, 128), (2, 8), (2, 16), (2, 32)]
            assert all(t > 0 for t in i["tuned_sched_ms"])
            assert i["tuned_sched_ms"][cands.index((i["schedule"], i["wg_per_cu"]))] == min(i["tuned_sched_ms"])
        b.run()
        first = [b.unit_mesh(u) for u in range(b.num_units)]
        for _ in range(3):
            b.run()
        ranges = check_units_against_plain(ctx, b, volumes)
        for u in range(b.num_units):
            v, i = b.unit_mesh(u, ranges)
            np.testing.assert_array_equal(v, first[u][0])
            np.testing.assert_array_equal(i, first[u][1])


def test_batch_without_graph_matches_graph(ctx, volumes):
    vols = [(v[2], v[3]) for v in volumes]
    with Batch(ctx, vols, block_dim=8, graph=False) as a, Batch(ctx, vols, block_dim=8, graph=True) as g:
        a.run()
        g.run()
        ra, rg = a.read_ranges(), g.read_ranges()
        for u in range(a.num_units):
            va, ia = a.unit_mesh(u, ra)
            vg, ig = g.unit_mesh(u, rg)
            np.testing.assert_array_equal(va, vg)
            np.testing.assert_array_equal(ia, ig)


def test_batch_follows_new_contents_and_flags_capacity(ctx):
    dims = (41, 38, 27)
    a = oracle.gen_blobs_u8(dims, 11, 6)
    buf = DeviceBuffer.from_numpy(a)
    vol = Volume(dims, "uint8", buf)
    f32 = oracle.gen_sdf_f32((30, 29, 31), 12, 6, 3, 0.1, 0.05)
    fbuf = DeviceBuffer.from_numpy(f32)
    fvol = Volume((30, 29, 31), "float32", fbuf)
    with Batch(ctx, [(vol, (128.0,)), (fvol, (0.0,))], block_dim=8) as b:
        b.run()
        r0 = b.read_ranges()
        # a constant volume: no active block, the other unit unchanged
        from meshtrex_amd import _lib as L
        const = np.full(a.size, 7, np.uint8)
        assert L.lib().mtx_memcpy_h2d(buf.ptr, const.ctypes.data, const.nbytes) == 0
        b.run()
        r1 = b.read_ranges()
        assert r1[0]["vertex_count"] == 0 and r1[0]["index_count"] == 0 and r1[0]["active_count"] == 0
        assert r1[1]["vertex_count"] == r0[1]["vertex_count"]
        # noise has far more surface than the planned blobs: flagged, not overrun
        noise = np.random.default_rng(5).integers(0, 256, a.size, dtype=np.uint8)
        assert L.lib().mtx_memcpy_h2d(buf.ptr, noise.ctypes.data, noise.nbytes) == 0
        b.run()
        with pytest.raises(MtxError, match="MTX_ECAPACITY"):
            b.read_ranges()
        # back to the planned contents: complete again
        assert L.lib().mtx_memcpy_h2d(buf.ptr, a.ctypes.data, a.nbytes) == 0
        b.run()
        r2 = b.read_ranges()
        assert r2[0]["vertex_count"] == r0[0]["vertex_count"] and r2[0]["flags"] == 0


def test_batch_rejects_bad_input(ctx):
    dims = (9, 9, 9)
    data = oracle.gen_blobs_u8(dims, 3, 2)
    host = Volume(dims, "uint8", data)
    with pytest.raises(ValueError):
        Batch(ctx, [(host, (1.0,))])
    dev = Volume(dims, "uint8", DeviceBuffer.from_numpy(data))
    with pytest.raises(MtxError, match="MTX_EINVAL"):
        Batch(ctx, [(dev, (1.0,))], block_dim=5)
    with pytest.raises(MtxError, match="MTX_EINVAL"):
        Batch(ctx, [(dev, ())])
    with pytest.raises(MtxError, match="MTX_EINVAL"):
        Batch(ctx, [(dev, (1.0,))], schedule=3)
