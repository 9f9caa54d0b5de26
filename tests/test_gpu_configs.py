"""The benchmark workloads at their stated sizes: the HIP path against the oracle.

blobs 256^3 u8, iso 128, 8^3 blocks (the reference CPU path's volume: its
    SHA-256 is the one the reference's computeMinMaxFromFile golden was made
    from): every output bit-exact, t within 1e-6, NDC within 2e-6, normals
    within 1e-4 rad, meshlets equal to the oracle's cut.
vessels 512^3 u8 (40 polylines of 8 segments), iso 80, 8^3 blocks: the
    GPU-generated volume read back; every output against the oracle as above.
SDF 1024^3 f32 (64 spheres, 32 tori, noise 0.05), iso 0, 8^3 blocks: on the
    whole volume, min/max pairs and the active list bit-exact, vertex and
    triangle counts equal to the oracle's, and the edge IDs distinct, active
    and in output order (with equal counts this makes them the oracle's list),
    positions within 1e-6 of the FP64 ones, the meshlet mesh equal to the
    shared mesh with no triangle dropped by the renderer's rule; every output
    bit-exact on the 1024 x 1024 x 65 slab z < 65.
batch of 48 volumes (225 units): every unit equals the single-volume path; the
    smallest volume of each dtype equals the oracle at each of its isovalues.
adjoint on 16 float32 blob volumes of 128^3-256^3: dL/dvolume and dL/diso of
    the target-sphere loss against the FP64 oracle adjoint, and against FP64
    central differences (h = 1e-3, Richardson-extrapolated) on 256 voxels per
    volume adjacent to the surface, 1e-4 relative.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from meshtrex_amd import (computeMinMaxMip, extractMeshletDescriptors, filterActiveBlocks, makeOptions,
                          makePushConstants, read_device, readActiveBlockCount, workloads)
from parity_checks import check_extraction, check_units_against_plain, renderer_accepts

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SEED = oracle.SEED_BASE


def generated(dims, kind, seed):
    """A volume from the GPU generator, read back to the host."""
    buf, _ = workloads.generate(dims, kind, seed)
    host = read_device(buf.ptr, np.uint8 if kind.startswith("u8") else np.float32, int(np.prod(dims)))
    buf.free()
    return host


def test_blobs_256_u8_iso128(ctx):
    dims = (256, 256, 256)
    data = oracle.gen_blobs_u8(dims, SEED, 16)
    g = json.load(open(os.path.join(HERE, "golden", "reference_cpu_golden.json")))
    assert hashlib.sha256(data.tobytes()).hexdigest() == g["volume"]["sha256"]
    _, f, ex = check_extraction(ctx, data, dims, 8, 128.0)
    assert ex.vertexCount > 50000 and ex.meshletCount > 0


def test_vessels_512_u8_iso80(ctx):
    dims = (512, 512, 512)
    data = generated(dims, "u8vessels", SEED + 1)
    _, f, ex = check_extraction(ctx, data, dims, 8, 80.0)
    # the sparse workload it stands for: a few percent of the blocks active
    assert 0.005 < f.activeBlockCount / 64 ** 3 < 0.15
    assert ex.vertexCount > 0


def owner_key(eid, dims, B):
    """Output-order key of each vertex: block ID, owner corner in block-local
    x-fastest order, axis."""
    e = eid.astype(np.int64)
    vox, ax = e // 3, e % 3
    X, Y = dims[0], dims[1]
    x, y, z = vox % X, (vox // X) % Y, vox // (X * Y)
    G = [-(-d // B) for d in dims]
    bid = x // B + G[0] * (y // B + G[1] * (z // B))
    local = x % B + B * (y % B + B * (z % B))
    return (bid * B ** 3 + local) * 3 + ax, x, y, z, ax


def test_sdf_1024_f32_iso0(ctx):
    dims = (1024, 1024, 1024)
    iso = 0.0
    buf, vol = workloads.generate(dims, "f32", SEED + 2)
    data = read_device(buf.ptr, np.float32, 1024 ** 3)
    pc = makePushConstants(dims, 8, iso)
    mm = computeMinMaxMip(ctx, vol, pc)
    f = filterActiveBlocks(ctx, mm, pc)
    readActiveBlockCount(ctx, f)
    ex = extractMeshletDescriptors(ctx, mm, f, pc, options=makeOptions(edge_ids=1, voxel_positions=1))
    ex.read_counts()
    pairs = oracle.minmax(data, dims, 8)
    np.testing.assert_array_equal(mm.read_pairs(0), pairs)
    ids = oracle.filter_ids(pairs, oracle.F32, iso)
    np.testing.assert_array_equal(f.read_ids(), ids)
    nv, nt = oracle.extract_count_threaded(data, dims, 8, iso)
    assert (ex.vertexCount, ex.indexCount) == (nv, 3 * nt)

    # the edge list: in output order (strictly increasing key, hence distinct),
    # every edge inside the volume and crossing the isovalue; nv of them, so
    # exactly the oracle's list
    eid = ex.read_edge_ids()
    key, x, y, z, ax = owner_key(eid, dims, 8)
    assert np.all(np.diff(key) > 0)
    end = np.choose(ax, [x, y, z])
    assert np.all(end + 1 < 1024)
    vox = (eid // 3).astype(np.int64)
    st = np.array([1, 1024, 1024 * 1024], np.int64)[ax]
    assert np.all((data[vox] <= iso) != (data[vox + st] <= iso))
    p64 = oracle.positions(data, dims, iso, eid)
    vp = ex.read_voxel_positions().astype(np.float64)
    assert np.all(np.abs(vp - p64) <= 2.0 ** -23 * np.maximum(np.abs(p64), 1.0) + 1e-6)
    del vox, st, key, x, y, z, ax, end, p64, vp, eid

    # the meshlet mesh is the shared mesh, whole under the renderer's rule
    idx = ex.read_indices().reshape(-1, 3)
    assert idx.size == 0 or int(idx.max()) < nv
    desc, mp, mind = ex.read_meshlets(), ex.read_meshlet_vertex_map(), ex.read_meshlet_indices().reshape(-1, 3)
    np.testing.assert_array_equal(mp[mind], idx)
    assert renderer_accepts(desc, mind, mp.size) == 0
    assert int(desc[-1, 1]) + 3 * int(desc[-1, 3]) == ex.indexCount
    del idx, desc, mp, mind
    ex.cleanup()
    f.cleanup()
    mm.cleanup()
    buf.free()

    # bit-exact on the slab of the first 65 z-planes
    slab = np.ascontiguousarray(data[: 1024 * 1024 * 65])
    del data
    _, f, ex = check_extraction(ctx, slab, (1024, 1024, 65), 8, iso)
    assert ex.vertexCount > 0 and ex.meshletCount > 0


def test_batch_48_volumes(ctx):
    from meshtrex_amd.batch import Batch
    vols = []
    for dims, kind, seed, fr in workloads.batched_specs(48):
        buf, vol = workloads.generate(dims, kind, seed)
        vols.append((None, dims, vol, workloads.isovalues(buf, dims, kind, fr), buf, kind))
    units = sum(len(v[3]) for v in vols)
    assert units == 225
    with Batch(ctx, [(v[2], v[3]) for v in vols], edge_ids=True, tune_iters=1) as b:
        b.run()
        ranges = check_units_against_plain(ctx, b, [v[:5] for v in vols])
        assert len(ranges) == units
        # the smallest volume of each dtype against the oracle at every isovalue
        first = np.cumsum([0] + [len(v[3]) for v in vols])
        for kind in ("u8", "f32"):
            k = min((j for j, v in enumerate(vols) if v[5] == kind), key=lambda j: int(np.prod(vols[j][1])))
            _, dims, _, isos, buf, _ = vols[k]
            data = read_device(buf.ptr, np.uint8 if kind == "u8" else np.float32, int(np.prod(dims)))
            for j, iso in enumerate(isos):
                u = int(first[k]) + j
                ref = oracle.extract(data, dims, ranges[u]["block_dim"], iso, normals=True)
                v, i = b.unit_mesh(u, ranges)
                np.testing.assert_array_equal(b.unit_edge_ids(u, ranges), ref["edge_ids"])
                np.testing.assert_array_equal(i.reshape(-1, 3), ref["tris"])
                if len(ref["edge_ids"]):
                    assert np.max(np.abs(v[:, :3] - ref["ndc"])) <= 2e-6


def local_fd(base, iso, vox, st, owner, ax, cc, R, i, sel, h):
    """Central difference of the target-sphere loss in voxel i, over the
    vertices `sel` whose edges have i as an endpoint (no other term of the
    loss depends on voxel i), Richardson-extrapolated from steps h and h/2."""
    a, b = vox[sel], vox[sel] + st[sel]

    def loss(dh):
        v0 = base[a] + np.where(a == i, dh, 0.0)
        v1 = base[b] + np.where(b == i, dh, 0.0)
        p = owner[sel].copy()
        p[np.arange(sel.size), ax[sel]] += (iso - v0) / (v1 - v0)
        return float(np.sum((np.linalg.norm(p - cc, axis=1) - R) ** 2))

    def fd(hh):
        return (loss(hh) - loss(-hh)) / (2 * hh)

    return (4 * fd(h / 2) - fd(h)) / 3


def test_adjoint_16_volumes():
    import torch
    from meshtrex_amd.autograd import Extraction, marching_cubes
    iso, h = 0.0, 1e-3
    specs = workloads.differentiable_specs(16)
    assert len(specs) == 16
    for n, (dims, seed) in enumerate(specs):
        v = generated(dims, "f32blobs", seed)
        vol = torch.from_numpy(v.reshape(dims[::-1])).cuda().requires_grad_(True)
        iso_t = torch.tensor(iso, device="cuda", requires_grad=True)
        c = torch.tensor([(d - 1) / 2.0 for d in dims], device="cuda", dtype=torch.float64)
        R = 0.3 * min(dims)
        pos, faces = marching_cubes(vol, iso_t)
        ((pos.double() - c).norm(dim=1) - R).pow(2).sum().backward()
        g_vol = vol.grad.cpu().numpy().reshape(-1).astype(np.float64)
        g_iso = float(iso_t.grad)
        e = Extraction(vol.detach(), iso, 8, edge_ids=True).edge_ids().cpu().numpy().astype(np.uint64)
        assert e.size == pos.shape[0] > 0
        del vol, pos, faces

        # the FP64 gradient of the same loss through the oracle's adjoint
        cc = np.array([(d - 1) / 2.0 for d in dims])
        p64 = oracle.positions(v, dims, iso, e)
        r = np.linalg.norm(p64 - cc, axis=1)
        ref_dv, ref_di = oracle.adjoint(v, dims, iso, e, (2 * (r - R) / r)[:, None] * (p64 - cc))
        scale = np.abs(ref_dv).max()
        assert np.all(np.abs(g_vol - ref_dv) <= 1e-4 * np.abs(ref_dv) + 1e-5 * scale), n
        assert abs(g_iso - ref_di) <= 1e-4 * abs(ref_di), n

        # central differences on 256 voxels that are endpoints of active edges
        ax = (e % 3).astype(np.int64)
        vox = (e // 3).astype(np.int64)
        st = np.array([1, dims[0], dims[0] * dims[1]])[ax]
        owner = np.stack([vox % dims[0], (vox // dims[0]) % dims[1], vox // (dims[0] * dims[1])],
                         1).astype(np.float64)
        base = v.astype(np.float64)
        ends = np.concatenate([vox, vox + st])
        order = np.argsort(ends, kind="stable")
        ends_sorted = ends[order]
        cand = np.unique(ends)
        cand = cand[np.abs(base[cand] - iso) > 10 * h]
        pick = np.random.default_rng(100 + n).choice(cand, min(256, cand.size), replace=False)
        floor = 1e-6 * np.abs(g_vol).max()
        worst = 0.0
        for i in pick.tolist():
            lo, hi = np.searchsorted(ends_sorted, [i, i + 1])
            sel = np.unique(order[lo:hi] % e.size)
            d = local_fd(base, iso, vox, st, owner, ax, cc, R, i, sel, h)
            worst = max(worst, (abs(d - g_vol[i]) - floor) / max(abs(d), 1e-300))
        assert worst <= 1e-4, (n, worst)
