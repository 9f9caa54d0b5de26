"""Pins the CPU oracle (oracle/mtx_oracle.c) before it is trusted as the checker.

Which check pins which oracle function:
  orc_tri_table / orc_edge_table   reference tables: FNV-1a64 fixtures from src/mc_tables.h,
                                   direct equality with the compiled reference header when
                                   oracle/_ref is built, and table invariants
  orc_edge_owner_table             the reference's edgeOwner constants
                                   (src/shaders/marching_cubes_pmb.mesh.glsl:39-42)
  orc_minmax_legacy                byte-identical to the reference's computeMinMaxFromFile
                                   (golden SHA-256 over 262,144 pairs + samples)
  orc_minmax (overlap 1)           numpy brute force over explicit windows; ramp closed form
  orc_pyramid                      numpy brute force
  orc_active / orc_filter          reference computeActiveBlockCountCPU golden counts;
                                   hand-built blocks for the rounding / min==max rules
  orc_extract counts               numpy count of sign-change edges (no tables); triangle
                                   count from the reference's numVerticesTable
  orc_extract topology             independent pure-Python MC keyed by endpoint voxel pairs;
                                   closed sphere: every edge shared by two triangles,
                                   Euler characteristic 2, consistent winding
  orc_extract geometry             interpolant at t equals the isovalue; hot-voxel known
                                   answer; linear-field normals
  orc_adjoint                      FP64 central differences of orc_positions
  generators                       golden SHA-256 of the 256^3 blob volume
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_cpu_golden.json")))


def fnv1a64(b):
    h = 0xCBF29CE484222325
    for x in np.frombuffer(b, dtype=np.uint8).tolist():
        h ^= x
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


# ---- tables ------------------------------------------------------------------

def test_tri_table_matches_reference_hash():
    assert fnv1a64(oracle.tri_table().astype("<i4").tobytes()) == GOLD["tables_fnv1a64_int32le"]["triTable"]


def test_edge_table_matches_reference_hash():
    assert fnv1a64(oracle.edge_table().astype("<i4").tobytes()) == GOLD["tables_fnv1a64_int32le"]["edgeTable"]


def test_num_vertices_table_is_three_per_triangle():
    tri = oracle.tri_table()
    nvt = np.array([3 * int(np.sum(tri[c] >= 0) // 3) for c in range(256)], "<i4")
    assert fnv1a64(nvt.tobytes()) == GOLD["tables_fnv1a64_int32le"]["numVerticesTable"]


@pytest.mark.skipif(not os.path.exists(oracle.REF_DRIVER), reason="reference tree not built here")
def test_tables_equal_compiled_reference_header(tmp_path):
    p = tmp_path / "t.bin"
    subprocess.check_call([oracle.REF_DRIVER, "tables", str(p)])
    raw = np.fromfile(p, dtype=np.int32)
    np.testing.assert_array_equal(oracle.tri_table().reshape(-1), raw[:4096])
    np.testing.assert_array_equal(oracle.edge_table(), raw[4096:4352])


def test_table_invariants():
    tri = oracle.tri_table()
    edge = oracle.edge_table()
    hist = np.zeros(6, int)
    for c in range(256):
        row = tri[c]
        n = int(np.sum(row >= 0))
        assert n % 3 == 0 and np.all(row[:n] >= 0) and np.all(row[n:] == -1)
        used = 0
        for e in row[:n]:
            used |= 1 << int(e)
        assert used == edge[c]          # edges used by the triangles == sign-change edges
        hist[n // 3] += 1
    assert hist.tolist() == [2, 16, 50, 80, 76, 32]


def test_edge_owner_matches_reference_constants():
    # src/shaders/marching_cubes_pmb.mesh.glsl:39-42
    ref = [(0, 0, 0, 0), (1, 0, 0, 1), (0, 1, 0, 0), (0, 0, 0, 1), (0, 0, 1, 0), (1, 0, 1, 1),
           (0, 1, 1, 0), (0, 0, 1, 1), (0, 0, 0, 2), (1, 0, 0, 2), (1, 1, 0, 2), (0, 1, 0, 2)]
    assert oracle.edge_owner_table().tolist() == [list(r) for r in ref]


# ---- generators / golden min-max ---------------------------------------------

@pytest.fixture(scope="module")
def blob256():
    return oracle.gen_blobs_u8((256, 256, 256), GOLD["volume"]["seed"], 16)


def test_generator_golden_hash(blob256):
    assert hashlib.sha256(blob256.tobytes()).hexdigest() == GOLD["volume"]["sha256"]


def test_minmax_legacy_matches_reference_cpu(blob256):
    pairs = oracle.minmax(blob256, (256, 256, 256), 4, overlap=0)
    g = GOLD["minmax_b4_overlap0"]
    assert hashlib.sha256(pairs.tobytes()).hexdigest() == g["sha256"]
    np.testing.assert_array_equal(pairs[np.array(g["sample_ids"])], np.array(g["sample_pairs"]))


def test_active_count_matches_reference_cpu(blob256):
    pairs = oracle.minmax(blob256, (256, 256, 256), 4, overlap=0)
    for iso, n in GOLD["active_count_b4"].items():
        assert oracle.filter_ids(pairs, oracle.U8, float(iso), oracle.RULE_CPU).size == n


# ---- min/max brute force ----------------------------------------------------------

def brute_minmax(vol3, B, overlap, keyfn):
    Z, Y, X = vol3.shape
    G = [(d + B - 1) // B for d in (X, Y, Z)]
    keys = keyfn(vol3)
    out = []
    span = B if overlap else B - 1
    for bz in range(G[2]):
        for by in range(G[1]):
            for bx in range(G[0]):
                w = keys[bz * B:min(bz * B + span, Z - 1) + 1, by * B:min(by * B + span, Y - 1) + 1,
                         bx * B:min(bx * B + span, X - 1) + 1]
                out.append((w.min(), w.max()))
    return np.array(out, np.uint32)


def f32_keys(a):
    b = a.view(np.uint32).astype(np.uint64)
    return np.where(b & 0x80000000, (~b) & 0xFFFFFFFF, b | 0x80000000).astype(np.uint32)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("overlap", [0, 1])
def test_minmax_brute_force(seed, overlap):
    rng = np.random.default_rng(seed)
    dims = tuple(int(x) for x in rng.integers(1, 21, 3))
    for B in (1, 2, 3, 4, 8, 16):
        u = rng.integers(0, 256, dims[::-1], dtype=np.uint8)
        got = oracle.minmax(u.reshape(-1), dims, B, overlap)
        np.testing.assert_array_equal(got, brute_minmax(u, B, overlap, lambda a: a.astype(np.uint32)))
        f = rng.standard_normal(dims[::-1]).astype(np.float32)
        got = oracle.minmax(f.reshape(-1), dims, B, overlap)
        np.testing.assert_array_equal(got, brute_minmax(f, B, overlap, f32_keys))


def test_minmax_ramp_closed_form():
    X, B = 29, 4
    v = np.tile(np.arange(X, dtype=np.uint8), 3 * 3)
    pairs = oracle.minmax(v, (X, 3, 3), B, 1)
    for bx in range((X + B - 1) // B):
        assert tuple(pairs[bx]) == (B * bx, min(B * bx + B, X - 1))


def test_f32_key_order():
    xs = np.array([-np.inf, -3e38, -1.0, -1e-40, -0.0, 0.0, 1e-40, 1.0, 3e38, np.inf], np.float32)
    keys = [oracle.key_f32(x) for x in xs]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    assert all(np.float32(oracle.unkey_f32(k)).tobytes() == x.tobytes() for k, x in zip(keys, xs))


def test_pyramid_brute_force():
    rng = np.random.default_rng(3)
    for G in [(5, 3, 7), (8, 8, 8), (13, 1, 2), (1, 1, 1)]:
        leaf = np.stack([rng.integers(0, 1000, np.prod(G)), rng.integers(1000, 2000, np.prod(G))], 1)
        allp, off, ext = oracle.pyramid(leaf.astype(np.uint32), G)
        assert len(ext) == 1 + int(np.floor(np.log2(max(G))))
        cur = leaf.reshape(G[2], G[1], G[0], 2)
        for lev in range(1, len(ext)):
            e = ext[lev]
            nxt = np.zeros((e[2], e[1], e[0], 2), np.uint32)
            for z in range(e[2]):
                for y in range(e[1]):
                    for x in range(e[0]):
                        c = cur[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2].reshape(-1, 2)
                        nxt[z, y, x] = (c[:, 0].min(), c[:, 1].max())
            np.testing.assert_array_equal(allp[off[lev]:off[lev + 1]], nxt.reshape(-1, 2))
            cur = nxt


# ---- activity rules -------------------------------------------------------------

def test_activity_rules_u8():
    A = oracle.active
    assert A(80, 81, oracle.U8, 80.6) and A(80, 81, oracle.U8, 80.4)
    assert A(79, 80, oracle.U8, 80.4)            # round(80.4)=80 in [79,80]
    assert not A(79, 80, oracle.U8, 80.6)        # round(80.6)=81
    assert not A(80, 80, oracle.U8, 80.0)        # min == max excluded (CPU rule)
    assert A(80, 80, oracle.U8, 80.0, oracle.RULE_GPU)
    assert not A(79, 80, oracle.U8, 80.4, oracle.RULE_GPU)
    assert not A(0, 10, oracle.U8, -3.0) and not A(0, 10, oracle.U8, float("nan"))


def test_activity_rules_f32():
    k = oracle.key_f32
    assert oracle.active(k(-1.0), k(2.0), oracle.F32, 0.0)
    assert not oracle.active(k(0.5), k(0.5), oracle.F32, 0.5)
    assert oracle.active(k(0.5), k(0.5), oracle.F32, 0.5, oracle.RULE_GPU)
    assert not oracle.active(k(0.1), k(2.0), oracle.F32, 0.0)


def test_filter_is_ascending_rule_subset():
    rng = np.random.default_rng(5)
    lo = rng.integers(0, 200, 5000)
    pairs = np.stack([lo, lo + rng.integers(0, 50, 5000)], 1).astype(np.uint32)
    ids = oracle.filter_ids(pairs, oracle.U8, 100.0)
    want = np.nonzero((pairs[:, 0] != pairs[:, 1]) & (pairs[:, 0] <= 100) & (pairs[:, 1] >= 100))[0]
    np.testing.assert_array_equal(ids, want)
    assert oracle.filter_ids(pairs, oracle.U8, -5.0).size == 0
    assert oracle.filter_ids(pairs, oracle.U8, 1000.0).size == 0


# ---- extraction ---------------------------------------------------------------

def sign_change_edges(vol3, iso):
    b = vol3.astype(np.float32) <= np.float32(iso)
    return (int(np.sum(b[:, :, 1:] != b[:, :, :-1])) + int(np.sum(b[:, 1:, :] != b[:, :-1, :]))
            + int(np.sum(b[1:, :, :] != b[:-1, :, :])))


def brute_ntri(vol3, iso):
    b = (vol3.astype(np.float32) <= np.float32(iso)).astype(np.int64)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    Z, Y, X = b.shape
    case = np.zeros((Z - 1, Y - 1, X - 1), np.int64)
    for i, (x, y, z) in enumerate(corners):
        case |= b[z:z + Z - 1, y:y + Y - 1, x:x + X - 1] << i
    ntri = np.array([int(np.sum(r >= 0)) // 3 for r in oracle.tri_table()])
    return int(ntri[case].sum())


@pytest.mark.parametrize("B", [8, 3, 5])
def test_extract_counts_brute_force(B):
    dims = (40, 31, 27)
    v = oracle.gen_blobs_u8(dims, 17, 6)
    v3 = v.reshape(dims[::-1])
    for iso in (128.0, 60.5, 200.0):
        nv, nt = oracle.extract_count(v, dims, B, iso)
        assert nv == sign_change_edges(v3, iso)
        assert nt == brute_ntri(v3, iso)
        # split over block rows, uneven row counts per thread: the same sums
        assert oracle.extract_count_threaded(v, dims, B, iso, threads=3) == (nv, nt)


def python_mc(vol3, iso):
    """Independent MC: vertices keyed by their (sorted) endpoint voxel pair."""
    tri = oracle.tri_table()
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    ends = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    Z, Y, X = vol3.shape
    out = []
    for z in range(Z - 1):
        for y in range(Y - 1):
            for x in range(X - 1):
                c = 0
                for i, (dx, dy, dz) in enumerate(corners):
                    if float(vol3[z + dz, y + dy, x + dx]) <= iso:
                        c |= 1 << i
                row = tri[c]
                for k in range(0, int(np.sum(row >= 0)), 3):
                    t = []
                    for e in row[k:k + 3]:
                        a, b = ends[e]
                        pa = (x + corners[a][0], y + corners[a][1], z + corners[a][2])
                        pb = (x + corners[b][0], y + corners[b][1], z + corners[b][2])
                        lo, hi = min(pa, pb, key=lambda p: (p[2], p[1], p[0])), max(pa, pb, key=lambda p: (p[2], p[1], p[0]))
                        axis = [i for i in range(3) if lo[i] != hi[i]][0]
                        t.append(3 * (lo[0] + X * (lo[1] + Y * lo[2])) + axis)
                    out.append(t)
    return np.array(out, np.uint64).reshape(-1, 3)


@pytest.mark.parametrize("seed", range(4))
def test_extract_topology_vs_independent_mc(seed):
    rng = np.random.default_rng(seed)
    dims = tuple(int(x) for x in rng.integers(2, 9, 3))
    v3 = rng.integers(0, 256, dims[::-1], dtype=np.uint8)
    for B in (2, 3, 8):
        r = oracle.extract(v3.reshape(-1), dims, B, 127.5)
        got = oracle.canonical_triangles(r["tris"], r["edge_ids"])
        ids = python_mc(v3, 127.5)
        want = oracle.canonical_triangles(np.arange(ids.size).reshape(-1, 3), ids.reshape(-1))
        np.testing.assert_array_equal(got, want)
        assert len(set(r["edge_ids"].tolist())) == len(r["edge_ids"])   # each edge once


def sphere_volume(n):
    # the reference's generateSphereVolume (src/meshtrex.cpp:23-52), iso 128
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = n * 0.4
    d = np.sqrt((x - n / 2.0) ** 2 + (y - n / 2.0) ** 2 + (z - n / 2.0) ** 2) - r
    v = np.clip((d / r) * 127.0 + 128.0 + 0.001, 0.0, 255.0)
    return v.astype(np.float32).astype(np.uint8).reshape(-1)


def test_sphere_is_closed_and_consistently_wound():
    n = 48
    v = sphere_volume(n)
    r = oracle.extract(v, (n, n, n), 8, 128.0)
    tris = r["tris"].astype(np.int64)
    V, F = len(r["edge_ids"]), len(tris)
    edges = np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert V - len(counts) + F == 2
    p = (r["ndc"] + 1.0) / 2.0 * n
    a, b, c = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    nrm = np.cross(b - a, c - a)
    radial = (a + b + c) / 3.0 - n / 2.0
    s = np.sign(np.sum(nrm * radial, axis=1))
    s = s[np.linalg.norm(nrm, axis=1) > 1e-9]     # voxels equal to the isovalue give t=0 slivers
    assert s.size > F // 2
    assert abs(s.sum()) == s.size                 # one orientation throughout


def test_vertices_interpolate_the_isovalue():
    dims = (33, 29, 31)
    v = oracle.gen_sdf_f32(dims, 5, 4, 2, 0.1)
    iso = 0.0
    r = oracle.extract(v, dims, 8, iso, normals=False)
    e = r["edge_ids"]
    vox = (e // 3).astype(np.int64)
    ax = (e % 3).astype(np.int64)
    stride = np.array([1, dims[0], dims[0] * dims[1]])[ax]
    v0 = v[vox].astype(np.float64)
    v1 = v[vox + stride].astype(np.float64)
    assert np.all((v0 <= iso) != (v1 <= iso))
    np.testing.assert_allclose(v0 + r["t"] * (v1 - v0), iso, atol=1e-12)


def test_hot_voxel_known_answer():
    dims = (17, 17, 17)
    d = np.zeros(dims[::-1], np.uint8)
    d[9, 8, 7] = 255
    r = oracle.extract(d.reshape(-1), dims, 8, 128.0)
    assert len(r["edge_ids"]) == 6 and len(r["tris"]) == 8
    p = (r["ndc"] + 1.0) / 2.0 * 17.0
    np.testing.assert_allclose(np.linalg.norm(p - [7, 8, 9], axis=1), 127 / 255, atol=1e-12)


def test_linear_field_normals():
    dims = (20, 20, 20)
    g = np.array([0.3, -0.5, 0.8])
    z, y, x = np.meshgrid(*[np.arange(20)] * 3, indexing="ij")
    v = (g[0] * x + g[1] * y + g[2] * z).astype(np.float32).reshape(-1)
    r = oracle.extract(v, dims, 8, 2.0)
    p = (r["ndc"] + 1.0) / 2.0 * 20.0
    inner = np.all((p > 1.5) & (p < 17.5), axis=1)
    assert inner.sum() > 50
    np.testing.assert_allclose(r["normal"][inner], np.tile(-g / np.linalg.norm(g), (inner.sum(), 1)),
                               atol=1e-6)


# ---- adjoint ---------------------------------------------------------------------

def test_adjoint_vs_central_differences():
    dims = (14, 12, 13)
    v = oracle.gen_blobs_f32(dims, 9, 4)
    iso = 0.0
    r = oracle.extract(v, dims, 4, iso, normals=False)
    e = r["edge_ids"]
    rng = np.random.default_rng(1)
    w = rng.standard_normal((e.size, 3))
    dv, di = oracle.adjoint(v, dims, iso, e, w)
    h = 1e-6   # FP64: truncation ~ (h / (v1-v0))^2, round-off ~ 1e-16 / h, both far below 1e-6
    e64 = e.astype(np.int64)
    vox = np.unique(np.concatenate([e64 // 3, e64 // 3 + np.array([1, dims[0], dims[0] * dims[1]])[e64 % 3]]))
    for i in rng.choice(vox, 40, replace=False).tolist():
        vp, vm = v.astype(np.float64).copy(), v.astype(np.float64).copy()
        vp[i] += h
        vm[i] -= h
        # positions in FP64 from perturbed values (topology frozen)
        def L(vals):
            vox_, ax = (e // 3).astype(np.int64), (e % 3).astype(np.int64)
            st = np.array([1, dims[0], dims[0] * dims[1]])[ax]
            t = (iso - vals[vox_]) / (vals[vox_ + st] - vals[vox_])
            return float(np.sum(w[np.arange(e.size), ax] * t))
        fd = (L(vp) - L(vm)) / (2 * h)
        assert abs(fd - dv[i]) <= 1e-6 * max(1.0, abs(fd))
    vals = v.astype(np.float64)
    vox_, ax = (e // 3).astype(np.int64), (e % 3).astype(np.int64)
    st = np.array([1, dims[0], dims[0] * dims[1]])[ax]
    def Li(iv):
        t = (iv - vals[vox_]) / (vals[vox_ + st] - vals[vox_])
        return float(np.sum(w[np.arange(e.size), ax] * t))
    fd = (Li(iso + h) - Li(iso - h)) / (2 * h)
    assert abs(fd - di) <= 1e-6 * max(1.0, abs(fd))
    # positions agree with the extraction's t along the owning edge
    p = oracle.positions(v, dims, iso, e)
    owner = np.stack([vox_ % dims[0], (vox_ // dims[0]) % dims[1], vox_ // (dims[0] * dims[1])], 1)
    np.testing.assert_allclose(p[np.arange(e.size), ax] - owner[np.arange(e.size), ax], r["t"], atol=1e-12)


# ---- meshlets -----------------------------------------------------------------
#
# The oracle's greedy cut (orc_meshlets) is pinned by properties that fix it
# uniquely, derived from the triangle list and the case indices rather than
# from a second greedy loop: meshlets are runs of consecutive cells of one
# block covering the triangles in order; each holds at most 64 vertices and
# 126 triangles; each but a block's last is maximal (its block's next cell
# would take it past a limit); its vertices are exactly those its triangles
# use, once each, in first-use order (cell order, then owner index u within
# the cell); map[meshlet index] is the shared mesh's index.

def uidx_of(dx, dy, dz, axis):
    """Owner index: 4 axis + d_p + 2 d_q, (p, q) the other axes ascending."""
    d = (dx, dy, dz)
    p, q = [a for a in range(3) if a != axis]
    return 4 * axis + d[p] + 2 * d[q]


def triangle_cells(data, dims, B, iso, ntris):
    """Global cell (x, y, z) and block ID of every triangle, from the case
    indices of the active blocks (triangles come block by block, cell by cell)."""
    G = oracle.grid(dims, B)
    ids = np.arange(G[0] * G[1] * G[2], dtype=np.uint32)
    cases = oracle.case_indices(data, dims, B, iso, ids).reshape(ids.size, -1)
    ntri = np.array([int(np.sum(r >= 0)) // 3 for r in oracle.tri_table()])
    cells, blocks = [], []
    for b in range(ids.size):
        bx, by, bz = b % G[0], (b // G[0]) % G[1], b // (G[0] * G[1])
        for c in np.nonzero(ntri[cases[b]])[0]:
            lx, ly, lz = c % B, (c // B) % B, c // (B * B)
            for _ in range(ntri[cases[b][c]]):
                cells.append((bx * B + lx, by * B + ly, bz * B + lz))
                blocks.append(b)
    assert len(cells) == ntris
    return np.array(cells, np.int64).reshape(-1, 3), np.array(blocks, np.int64)


def check_meshlet_contract(data, dims, B, iso):
    ref = oracle.extract(data, dims, B, iso, normals=False)
    m = oracle.meshlets(data, dims, B, iso)
    tris, eid = ref["tris"].astype(np.int64), ref["edge_ids"].astype(np.int64)
    desc, mp, mind = m["desc"].astype(np.int64), m["map"].astype(np.int64), m["mind"].astype(np.int64)
    np.testing.assert_array_equal(mp[mind], tris)
    if len(tris) == 0:
        assert len(desc) == 0 and mp.size == 0
        return
    cells, blocks = triangle_cells(data, dims, B, iso, len(tris))
    X, Y = dims[0], dims[1]
    cell_lin = cells[:, 0] + X * (cells[:, 1] + Y * cells[:, 2])
    # runs: contiguous triangles, contiguous meshlet vertices, whole cells, one block
    assert desc[0, 1] == 0 and desc[0, 0] == 0
    assert np.all(desc[1:, 1] == desc[:-1, 1] + 3 * desc[:-1, 3])
    assert np.all(desc[1:, 0] == desc[:-1, 0] + desc[:-1, 2])
    assert desc[-1, 1] + 3 * desc[-1, 3] == 3 * len(tris) and desc[-1, 0] + desc[-1, 2] == mp.size
    starts = desc[:, 1] // 3
    for k in range(len(desc)):
        t0, t1 = starts[k], starts[k] + desc[k, 3]
        vo, vc = desc[k, 0], desc[k, 2]
        assert 1 <= vc <= 64 and 1 <= desc[k, 3] <= 126
        assert len(set(blocks[t0:t1])) == 1
        if t0 > 0:
            assert cell_lin[t0] != cell_lin[t0 - 1]          # a cell is never split
        # the renderer keeps every triangle
        assert np.all((mind[t0:t1] >= vo) & (mind[t0:t1] < vo + vc))
        # exactly the vertices the triangles use, once each
        used = tris[t0:t1].reshape(-1)
        assert sorted(mp[vo:vo + vc].tolist()) == sorted(set(used.tolist()))
        # first-use order: first cell using the vertex, then its owner index in that cell
        key = {}
        for t in range(t0, t1):
            for v in tris[t]:
                if v in key:
                    continue
                e = eid[v]
                vox, ax = e // 3, e % 3
                o = np.array([vox % X, (vox // X) % Y, vox // (X * Y)]) - cells[t]
                assert np.all((o >= 0) & (o <= 1))
                key[v] = (cell_lin[t], uidx_of(o[0], o[1], o[2], ax))
        assert mp[vo:vo + vc].tolist() == sorted(key, key=lambda v: key[v])
        # maximal: the block's next cell would take the meshlet past a limit
        if k + 1 < len(desc) and blocks[starts[k + 1]] == blocks[t0]:
            n0 = starts[k + 1]
            n1 = n0
            while n1 < len(tris) and cell_lin[n1] == cell_lin[n0]:
                n1 += 1
            more = set(tris[n0:n1].reshape(-1).tolist()) - set(used.tolist())
            assert vc + len(more) > 64 or desc[k, 3] + (n1 - n0) > 126


@pytest.mark.parametrize("B", [8, 4, 5, 3])
def test_meshlet_contract_blobs(B):
    dims = (41, 37, 30)
    check_meshlet_contract(oracle.gen_blobs_u8(dims, oracle.SEED_BASE + 9, 10), dims, B, 128.0)


def test_meshlet_contract_noise_and_sdf():
    # noise: dense cells, many meshlets cut by the vertex limit
    rng = np.random.default_rng(3)
    dims = (21, 19, 17)
    check_meshlet_contract(rng.integers(0, 256, int(np.prod(dims)), dtype=np.uint8), dims, 8, 127.5)
    dims = (44, 40, 36)
    check_meshlet_contract(oracle.gen_sdf_f32(dims, 41, 6, 3, 0.15), dims, 8, 0.0)


def test_meshlet_known_answers():
    """One hot voxel in a field of 0 (iso 128): its 8 cells each give one
    triangle; with the voxel at a block's corner the cells fall in 8 blocks,
    so 8 meshlets of one triangle and 3 vertices each."""
    dims = (17, 17, 17)
    data = np.zeros(dims[::-1], np.uint8)
    data[8, 8, 8] = 255
    m = oracle.meshlets(data.reshape(-1), dims, 8, 128.0)
    assert len(m["desc"]) == 8 and np.all(m["desc"][:, 2:] == [3, 1])
    # an empty surface: no meshlets
    m = oracle.meshlets(np.zeros(8 * 8 * 8, np.uint8), (8, 8, 8), 8, 128.0)
    assert len(m["desc"]) == 0 and m["map"].size == 0
