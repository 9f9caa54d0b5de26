"""CPU oracle (TEST INFRASTRUCTURE ONLY).

ctypes wrapper of oracle/mtx_oracle.c, the scalar C restatement of the
reference's semantics.  Only tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py use it, as the checker; the product library never
calls it.  Build it with `make -C oracle` (done by __graft_entry__.build()).

Each function names the reference code it restates in mtx_oracle.c's header;
tests/test_oracle.py lists which check pins which function.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmtx_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
SEED_BASE = 0x6D747278
U8, F32 = 0, 1
RULE_CPU, RULE_GPU = 0, 1

_lib = None


def build(ref=False):
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.orc_key_f32.restype = ctypes.c_uint32
        _lib.orc_key_f32.argtypes = [ctypes.c_float]
        _lib.orc_unkey_f32.restype = ctypes.c_float
        _lib.orc_unkey_f32.argtypes = [ctypes.c_uint32]
        _lib.orc_active.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_int]
        _lib.orc_filter.restype = ctypes.c_uint64
        _lib.orc_filter.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int, ctypes.c_void_p]
        _lib.orc_pyramid_layout.restype = ctypes.c_uint32
        for n in ("orc_gen_blobs_u8", "orc_gen_blobs_f32"):
            getattr(_lib, n).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
        _lib.orc_gen_vessels_u8.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                            ctypes.c_int, ctypes.c_int]
        _lib.orc_gen_sdf_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double]
        _lib.orc_extract_count.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                           ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
        _lib.orc_extract_count_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                                ctypes.c_void_p]
        _lib.orc_extract.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                     ctypes.c_float] + [ctypes.c_void_p] * 5
        _lib.orc_case_indices.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                          ctypes.c_float, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        _lib.orc_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                     ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.orc_positions.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                       ctypes.c_uint64, ctypes.c_void_p]
        _lib.orc_minmax.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                    ctypes.c_void_p]
        _lib.orc_minmax_legacy.argtypes = _lib.orc_minmax.argtypes
        _lib.orc_pyramid.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.orc_meshlets.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_void_p] * 3
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _dims(d):
    return (ctypes.c_uint32 * 3)(*[int(x) for x in d])


def _dt(vol):
    return U8 if vol.dtype == np.uint8 else F32


def grid(dims, B):
    return tuple((int(d) + B - 1) // B for d in dims)


# ---- generators ---------------------------------------------------------------

def gen_blobs_u8(dims, seed=SEED_BASE, nblobs=16):
    out = np.zeros(int(np.prod(dims)), np.uint8)
    assert lib().orc_gen_blobs_u8(_p(out), _dims(dims), seed, nblobs) == 0
    return out


def gen_blobs_f32(dims, seed=SEED_BASE, nblobs=16):
    out = np.zeros(int(np.prod(dims)), np.float32)
    assert lib().orc_gen_blobs_f32(_p(out), _dims(dims), seed, nblobs) == 0
    return out


def gen_vessels_u8(dims, seed=SEED_BASE + 1, npoly=40, nseg=8):
    out = np.zeros(int(np.prod(dims)), np.uint8)
    assert lib().orc_gen_vessels_u8(_p(out), _dims(dims), seed, npoly, nseg) == 0
    return out


def gen_sdf_f32(dims, seed=SEED_BASE + 2, nsph=64, ntor=32, scale=1.0, sigma=0.05):
    out = np.zeros(int(np.prod(dims)), np.float32)
    assert lib().orc_gen_sdf_f32(_p(out), _dims(dims), seed, nsph, ntor, scale, sigma) == 0
    return out


# ---- min/max / filter -----------------------------------------------------------

def minmax(vol, dims, B, overlap=1):
    vol = np.ascontiguousarray(vol)
    G = grid(dims, B)
    pairs = np.zeros(2 * G[0] * G[1] * G[2], np.uint32)
    fn = lib().orc_minmax if overlap else lib().orc_minmax_legacy
    fn(_p(vol), _dt(vol), _dims(dims), B, _p(pairs))
    return pairs.reshape(-1, 2)


def pyramid(leaf_pairs, G):
    """All levels from level 0; returns (pairs (n,2), offsets, extents)."""
    ext = ((ctypes.c_uint32 * 3) * 32)()
    off = (ctypes.c_uint64 * 33)()
    L = lib().orc_pyramid_layout(_dims(G), ext, off)
    allp = np.zeros(2 * off[L], np.uint32)
    allp[: leaf_pairs.size] = leaf_pairs.reshape(-1)
    lib().orc_pyramid(_p(allp), _dims(G))
    return allp.reshape(-1, 2), [off[i] for i in range(L + 1)], [tuple(ext[i]) for i in range(L)]


def key_f32(x):
    return lib().orc_key_f32(float(x))


def unkey_f32(k):
    return lib().orc_unkey_f32(int(k))


def active(lo, hi, dtype, iso, rule=RULE_CPU):
    return bool(lib().orc_active(int(lo), int(hi), dtype, float(iso), rule))


def filter_ids(pairs, dtype, iso, rule=RULE_CPU):
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
    nb = pairs.shape[0]
    ids = np.zeros(max(nb, 1), np.uint32)
    n = lib().orc_filter(_p(pairs), nb, dtype, iso, rule, _p(ids))
    return ids[:n].copy()


# ---- extraction ----------------------------------------------------------------

def extract_count(vol, dims, B, iso):
    vol = np.ascontiguousarray(vol)
    nv = ctypes.c_uint64()
    nt = ctypes.c_uint64()
    lib().orc_extract_count(_p(vol), _dt(vol), _dims(dims), B, iso, ctypes.byref(nv), ctypes.byref(nt))
    return nv.value, nt.value


def extract_count_threaded(vol, dims, B, iso, threads=16):
    """extract_count split over block rows (orc_extract_count_rows) on a
    thread pool; ctypes releases the GIL, so the rows run in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    vol = np.ascontiguousarray(vol)
    rows = -(-dims[2] // B)
    cuts = np.linspace(0, rows, min(threads, rows) + 1).astype(np.int64)
    d = _dims(dims)

    def part(k):
        nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
        lib().orc_extract_count_rows(_p(vol), _dt(vol), d, B, iso, int(cuts[k]), int(cuts[k + 1]),
                                     ctypes.byref(nv), ctypes.byref(nt))
        return nv.value, nt.value

    with ThreadPoolExecutor(len(cuts) - 1) as ex:
        parts = list(ex.map(part, range(len(cuts) - 1)))
    return sum(p[0] for p in parts), sum(p[1] for p in parts)


def extract(vol, dims, B, iso, normals=True):
    """dict: edge_ids (uint64), t (f64), ndc (n,3 f64), normal (n,3 f64), tris (m,3 uint32)."""
    vol = np.ascontiguousarray(vol)
    nv, nt = extract_count(vol, dims, B, iso)
    out = {
        "edge_ids": np.zeros(max(nv, 1), np.uint64),
        "t": np.zeros(max(nv, 1), np.float64),
        "ndc": np.zeros((max(nv, 1), 3), np.float64),
        "normal": np.zeros((max(nv, 1), 3), np.float64) if normals else None,
        "tris": np.zeros((max(nt, 1), 3), np.uint32),
    }
    r = lib().orc_extract(_p(vol), _dt(vol), _dims(dims), B, iso, _p(out["edge_ids"]), _p(out["t"]),
                          _p(out["ndc"]), _p(out["normal"]), _p(out["tris"]))
    if r != 0:
        raise RuntimeError("oracle extract failed: %d" % r)
    for k in ("edge_ids", "t", "ndc", "normal"):
        if out[k] is not None:
            out[k] = out[k][:nv]
    out["tris"] = out["tris"][:nt]
    return out


def meshlets(vol, dims, B, iso):
    """Renderer-layout meshlets (orc_meshlets): dict desc (m, 4) uint32
    {vertexOffset, indexOffset, vertexCount, primitiveCount}, map (s,) uint32
    meshlet vertex -> shared-mesh vertex, mind (t, 3) uint32 meshlet vertex of
    each triangle vertex."""
    vol = np.ascontiguousarray(vol)
    nm, ns = ctypes.c_uint64(), ctypes.c_uint64()
    r = lib().orc_meshlets(_p(vol), _dt(vol), _dims(dims), B, iso, ctypes.byref(nm), ctypes.byref(ns), None, None,
                           None)
    if r != 0:
        raise RuntimeError("oracle meshlets failed: %d" % r)
    _, nt = extract_count(vol, dims, B, iso)
    desc = np.zeros((max(nm.value, 1), 4), np.uint32)
    mp = np.zeros(max(ns.value, 1), np.uint32)
    mind = np.zeros((max(nt, 1), 3), np.uint32)
    r = lib().orc_meshlets(_p(vol), _dt(vol), _dims(dims), B, iso, ctypes.byref(nm), ctypes.byref(ns), _p(desc),
                           _p(mp), _p(mind))
    if r != 0:
        raise RuntimeError("oracle meshlets failed: %d" % r)
    return {"desc": desc[:nm.value], "map": mp[:ns.value], "mind": mind[:nt]}


def case_indices(vol, dims, B, iso, ids):
    vol = np.ascontiguousarray(vol)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    out = np.zeros(max(ids.size * B ** 3, 1), np.uint8)
    lib().orc_case_indices(_p(vol), _dt(vol), _dims(dims), B, iso, _p(ids), ids.size, _p(out))
    return out[: ids.size * B ** 3]


def adjoint(vol, dims, iso, edge_ids, dLdp):
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    e = np.ascontiguousarray(edge_ids, dtype=np.uint64)
    g = np.ascontiguousarray(dLdp, dtype=np.float64).reshape(-1)
    dv = np.zeros(vol.size, np.float64)
    di = ctypes.c_double()
    lib().orc_adjoint(_p(vol), _dims(dims), iso, _p(e), e.size, _p(g), _p(dv), ctypes.byref(di))
    return dv, di.value


def positions(vol, dims, iso, edge_ids):
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    e = np.ascontiguousarray(edge_ids, dtype=np.uint64)
    p = np.zeros((e.size, 3), np.float64)
    lib().orc_positions(_p(vol), _dims(dims), iso, _p(e), e.size, _p(p))
    return p


def tri_table():
    t = np.zeros(4096, np.int32)
    lib().orc_tri_table(_p(t))
    return t.reshape(256, 16)


def edge_table():
    t = np.zeros(256, np.int32)
    lib().orc_edge_table(_p(t))
    return t


def edge_owner_table():
    t = np.zeros(48, np.uint32)
    lib().orc_edge_owner_table(_p(t))
    return t.reshape(12, 4)


def canonical_triangles(tris, edge_ids):
    """Triangles as edge-ID triples, rotated so the smallest ID is first
    (winding kept), rows sorted: comparable across vertex orderings."""
    e = np.asarray(edge_ids, dtype=np.uint64)[np.asarray(tris, dtype=np.int64)]
    if e.size == 0:
        return e.reshape(0, 3)
    r = np.argmin(e, axis=1)
    idx = (r[:, None] + np.arange(3)[None, :]) % 3
    e = np.take_along_axis(e, idx, axis=1)
    order = np.lexsort((e[:, 2], e[:, 1], e[:, 0]))
    return e[order]
