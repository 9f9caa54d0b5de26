"""The C++ host API (include/meshtrex_amd/managers.hpp) end to end on the GPU:
loadVolume -> computeMinMaxMip -> filterActiveBlocks -> readActiveBlockCount ->
extractMeshletDescriptors -> writeGPUExtractionToOBJ, against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "cpp", "managers_demo")


def sphere_volume(w, h, d):
    # the reference's generateSphereVolume (src/meshtrex.cpp:23-52), in float32
    z, y, x = np.meshgrid(np.arange(d, dtype=np.float32), np.arange(h, dtype=np.float32),
                          np.arange(w, dtype=np.float32), indexing="ij")
    r = np.float32(w * 0.4)
    dist = np.sqrt((x - np.float32(w / 2)) ** 2 + (y - np.float32(h / 2)) ** 2 + (z - np.float32(d / 2)) ** 2) - r
    v = np.clip(dist / r * np.float32(127.0) + np.float32(128.0) + np.float32(0.001), 0, 255)
    return v.astype(np.uint8).reshape(-1)


def test_cpp_managers_demo(tmp_path):
    dims = (96, 80, 72)
    vol = sphere_volume(*dims)
    path = tmp_path / ("sphere_%dx%dx%d_uint8.raw" % dims)
    vol.tofile(path)
    obj = tmp_path / "sphere.obj"
    out = subprocess.run([DEMO, str(path), "128", "8", str(obj)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    ref = oracle.extract(vol, dims, 8, 128.0, normals=False)
    ids = oracle.filter_ids(oracle.minmax(vol, dims, 8), oracle.U8, 128.0)
    assert res["activeBlockCount"] == ids.size
    assert res["vertexCount"] == len(ref["edge_ids"])
    assert res["indexCount"] == 3 * len(ref["tris"])
    assert set(res["regions"]) == {"MinMax_Leaf_Pass", "MinMax_Octree_Reduction", "Active_Block_Filtering",
                                   "Mesh_Extraction"}
    # the reference-named members (src/filteringOutput.h:12-13, src/minMaxOutput.h:11-13,
    # src/extractionOutput.h:11-16) read back through their device pointers
    nb = int(np.prod([(d + 7) // 8 for d in dims]))
    assert res["activeBlockCountBuffer"] == ids.size
    assert res["firstActiveBlockId"] == int(ids[0])
    assert res["compactedBlockIdBuffer_size"] == 4 * nb
    assert res["volumeImage_size"] == vol.size
    assert res["levels"] == int(np.floor(np.log2(max((d + 7) // 8 for d in dims)))) + 1
    assert res["vertexCountBuffer"] == res["vertexCount"]
    assert res["indexCountBuffer"] == res["indexCount"]
    assert res["meshletDescriptorCountBuffer"] == res["meshletCount"]
    assert res["vertexBuffer_size"] == 32 * res["vertexCount"]
    assert res["indexBuffer_size"] == 4 * res["indexCount"]
    assert res["meshletDescriptorBuffer_size"] == 16 * res["meshletCount"]
    # the renderer layout through its members: every triangle covered once, none dropped
    assert res["meshletIndexBuffer_size"] == 4 * res["indexCount"]
    assert res["meshletVertexBuffer_size"] >= res["vertexBuffer_size"]
    assert res["rendererTriangles"] == res["indexCount"] // 3
    assert res["rendererDropped"] == 0 and res["overfullMeshlets"] == 0
    lines = obj.read_text().splitlines()
    vs = [l for l in lines if l.startswith("v ")]
    vns = [l for l in lines if l.startswith("vn ")]
    fs = [l for l in lines if l.startswith("f ")]
    assert len(vs) == len(vns) == res["vertexCount"] and len(fs) == len(ref["tris"])
    faces = np.array([[int(t.split("//")[0]) - 1 for t in f.split()[1:]] for f in fs])
    np.testing.assert_array_equal(faces, ref["tris"])
    pos = np.array([[float(c) for c in l.split()[1:]] for l in vs])
    np.testing.assert_allclose(pos, ref["ndc"], atol=1e-5)


def test_cpp_demo_rejects_bad_volume_name(tmp_path):
    p = tmp_path / "volume.raw"
    p.write_bytes(b"\0" * 8)
    out = subprocess.run([DEMO, str(p), "1", "8", str(tmp_path / "x.obj")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 1 and "naming scheme" in out.stderr
