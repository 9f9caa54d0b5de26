"""The C-ABI library loads on a host without a GPU, exports every function
include/meshtrex_amd/mtx.h declares, and the ctypes mirrors match the C
layouts.  No compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

from meshtrex_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include", "meshtrex_amd")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", "meshtrex_amd", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(mtx_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_path():
    names = declared_functions()
    for n in ("mtx_compute_min_max", "mtx_filter_active_blocks", "mtx_read_active_count",
              "mtx_extract", "mtx_context_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    assert set(declared_functions()) <= set(L.exported_symbols())


@pytest.mark.parametrize("name,cls", [
    ("mtx_push_constants", L.PushConstants), ("mtx_volume", L.CVolume),
    ("mtx_options", L.Options), ("mtx_minmax_output", L.CMinMaxOutput),
    ("mtx_filtering_output", L.CFilteringOutput), ("mtx_extraction_output", L.CExtractionOutput),
    ("mtx_batch_volume", L.CBatchVolume), ("mtx_batch_options", L.BatchOptions),
    ("mtx_batch_range", L.BatchRange), ("mtx_batch_output", L.CBatchOutput),
    ("mtx_batch_class_info", L.BatchClassInfo)])
def test_struct_layouts(name, cls):
    assert L.lib().mtx_struct_size(name.encode()) == ctypes.sizeof(cls)


def test_push_constants_are_64_bytes_like_the_reference():
    # src/common.h:29-34: alignas(16) { uvec4 volumeDim, blockDim, blockGridDim; float iso }
    assert ctypes.sizeof(L.PushConstants) == 64
    assert L.lib().mtx_struct_size(b"mtx_vertex") == 32      # VertexData
    assert L.lib().mtx_struct_size(b"mtx_meshlet") == 16     # MeshletDescriptor


def test_push_constants_init_matches_reference_main():
    from meshtrex_amd import makePushConstants
    pc = makePushConstants((1024, 1024, 314), 3, 80.0)   # src/meshtrex.cpp:76-84
    assert list(pc.blockGridDim)[:3] == [342, 342, 105]
    assert list(pc.volumeDim) == [1024, 1024, 314, 1] and pc.isovalue == 80.0


def test_key_roundtrip_and_order():
    import numpy as np
    xs = np.array([-np.inf, -1e30, -1.5, -0.0, 0.0, 1e-40, 2.5, np.inf], np.float32)
    keys = [L.lib().mtx_key_f32(float(x)) for x in xs]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    for x, k in zip(xs, keys):
        assert np.float32(L.lib().mtx_unkey_f32(k)).tobytes() == x.tobytes()


def test_context_create_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert L.lib().mtx_context_create(0, ctypes.byref(h)) == L.MTX_EDEVICE


def test_load_volume_naming_contract(tmp_path):
    import numpy as np
    from meshtrex_amd import loadVolume
    p = tmp_path / "bonsai_4x3x2_uint8.raw"
    np.arange(24, dtype=np.uint8).tofile(p)
    v = loadVolume(p)
    assert v.volume_dims == (4, 3, 2) and v.volume_type == "uint8" and v.volume_data[5] == 5
    q = tmp_path / "sdf_2x2x2_float32.raw"
    np.linspace(-1, 1, 8, dtype=np.float32).tofile(q)
    assert loadVolume(q).volume_data.dtype == np.float32
    with pytest.raises(RuntimeError, match="naming scheme"):
        loadVolume(tmp_path / "noname.raw")
    short = tmp_path / "s_4x4x4_uint8.raw"
    short.write_bytes(b"\0" * 10)
    with pytest.raises(RuntimeError, match="Failed to read"):
        loadVolume(short)


def test_cpp_header_compiles_standalone(tmp_path):
    # managers.hpp + mtx.h need only a C++17 host compiler (no HIP headers)
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text('#include "meshtrex_amd/managers.hpp"\nint main(){ meshtrex::PushConstants pc; return (int)sizeof(pc) - 64; }\n')
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                           "-I" + os.path.join(ROOT, "include"), str(src)])


@pytest.mark.parametrize("torch_first", [False, True])
def test_one_hip_runtime_whatever_the_import_order(torch_first):
    """torch ships its own HIP/HSA runtimes; with either import order the
    process must hold one copy of each (two cannot both open the GPU)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n" % ROOT +
            ("import torch\n" if torch_first else "") +
            "from meshtrex_amd import _lib as L\nL.lib()\n" +
            ("" if torch_first else "import torch\n") +
            "maps = {l.split()[-1] for l in open('/proc/self/maps') if 'amdhip64' in l or 'hsa-runtime' in l}\n"
            "print(sum('libamdhip64' in m for m in maps), sum('libhsa-runtime64' in m for m in maps))\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["1", "1"], out.stdout
