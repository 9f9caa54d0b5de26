"""Write tests/golden/reference_cpu_golden.json from the reference's own CPU code.

Run in a container that has the reference tree: it builds oracle/_ref/ref_driver
(oracle/Makefile, target `ref`), which compiles the reference's
src/testMinMax.h and src/mc_tables.h where they lie, and records:
  * the 256^3 u8 blob volume (seed 0x6D747278) hash, generated by the oracle's
    generator (the reference has no volume files: they are not in its tree);
  * computeMinMaxFromFile on that volume (B=4, no overlap, src/testMinMax.h:28-135):
    FNV-1a64 and SHA-256 of all 262,144 pairs plus 1024 sampled pairs;
  * computeActiveBlockCountCPU (src/testMinMax.h:220-242) at isovalues 80, 128;
  * FNV-1a64 of the little-endian int32 bytes of triTable, edgeTable and
    numVerticesTable (src/mc_tables.h).
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEED = 0x6D747278


def fnv1a64(b: bytes) -> str:
    a = np.frombuffer(b, dtype=np.uint8)
    h = 0xCBF29CE484222325
    for x in a.tolist():
        h ^= x
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "ref"])
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libmtx_oracle.so"))
    dims = (ctypes.c_uint32 * 3)(256, 256, 256)
    vol = np.zeros(256 ** 3, np.uint8)
    assert orc.orc_gen_blobs_u8(vol.ctypes.data_as(ctypes.c_void_p), dims, ctypes.c_uint64(SEED), 16) == 0
    out = {"volume": {"recipe": "blobs u8 256^3, 16 blobs", "seed": SEED,
                      "sha256": hashlib.sha256(vol.tobytes()).hexdigest()}}
    with tempfile.TemporaryDirectory() as td:
        vpath = os.path.join(td, "blobs_256x256x256_uint8.raw")
        ppath = os.path.join(td, "pairs.bin")
        vol.tofile(vpath)
        subprocess.check_call([drv, "minmax", vpath, ppath], stdout=subprocess.DEVNULL)
        pairs = np.fromfile(ppath, dtype=np.uint32)
        assert pairs.size == 2 * 64 ** 3
        rng = np.random.default_rng(7)
        idx = np.sort(rng.choice(64 ** 3, 1024, replace=False))
        out["minmax_b4_overlap0"] = {
            "block": 4, "grid": [64, 64, 64],
            "fnv1a64": fnv1a64(pairs.tobytes()),
            "sha256": hashlib.sha256(pairs.tobytes()).hexdigest(),
            "sample_ids": idx.tolist(),
            "sample_pairs": pairs.reshape(-1, 2)[idx].tolist(),
        }
        counts = {}
        for iso in ("80", "128"):
            txt = subprocess.check_output([drv, "count", ppath, iso]).decode()
            line = [l for l in txt.splitlines() if l.startswith("ACTIVE ")][0]
            counts[iso] = int(line.split()[1])
        out["active_count_b4"] = counts
        tpath = os.path.join(td, "tables.bin")
        subprocess.check_call([drv, "tables", tpath])
        raw = np.fromfile(tpath, dtype=np.int32)
        out["tables_fnv1a64_int32le"] = {
            "triTable": fnv1a64(raw[:4096].tobytes()),
            "edgeTable": fnv1a64(raw[4096:4352].tobytes()),
            "numVerticesTable": fnv1a64(raw[4352:4608].tobytes()),
        }
    with open(os.path.join(HERE, "reference_cpu_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps({k: v for k, v in out.items() if k != "minmax_b4_overlap0"}, indent=1))


if __name__ == "__main__":
    sys.exit(main())
