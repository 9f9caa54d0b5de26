"""Build the HIP libraries in-tree for gfx950.

    python -m meshtrex_amd.build        # or meshtrex_amd.build.build()

Produces, next to this file:
  libmeshtrex_amd.so   the extraction path behind the C ABI (include/meshtrex_amd/mtx.h)
  libmtx_synth.so      the synthetic-volume generator used by the benchmark and tests
Every build recompiles (the libraries are small); objects go to _build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("MTX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Wall",
          "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")] + DEVICE_FLAGS

LIBS = {
    "libmeshtrex_amd.so": ["mtx_kernels.hip", "mtx_batch.hip", "mtx_api.hip", "mtx_batch_api.hip"],
    "libmtx_synth.so": ["synth.hip"],
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_demo(verbose=False):
    """The C++ host-API demo (tests/cpp/managers_demo.cpp): host code only."""
    src = os.path.join(ROOT, "tests", "cpp", "managers_demo.cpp")
    if not os.path.exists(src):
        return
    out = os.path.join(ROOT, "tests", "cpp", "managers_demo")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-o", out,
          "-L" + HERE, "-lmeshtrex_amd", "-Wl,-rpath,$ORIGIN/../../meshtrex_amd"])
    if verbose:
        print("built", out)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for v in LIBS.values() for s in v]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    for lib, parts in LIBS.items():
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] +
             [objs[p] for p in parts])
        os.replace(tmp, out)
        if verbose:
            print("built", out)
    build_demo(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
