"""meshtrex_amd: isosurface extraction on AMD MI355X (gfx950).

Block min/max -> active-block filter -> PMB marching cubes, as hand-written
HIP kernels behind a C ABI (include/meshtrex_amd/mtx.h).  This package holds
the kernels (csrc/), the in-tree build (build.py), the ctypes view of the ABI
(_lib.py), the host API with the reference's manager names (managers.py) and
the batched form (batch.py).
"""
from .managers import (  # noqa: F401
    Context, DeviceBuffer, Volume, MinMaxOutput, FilteringOutput, ExtractionOutput, MtxError,
    computeMinMaxMip, filterActiveBlocks, readActiveBlockCount, extractMeshletDescriptors,
    loadVolume, makePushConstants, makeOptions, read_device, MtxProfiler, writeGPUExtractionToOBJ,
    PinnedBuffer, pinned_array,
)
from .batch import Batch  # noqa: F401,E402
