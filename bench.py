"""Benchmark of the isosurface-extraction path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--skip-small] [--skip-batched] [--skip-adjoint] [--skip-ingest]

Headline (`value`): a step is one pass of the path over a resident 1024^3
float32 SDF (64 spheres + 32 tori, N(0, 0.05^2) noise, isovalue 0, 8^3-cell
blocks): block min/max with the fused activity mask, the min/max pyramid,
ordered active-block compaction and PMB marching cubes (count, scan, emit of
vertices with normals, index buffer, meshlets), all on one stream with no
host synchronisation, into buffers sized by a warm-up pass.  `value` is
Gvoxels/s over the K timed steps (hipEvents bracketing them); each step's own
events give the median beside it.  `roofline` describes the block min/max
kernel (the dominant, HBM-bound kernel) from HIP events around its launch
inside the timed steps; stage times come from a separate pass.
`cpu_baseline` times the CPU oracle on a z-slab of the same volume on this
host at T threads (1 thread beside it).

`extra.small_volumes`: the 256^3 uint8 blob volume (isovalue 128) and the
512^3 uint8 vessel volume (isovalue 80), same stages, median of >= 20 steps
with a 512 MiB flush before each; the CPU oracle on the whole volumes at 1
and T threads; the reference's own CPU min/max (compiled from its headers)
against the GPU min/max in the same mode.

`extra.batched`: 48 volumes of uneven dimensions (64-384 per axis, none a
multiple of 4; uint8 blobs and float32 SDFs alternating) with 1-8 isovalues
each, run as one captured launch sequence (mtx_batch_run), against the same
units through the single-volume path launched one after another -- once as the
reference's main calls it (allocating, blocking count readback) and once with
preallocated outputs and no host synchronisation.

`extra.fwd_adjoint`: 16 cubic float32 blob fields (128-256 voxels a side),
isovalue 0, loss sum over vertices of (|p - c| - R)^2; device time of the
forward pass (min/max, compaction, extraction with voxel positions) against
forward + loss gradient + adjoint, and the adjoint's gradient against FP64
central differences on 256 voxels of the first volume.

Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxels/s + %HBM peak, 1024³ f32; batched & fwd+adjoint vs plain, same GPU"
HBM_PEAK_GBS = 8000.0
SEED = 0x6D747278 + 2
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_minmax_b8_1024cube_f32.txt")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse_traffic():
    """HBM bytes per launch of the min/max kernel, from the committed PMC summary."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def cpu_threads():
    """T for the threaded CPU runs: the CPUs this process may use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_path(host, dims, iso, threads):
    """The CPU oracle's min/max, filter and extraction (with normals) over
    z-slabs of 8-plane block rows, one slab per thread.  Each slab carries the
    plane above its last block row, so its min/max pairs and its cells are the
    whole volume's (only the x/y edges of that shared plane are extracted
    twice).  Stages run one after another, each across all threads."""
    import numpy as np
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    X, Y, Z = dims
    rows = -(-Z // 8)
    cuts = np.linspace(0, rows, min(threads, rows) + 1).astype(int)
    slabs, own = [], []
    for a, b in zip(cuts[:-1], cuts[1:]):
        z0, z1 = 8 * int(a), min(Z, 8 * int(b) + 1)
        slabs.append((host[z0 * X * Y: z1 * X * Y], (X, Y, z1 - z0)))
        own.append(-(-X // 8) * -(-Y // 8) * int(b - a))   # blocks of the slab's own rows
    dt = oracle.U8 if host.dtype == np.uint8 else oracle.F32
    with ThreadPoolExecutor(len(slabs)) as ex:
        t0 = time.perf_counter()
        pairs = list(ex.map(lambda sl: oracle.minmax(sl[0], sl[1], 8, 1), slabs))
        t1 = time.perf_counter()
        ids = list(ex.map(lambda pr: oracle.filter_ids(pr, dt, iso), pairs))
        t2 = time.perf_counter()
        meshes = list(ex.map(lambda sl: oracle.extract(sl[0], sl[1], 8, iso, normals=True), slabs))
        t3 = time.perf_counter()
    nvox = X * Y * Z
    tris = sum(len(m["tris"]) for m in meshes)
    return {"threads": len(slabs), "minmax_s": round(t1 - t0, 4), "filter_s": round(t2 - t1, 4),
            "extract_s": round(t3 - t2, 4), "total_s": round(t3 - t0, 4),
            "Gvoxels_per_s": round(nvox / (t3 - t0) / 1e9, 5), "triangles_per_s": round(tris / (t3 - t0), 1),
            "triangles": tris, "active_blocks": sum(int(np.count_nonzero(x < n)) for x, n in zip(ids, own))}


def cpu_both(host, dims, iso, T):
    one = cpu_path(host, dims, iso, 1)
    many = cpu_path(host, dims, iso, T) if T > 1 else one
    return {"1": one, "T": many}


class Path:
    """One resident volume through the path -- block min/max with the fused
    activity mask, the min/max pyramid, ordered compaction, PMB extraction
    with normals (and meshlets unless meshlets=0) -- into outputs sized by a
    first, allocating pass; one stream, no host synchronisation in a step."""

    def __init__(self, ctx, stream, vol_ptr, dims, dtype, iso, meshlets=1):
        from meshtrex_amd import _lib as L
        from meshtrex_amd import hiprt
        from meshtrex_amd.managers import (ExtractionOutput, FilteringOutput, MinMaxOutput, _check, makeOptions,
                                           makePushConstants)
        self.ctx, self.stream, self.dims = ctx, stream, dims
        self.nvox = dims[0] * dims[1] * dims[2]
        self.esize = 4 if dtype == "float32" else 1
        self.pc = makePushConstants(dims, 8, iso)
        self.cvol = L.CVolume()
        self.cvol.data, self.cvol.dims[:], self.cvol.on_device = vol_ptr, dims, 1
        self.cvol.dtype = L.MTX_F32 if dtype == "float32" else L.MTX_U8
        self.o_mm = makeOptions(build_pyramid=0, reuse=1)
        self.o_ex = makeOptions(reuse=1, meshlets=meshlets)
        self.mm, self.filt, self.ex = MinMaxOutput(ctx), FilteringOutput(ctx), ExtractionOutput(ctx)
        self.lib, self.h, self._check = L.lib(), ctx.handle, _check
        self._call(self.lib.mtx_compute_min_max, ctypes.byref(self.cvol), ctypes.byref(self.pc),
                   ctypes.byref(self.o_mm), stream, ctypes.byref(self.mm.c))
        self._call(self.lib.mtx_build_pyramid, ctypes.byref(self.mm.c), stream)
        self._call(self.lib.mtx_filter_active_blocks, ctypes.byref(self.mm.c), ctypes.byref(self.pc), None, stream,
                   ctypes.byref(self.filt.c))
        self._call(self.lib.mtx_extract, ctypes.byref(self.mm.c), ctypes.byref(self.filt.c), ctypes.byref(self.pc),
                   ctypes.byref(makeOptions(meshlets=meshlets)), stream, ctypes.byref(self.ex.c))
        hiprt.stream_sync(stream)

    def _call(self, fn, *a):
        self._check(self.ctx, fn(self.h, *a))

    def step(self, ev=(None,) * 5):
        """ev: events recorded before min/max, after min/max, after the
        pyramid, after the filter and after the extraction (None: skipped)."""
        s = self.stream
        if ev[0]:
            ev[0].record(s)
        self._call(self.lib.mtx_compute_min_max, ctypes.byref(self.cvol), ctypes.byref(self.pc),
                   ctypes.byref(self.o_mm), s, ctypes.byref(self.mm.c))
        if ev[1]:
            ev[1].record(s)
        self._call(self.lib.mtx_build_pyramid, ctypes.byref(self.mm.c), s)
        if ev[2]:
            ev[2].record(s)
        self._call(self.lib.mtx_filter_active_blocks, ctypes.byref(self.mm.c), ctypes.byref(self.pc), None, s,
                   ctypes.byref(self.filt.c))
        if ev[3]:
            ev[3].record(s)
        self._call(self.lib.mtx_extract, ctypes.byref(self.mm.c), ctypes.byref(self.filt.c), ctypes.byref(self.pc),
                   ctypes.byref(self.o_ex), s, ctypes.byref(self.ex.c))
        if ev[4]:
            ev[4].record(s)

    def counts(self):
        nv, ni, nm = self.ex.read_counts()
        self._check(self.ctx, self.lib.mtx_read_active_count(self.h, ctypes.byref(self.filt.c)))
        return {"active_blocks": int(self.filt.activeBlockCount), "vertices": nv, "triangles": ni // 3,
                "meshlets": nm, "meshlet_vertices": self.ex.meshletVertexCount}

    def close(self):
        self.ex.cleanup()
        self.filt.cleanup()
        self.mm.cleanup()


def time_path(path, steps, warmup, flush=None):
    """hipEvent times of `steps` steps after `warmup`: the bracket over all of
    them, each step's time (from its own start event to the next step's, the
    last to the closing event) and its min/max launch's time (events before
    and after that launch).  An event recorded inside a step is a marker the
    GPU stops at (a few microseconds each on MI355X), so a step carries just
    these two.  With `flush`, a 512 MiB device write runs before each step,
    outside the step's events (and inside the bracket, which then is not a
    per-step figure)."""
    import statistics
    from meshtrex_amd import hiprt
    s = path.stream
    for _ in range(warmup):
        if flush:
            flush()
        path.step()
    hiprt.stream_sync(s)
    evs = [[hiprt.Event() for _ in range(2)] for _ in range(steps)]
    ends = [hiprt.Event() for _ in range(steps)] if flush else None
    start, stop = hiprt.Event(), hiprt.Event()
    w0 = time.perf_counter()
    start.record(s)
    for k in range(steps):
        if flush:
            flush()
        path.step((evs[k][0], evs[k][1], None, None, ends[k] if flush else None))
    stop.record(s)
    stop.sync()
    wall = time.perf_counter() - w0
    if flush:
        per = [ends[k].ms_since(evs[k][0]) for k in range(steps)]
    else:
        per = [(evs[k + 1][0] if k + 1 < steps else stop).ms_since(evs[k][0]) for k in range(steps)]
    mm = [e[1].ms_since(e[0]) for e in evs]
    return {"bracket_ms": stop.ms_since(start) / steps, "median_ms": statistics.median(per),
            "mean_ms": statistics.fmean(per), "min_ms": min(per), "max_ms": max(per),
            "minmax_mean_ms": statistics.fmean(mm), "minmax_median_ms": statistics.median(mm), "wall_s": wall}


def stage_times(path, steps):
    """Mean time of each stage, from events between the stages (a pass of its own)."""
    from meshtrex_amd import hiprt
    evs = [[hiprt.Event() for _ in range(5)] for _ in range(steps)]
    for k in range(steps):
        path.step(evs[k])
    evs[-1][4].sync()
    names = ("minmax_leaf", "pyramid", "filter", "extract")
    return {n: round(sum(e[i + 1].ms_since(e[i]) for e in evs) / steps, 4) for i, n in enumerate(names)}


def make_flush(stream):
    """A 512 MiB device copy: evicts the 256 MiB Infinity Cache between steps."""
    from meshtrex_amd import _lib as L
    from meshtrex_amd.managers import DeviceBuffer
    n = 512 << 20
    a, b = DeviceBuffer(n), DeviceBuffer(n)

    def flush():
        assert L.synth().mtx_bench_copy(b.ptr, a.ptr, n, stream) == 0
    flush.buffers = (a, b)
    return flush


def bench_plain(args, ctx, stream):
    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt
    from meshtrex_amd.managers import DeviceBuffer
    S = args.size
    dims = (S, S, S)
    nvox = S ** 3
    vol = DeviceBuffer(4 * nvox)
    d3 = (L.u32 * 3)(*dims)
    t0 = time.perf_counter()
    rc = L.synth().mtx_synth_sdf_f32(vol.ptr, d3, SEED, 64, 32, S / 1024.0, 0.05, stream)
    assert rc == 0, rc
    gen_s = time.perf_counter() - t0
    path = Path(ctx, stream, vol.ptr, dims, "float32", 0.0, meshlets=1)
    t = time_path(path, args.steps, args.warmup)
    stages = stage_times(path, args.steps)
    cnt = path.counts()
    ms = t["bracket_ms"]

    # the same step writing the shared mesh only (no meshlets)
    p0 = Path(ctx, stream, vol.ptr, dims, "float32", 0.0, meshlets=0)
    t0m = time_path(p0, args.steps, args.warmup)
    st0 = stage_times(p0, args.steps)
    p0.close()

    # achievable HBM rates on this GPU: a read-only stream over the 4 GiB volume
    # and a plain 1 GiB device copy (read + write), same stream
    sink = DeviceBuffer(16384 * 4 * 4)
    for _ in range(2):
        assert L.synth().mtx_bench_read(vol.ptr, 4 * nvox, sink.ptr, stream) == 0
    c0, c1 = hiprt.Event(), hiprt.Event()
    c0.record(stream)
    for _ in range(5):
        assert L.synth().mtx_bench_read(vol.ptr, 4 * nvox, sink.ptr, stream) == 0
    c1.record(stream)
    c1.sync()
    read_gbs = 4 * nvox * 5 / (c1.ms_since(c0) * 1e-3) / 1e9
    sink.free()
    cbytes = 1 << 30
    ca, cb = DeviceBuffer(cbytes), DeviceBuffer(cbytes)
    for _ in range(2):
        assert L.synth().mtx_bench_copy(cb.ptr, ca.ptr, cbytes, stream) == 0
    c0.record(stream)
    for _ in range(5):
        assert L.synth().mtx_bench_copy(cb.ptr, ca.ptr, cbytes, stream) == 0
    c1.record(stream)
    c1.sync()
    copy_gbs = 2 * cbytes * 5 / (c1.ms_since(c0) * 1e-3) / 1e9
    ca.free()
    cb.free()

    nv, ntri, nm, nmv, na = (cnt["vertices"], cnt["triangles"], cnt["meshlets"], cnt["meshlet_vertices"],
                             cnt["active_blocks"])
    nb = ((S + 7) // 8) ** 3
    leaf_bytes = 4 * nvox + 8 * nb
    mmf_bytes = 4 * nvox + 8 * nb + 4 * na + 4
    ext_bytes = na * 8 ** 3 * 4 + 32 * nv + 12 * ntri + 16 * nm
    # everything the extraction writes: the shared mesh, the meshlet descriptors
    # and the renderer layout (meshlet vertices, their map, meshlet indices)
    ext_out_bytes = 32 * nv + 12 * ntri + 16 * nm + (32 + 4) * nmv + 12 * ntri
    # the roofline kernel's time: its own events, inside the timed steps
    leaf_gbs = leaf_bytes / (t["minmax_mean_ms"] * 1e-3) / 1e9
    mmf_ms = stages["minmax_leaf"] + stages["filter"]
    mmf_gbs = mmf_bytes / (mmf_ms * 1e-3) / 1e9
    res = {
        "ms": ms, "nvox": nvox, "dims": dims,
        "config": {"workload": "plain single extraction, %d^3 float32 SDF (64 spheres + 32 tori + "
                               "N(0,0.05^2)), isovalue 0, 8^3-cell blocks; min-max + pyramid + "
                               "filter + PMB extraction with normals and meshlets" % S,
                   "volume_bytes": 4 * nvox, "blocks": nb, "active_blocks": na,
                   "vertices": nv, "triangles": ntri, "meshlets": nm, "meshlet_vertices": nmv},
        "roofline": {"bound": "hbm", "achieved": round(leaf_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(leaf_gbs / HBM_PEAK_GBS, 4), "traffic": parse_traffic(),
                     "kernel": "k_minmax_b8<float> (block min/max + fused activity mask)",
                     "kernel_ms": round(t["minmax_mean_ms"], 4),
                     "algorithmic_bytes_per_launch": leaf_bytes,
                     "achievable_read_GBps": round(read_gbs, 1),
                     "frac_of_achievable_read": round(leaf_gbs / read_gbs, 4),
                     "achievable_copy_GBps": round(copy_gbs, 1)},
        "extra": {
            "step_ms": {k: round(t[k], 4) for k in ("bracket_ms", "median_ms", "mean_ms", "min_ms", "max_ms")},
            "Gvoxels_per_s_at_median": round(nvox / (t["median_ms"] * 1e-3) / 1e9, 3),
            "stage_ms": stages,
            "stage_ms_note": "a separate pass with events between the stages, after the timed steps",
            "minmax_plus_filter": {"bytes": mmf_bytes, "ms": round(mmf_ms, 4),
                                   "GBps": round(mmf_gbs, 1), "frac_of_peak": round(mmf_gbs / HBM_PEAK_GBS, 4),
                                   "target_frac": 0.70},
            "extraction": {"algorithmic_bytes": ext_bytes, "ms": stages["extract"],
                           "GBps": round(ext_bytes / (stages["extract"] * 1e-3) / 1e9, 1),
                           "bytes_in_plus_all_outputs": na * 8 ** 3 * 4 + ext_out_bytes,
                           "GBps_all_outputs": round((na * 8 ** 3 * 4 + ext_out_bytes) /
                                                     (stages["extract"] * 1e-3) / 1e9, 1)},
            "shared_mesh_only": {"ms": round(t0m["bracket_ms"], 4), "median_ms": round(t0m["median_ms"], 4),
                                 "Gvoxels_per_s": round(nvox / (t0m["bracket_ms"] * 1e-3) / 1e9, 3),
                                 "extract_ms": st0["extract"],
                                 "note": "the same step with meshlets=0: vertices, normals and the index "
                                         "buffer, no meshlet descriptors or renderer layout"},
            "triangles_per_s": round(ntri / (ms * 1e-3), 1),
            "wall_s_timed": round(t["wall_s"], 4),
            "generate_s": round(gen_s, 3),
        },
        "_keep": (vol, path),
    }
    return res


def bench_small(args, ctx, stream, kind, dims, seed, iso, flush, name, workload):
    """A workload that fits the 256 MiB Infinity Cache: median of >= 20 steps,
    each after a 512 MiB flush."""
    from meshtrex_amd import workloads
    buf, vol = workloads.generate(dims, kind, seed, stream)
    path = Path(ctx, stream, buf.ptr, dims, vol.volume_type, iso, meshlets=1)
    steps = max(20, args.steps)
    t = time_path(path, steps, 3, flush=flush)
    cnt = path.counts()
    path.close()
    nvox = path.nvox
    med = t["median_ms"]
    log("%s: median %.4f ms/step, %.1f Gvox/s" % (name, med, nvox / (med * 1e-3) / 1e9))
    out = {"workload": workload, "steps": steps, "warmup": 3,
           "median_ms": round(med, 4), "mean_ms": round(t["mean_ms"], 4), "min_ms": round(t["min_ms"], 4),
           "Gvoxels_per_s": round(nvox / (med * 1e-3) / 1e9, 3),
           "triangles_per_s": round(cnt["triangles"] / (med * 1e-3), 1),
           "minmax_median_ms": round(t["minmax_median_ms"], 4),
           "flush": "512 MiB device copy before each step, outside the step's events", **cnt}
    return out, buf


def bench_reference_minmax(ctx, stream, buf, dims, flush):
    """The reference's own CPU min/max + active count (src/testMinMax.h:28-135,
    220-242; compiled from the reference's headers) on the 256^3 blob volume,
    beside the GPU min/max in the same mode (4^3 blocks, no overlap)."""
    import statistics
    import subprocess
    import tempfile
    import numpy as np
    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt
    from meshtrex_amd.managers import MinMaxOutput, makeOptions, makePushConstants, read_device
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    res = {}
    if os.path.exists(drv):
        host = read_device(buf.ptr, np.uint8, int(np.prod(dims)))
        with tempfile.TemporaryDirectory() as d:
            f = os.path.join(d, "blobs256.raw")
            host.tofile(f)
            r = subprocess.run([drv, "time", f, "128", "21"], capture_output=True, text=True, timeout=120)
        if r.returncode == 0 and r.stdout.startswith("SECONDS"):
            w = r.stdout.split()
            res["reference_cpu"] = {"ms": round(1e3 * float(w[1]), 3), "active_blocks": int(w[3]), "threads": 1,
                                    "what": "computeMinMaxFromFile (file read included) + "
                                            "computeActiveBlockCountCPU, median of 21 runs"}
        else:
            res["reference_cpu"] = {"error": (r.stdout + r.stderr)[-300:]}
    else:
        res["reference_cpu"] = {"error": "oracle/_ref/ref_driver not built"}
    pc = makePushConstants(dims, 4, 128.0)
    cv = L.CVolume()
    cv.data, cv.dims[:], cv.dtype, cv.on_device = buf.ptr, dims, L.MTX_U8, 1
    o = makeOptions(build_pyramid=0, reuse=1, overlap=0, fuse_filter=0)
    mm = MinMaxOutput(ctx)
    lib = L.lib()
    assert lib.mtx_compute_min_max(ctx.handle, ctypes.byref(cv), ctypes.byref(pc), ctypes.byref(o), stream,
                                   ctypes.byref(mm.c)) == 0
    per = []
    for _ in range(21):
        flush()
        e0, e1 = hiprt.Event(), hiprt.Event()
        e0.record(stream)
        assert lib.mtx_compute_min_max(ctx.handle, ctypes.byref(cv), ctypes.byref(pc), ctypes.byref(o), stream,
                                       ctypes.byref(mm.c)) == 0
        e1.record(stream)
        e1.sync()
        per.append(e1.ms_since(e0))
    mm.cleanup()
    res["gpu_minmax_same_mode"] = {"ms": round(statistics.median(per), 4),
                                   "what": "mtx_compute_min_max, 4^3 blocks, no overlap, median of 21 after a "
                                           "512 MiB flush each"}
    if "ms" in res["reference_cpu"]:
        res["speedup"] = round(res["reference_cpu"]["ms"] / res["gpu_minmax_same_mode"]["ms"], 1)
    return res


def bench_ingest(args, ctx, vol_ptr, dims, minmax_ms):
    """Host volume -> min/max (two slabs, the head's min/max running while the
    tail uploads), from pageable and pinned host memory, beside a bare pinned
    copy and the device-resident min/max; and a copy of the whole volume with
    back-to-back min/max launches over the resident volume on another stream,
    which shows what the overlap costs the copy."""
    import numpy as np
    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt
    from meshtrex_amd.managers import (DeviceBuffer, MinMaxOutput, Volume, _check, makeOptions, makePushConstants,
                                       pinned_array, read_device)
    nvox = dims[0] * dims[1] * dims[2]
    nbytes = 4 * nvox
    host = read_device(vol_ptr, np.float32, nvox)
    pin = pinned_array(nvox, np.float32)
    pin[:] = host
    lib, h = L.lib(), ctx.handle
    pc = makePushConstants(dims, 8, 0.0)
    o = makeOptions(build_pyramid=0, reuse=1)
    res = {"volume_bytes": nbytes}
    for name, arr in (("pageable", host), ("pinned", pin)):
        v = Volume(dims, "float32", arr)._c()
        mm = MinMaxOutput(ctx)
        times = []
        for _ in range(4):
            t0 = time.perf_counter()
            _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(v), ctypes.byref(pc), ctypes.byref(o), None,
                                                ctypes.byref(mm.c)))
            times.append(time.perf_counter() - t0)
        mm.cleanup()
        best = min(times[1:])   # the first call also allocates the device copy
        res[name] = {"ms": round(best * 1e3, 2), "GBps": round(nbytes / best / 1e9, 2)}
    d = DeviceBuffer(nbytes)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        assert lib.mtx_memcpy_h2d(d.ptr, pin.ctypes.data, nbytes) == 0
        times.append(time.perf_counter() - t0)
    copy_s = min(times)
    res["pinned_copy_only"] = {"ms": round(copy_s * 1e3, 2), "GBps": round(nbytes / copy_s / 1e9, 2),
                               "what": "one synchronous copy of the whole volume"}
    res["device_resident_minmax_ms"] = round(minmax_ms, 3)
    res["pinned_minus_copy_only_ms"] = round(res["pinned"]["ms"] - copy_s * 1e3, 3)
    # the copy (events on its stream) alone, then with min/max launches over
    # the resident volume back to back on a second stream during it
    hip = hiprt.hip()
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    sc, sk = hiprt.stream_create(), hiprt.stream_create()
    e0, e1 = hiprt.Event(), hiprt.Event()
    cv = L.CVolume()
    cv.data, cv.dims[:], cv.on_device, cv.dtype = vol_ptr, dims, 1, L.MTX_F32
    mm = MinMaxOutput(ctx)
    _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(cv), ctypes.byref(pc), ctypes.byref(o), sk, ctypes.byref(mm.c)))
    hiprt.stream_sync(sk)
    k0, k1 = hiprt.Event(), hiprt.Event()
    out = {}
    for label, nk in (("alone", 0), ("with_minmax", 60)):
        best = None
        for _ in range(2):
            e0.record(sc)
            assert hip.hipMemcpyAsync(d.ptr, pin.ctypes.data, nbytes, 1, sc) == 0
            e1.record(sc)
            k0.record(sk)
            for _ in range(nk):
                _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(cv), ctypes.byref(pc), ctypes.byref(o), sk,
                                                    ctypes.byref(mm.c)))
            k1.record(sk)
            hiprt.stream_sync(sc)
            hiprt.stream_sync(sk)
            ms = e1.ms_since(e0)
            best = ms if best is None else min(best, ms)
            kms = k1.ms_since(k0)
        out[label] = {"copy_ms": round(best, 3), "minmax_launches": nk, "minmax_ms_total": round(kms, 3)}
    a, w = out["alone"], out["with_minmax"]
    # how much longer the copy took per ms of min/max running beside it
    out["copy_slowdown_ms_per_minmax_ms"] = round((w["copy_ms"] - a["copy_ms"]) / max(w["minmax_ms_total"], 1e-9), 3)
    out["note"] = ("copy_slowdown_ms_per_minmax_ms near 0: a full-rate min/max beside the copy does not slow it, "
                   "so the upload's min/max can run under the copy of the rest")
    res["copy_under_minmax"] = out
    mm.cleanup()
    hiprt.stream_destroy(sc)
    hiprt.stream_destroy(sk)
    d.free()
    return res


def bench_batched(args, ctx, stream):
    import numpy as np
    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt, workloads
    from meshtrex_amd.batch import Batch
    from meshtrex_amd.managers import (DeviceBuffer, ExtractionOutput, FilteringOutput, MinMaxOutput, _check,
                                       computeMinMaxMip, extractMeshletDescriptors, filterActiveBlocks,
                                       makeOptions, makePushConstants, readActiveBlockCount)
    specs = workloads.batched_specs(args.batched_volumes)
    vols = []
    t0 = time.perf_counter()
    for dims, kind, seed, fr in specs:
        buf, vol = workloads.generate(dims, kind, seed, stream)
        vols.append((buf, vol, dims, kind, workloads.isovalues(buf, dims, kind, fr)))
    gen_s = time.perf_counter() - t0
    units = sum(len(v[4]) for v in vols)
    voxels = sum(int(np.prod(v[2])) for v in vols)
    unit_voxels = sum(int(np.prod(v[2])) * len(v[4]) for v in vols)
    nbytes = sum(int(np.prod(v[2])) * (1 if v[3] == "u8" else 4) for v in vols)
    log("batched: %d volumes, %d units, %.3g voxels, generated in %.1f s" % (len(vols), units, voxels, gen_s))

    t0 = time.perf_counter()
    b = Batch(ctx, [(v[1], v[4]) for v in vols], block_dim=args.batch_block, tune_iters=3)
    plan_s = time.perf_counter() - t0
    log("batched: planned (autotuned) in %.1f s" % plan_s)
    for _ in range(args.warmup):
        b.run(stream)
    hiprt.stream_sync(stream)
    e0, e1 = hiprt.Event(), hiprt.Event()
    # a 4 KiB copy kernel (k_copy) before and after the timed replays marks
    # them in a kernel trace
    marker = DeviceBuffer(1 << 16)
    assert L.synth().mtx_bench_copy(marker.ptr, marker.ptr + 8192, 4096, stream) == 0
    e0.record(stream)
    for _ in range(args.batched_steps):
        b.run(stream)
    e1.record(stream)
    assert L.synth().mtx_bench_copy(marker.ptr, marker.ptr + 8192, 4096, stream) == 0
    e1.sync()
    marker.free()
    batched_ms = e1.ms_since(e0) / args.batched_steps
    ranges = b.read_ranges()
    tris = sum(r["index_count"] for r in ranges) // 3
    verts = sum(r["vertex_count"] for r in ranges)

    # the same units through the single-volume path, preallocated, stream-ordered
    lib, h = L.lib(), ctx.handle
    o_mm = makeOptions(build_pyramid=0, reuse=1)
    # the batch writes the shared mesh only (no meshlets): so do these
    o_ex_alloc, o_ex = makeOptions(meshlets=0), makeOptions(reuse=1, meshlets=0)
    plan = []
    for buf, vol, dims, kind, isos in vols:
        cv = vol._c()
        pcs = [makePushConstants(dims, 8, iso) for iso in isos]
        plan.append((cv, pcs, MinMaxOutput(ctx), [FilteringOutput(ctx) for _ in isos],
                     [ExtractionOutput(ctx) for _ in isos]))

    def sequential(o_extract):
        for cv, pcs, mm, fs, exs in plan:
            _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(cv), ctypes.byref(pcs[0]), ctypes.byref(o_mm),
                                                stream, ctypes.byref(mm.c)))
            for pc, f, ex in zip(pcs, fs, exs):
                _check(ctx, lib.mtx_filter_active_blocks(h, ctypes.byref(mm.c), ctypes.byref(pc), None, stream,
                                                         ctypes.byref(f.c)))
                _check(ctx, lib.mtx_extract(h, ctypes.byref(mm.c), ctypes.byref(f.c), ctypes.byref(pc),
                                            ctypes.byref(o_extract), stream, ctypes.byref(ex.c)))

    sequential(o_ex_alloc)          # sizes every unit's outputs
    for _ in range(args.warmup):
        sequential(o_ex)
    hiprt.stream_sync(stream)
    e0.record(stream)
    h0 = time.perf_counter()
    for _ in range(args.batched_steps):
        sequential(o_ex)
    host_ms = (time.perf_counter() - h0) * 1e3 / args.batched_steps
    e1.record(stream)
    e1.sync()
    seq_ms = e1.ms_since(e0) / args.batched_steps
    seq_tris = 0
    for _, _, mm, fs, exs in plan:
        for ex in exs:
            seq_tris += ex.read_counts()[1] // 3
    for _, _, mm, fs, exs in plan:
        for ex in exs:
            ex.cleanup()
        for f in fs:
            f.cleanup()
        mm.cleanup()

    # the reference's call pattern per unit (src/meshtrex.cpp:107-132): allocate,
    # min/max with pyramid, filter, blocking count readback, extract, free
    hiprt.stream_sync(stream)
    w0 = time.perf_counter()
    for buf, vol, dims, kind, isos in vols:
        for iso in isos:
            pc = makePushConstants(dims, 8, iso)
            mm = computeMinMaxMip(ctx, vol, pc)
            f = filterActiveBlocks(ctx, mm, pc)
            readActiveBlockCount(ctx, f)
            ex = extractMeshletDescriptors(ctx, mm, f, pc)
            ex.cleanup()
            f.cleanup()
            mm.cleanup()
    api_ms = (time.perf_counter() - w0) * 1e3
    infos = b.class_info()
    b.close()
    for v in vols:
        v[0].free()
    gv = lambda ms: round(unit_voxels / (ms * 1e-3) / 1e9, 2)
    return {
        "workload": "%d volumes (dims 64-384 per axis, no multiple of 4; uint8 16-blob fields and float32 "
                    "SDFs alternating), 1-8 isovalues each in [min, max]: %d (volume, isovalue) units, "
                    "%d voxels, %d voxel-isovalue pairs, %.3f GB of voxels"
                    % (len(vols), units, voxels, unit_voxels, nbytes / 1e9),
        "units": units, "vertices": verts, "triangles": tris, "triangles_sequential": seq_tris,
        "batched": {"ms": round(batched_ms, 3), "Gvoxels_per_s": gv(batched_ms),
                    "triangles_per_s": round(tris / (batched_ms * 1e-3), 1),
                    "plan_s": round(plan_s, 2), "steps": args.batched_steps,
                    "classes": [{k: (round(x, 4) if isinstance(x, float) else
                                     [round(y, 4) for y in x] if isinstance(x, list) else x)
                                 for k, x in c.items()} for c in infos]},
        "sequential_plain_preallocated": {"ms": round(seq_ms, 3), "Gvoxels_per_s": gv(seq_ms),
                                          "host_enqueue_ms": round(host_ms, 3), "block_dim": 8},
        "sequential_plain_api": {"ms": round(api_ms, 3), "Gvoxels_per_s": gv(api_ms), "block_dim": 8,
                                 "note": "wall time, allocating outputs and blocking on the count per unit"},
        "speedup_vs_sequential_preallocated": round(seq_ms / batched_ms, 2),
        "speedup_vs_sequential_api": round(api_ms / batched_ms, 2),
        "target_speedup": 3.0,
        "generate_s": round(gen_s, 2),
    }


def bench_adjoint(args, ctx):
    import numpy as np
    import torch
    from meshtrex_amd import _lib as L
    from meshtrex_amd import workloads
    from meshtrex_amd.autograd import Extraction, marching_cubes
    from meshtrex_amd.managers import (ExtractionOutput, FilteringOutput, MinMaxOutput, _check, makeOptions,
                                       makePushConstants)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream().cuda_stream
    lib, h = L.lib(), ctx.handle
    specs = workloads.differentiable_specs(args.adjoint_volumes)
    items = []
    o_mm = makeOptions(build_pyramid=0, reuse=1)
    o_alloc = makeOptions(voxel_positions=1, normals=0, meshlets=0)
    o_ex = makeOptions(voxel_positions=1, normals=0, reuse=1, meshlets=0)
    for dims, seed in specs:
        vol = torch.empty(dims[::-1], dtype=torch.float32, device=dev)
        d3 = (L.u32 * 3)(*dims)
        assert L.synth().mtx_synth_blobs_f32(vol.data_ptr(), d3, seed, 16, stream) == 0
        cv = L.CVolume()
        cv.data, cv.dims[:], cv.dtype, cv.on_device = vol.data_ptr(), dims, L.MTX_F32, 1
        pc = makePushConstants(dims, 8, 0.0)
        mm, f, ex = MinMaxOutput(ctx), FilteringOutput(ctx), ExtractionOutput(ctx)
        _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(cv), ctypes.byref(pc), ctypes.byref(o_mm), stream,
                                            ctypes.byref(mm.c)))
        _check(ctx, lib.mtx_filter_active_blocks(h, ctypes.byref(mm.c), ctypes.byref(pc), None, stream,
                                                 ctypes.byref(f.c)))
        _check(ctx, lib.mtx_extract(h, ctypes.byref(mm.c), ctypes.byref(f.c), ctypes.byref(pc),
                                    ctypes.byref(o_alloc), stream, ctypes.byref(ex.c)))
        nv = ex.c.vertexCount
        c = torch.tensor([(d - 1) / 2.0 for d in dims], dtype=torch.float32, device=dev)
        items.append({"vol": vol, "cv": cv, "pc": pc, "mm": mm, "f": f, "ex": ex, "nv": nv, "c": c,
                      "R": 0.3 * min(dims), "pos": torch.empty((nv, 3), dtype=torch.float32, device=dev),
                      "dvol": torch.empty_like(vol), "diso": torch.empty((), dtype=torch.float32, device=dev)})

    def fwd(it):
        _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(it["cv"]), ctypes.byref(it["pc"]), ctypes.byref(o_mm),
                                            stream, ctypes.byref(it["mm"].c)))
        _check(ctx, lib.mtx_filter_active_blocks(h, ctypes.byref(it["mm"].c), ctypes.byref(it["pc"]), None,
                                                 stream, ctypes.byref(it["f"].c)))
        _check(ctx, lib.mtx_extract(h, ctypes.byref(it["mm"].c), ctypes.byref(it["f"].c), ctypes.byref(it["pc"]),
                                    ctypes.byref(o_ex), stream, ctypes.byref(it["ex"].c)))

    def adj(it):
        nv = it["nv"]
        if nv:
            _check(ctx, lib.mtx_memcpy_d2d_async(it["pos"].data_ptr(), it["ex"].c.vertex_voxel_pos, 12 * nv,
                                                 stream))
        d = it["pos"] - it["c"]
        r = d.norm(dim=1, keepdim=True)
        g = (2.0 * (r - it["R"]) / r) * d
        _check(ctx, lib.mtx_adjoint(h, ctypes.byref(it["mm"].c), ctypes.byref(it["f"].c), ctypes.byref(it["ex"].c),
                                    ctypes.byref(it["pc"]), g.data_ptr() if nv else None, it["dvol"].data_ptr(),
                                    it["diso"].data_ptr(), stream))

    for _ in range(args.warmup):
        for it in items:
            fwd(it)
            adj(it)
    torch.cuda.synchronize()
    reps = args.adjoint_steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in items for _ in range(reps)]
    k = 0
    for _ in range(reps):
        for it in items:
            ev[k][0].record()
            fwd(it)
            ev[k][1].record()
            adj(it)
            ev[k][2].record()
            k += 1
    torch.cuda.synchronize()
    f_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / reps
    a_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / reps
    voxels = sum(int(np.prod(d)) for d, _ in specs)
    verts = sum(it["nv"] for it in items)

    # accuracy on the first volume through the torch.autograd binding: FP64
    # central differences (Richardson-extrapolated, h = 1e-3) of the FP64 loss
    # with the topology frozen, at 256 voxels that are endpoints of active edges
    it = items[0]
    dims = specs[0][0]
    vol = it["vol"].clone().requires_grad_(True)
    pos, _ = marching_cubes(vol, 0.0)
    cc = torch.tensor([(d - 1) / 2.0 for d in dims], dtype=torch.float64, device=dev)
    loss = ((pos.double() - cc).norm(dim=1) - it["R"]).pow(2).sum()
    loss.backward()
    g = vol.grad.detach().cpu().numpy().reshape(-1).astype(np.float64)
    e = Extraction(vol.detach(), 0.0, 8, edge_ids=True).edge_ids().cpu().numpy().astype(np.int64)
    base = vol.detach().cpu().numpy().reshape(-1).astype(np.float64)
    ax, vox = e % 3, e // 3
    st = np.array([1, dims[0], dims[0] * dims[1]])[ax]
    owner = np.stack([vox % dims[0], (vox // dims[0]) % dims[1], vox // (dims[0] * dims[1])], 1).astype(np.float64)
    c64 = np.array([(d - 1) / 2.0 for d in dims])
    R = it["R"]

    def L64(vals):
        t = (0.0 - vals[vox]) / (vals[vox + st] - vals[vox])
        p = owner.copy()
        p[np.arange(e.size), ax] += t
        return float(np.sum((np.linalg.norm(p - c64, axis=1) - R) ** 2))

    hstep = 1e-3
    cand = np.unique(np.concatenate([vox, vox + st]))
    cand = cand[np.abs(base[cand]) > 10 * hstep]
    pick = np.random.default_rng(7).choice(cand, min(256, cand.size), replace=False)
    gmax = np.abs(g).max()
    floor = 1e-6 * gmax
    worst = 0.0
    raw = 0.0
    absmax = 0.0
    for i in pick.tolist():
        def f(dh):
            vv = base.copy()
            vv[i] += dh
            return L64(vv)

        def fd(hh):
            return (f(hh) - f(-hh)) / (2 * hh)
        d = (4 * fd(hstep / 2) - fd(hstep)) / 3
        worst = max(worst, (abs(d - g[i]) - floor) / max(abs(d), 1e-300))
        absmax = max(absmax, abs(d - g[i]))
        if abs(d) >= 1e-3 * gmax:
            raw = max(raw, abs(d - g[i]) / abs(d))
    for it in items:
        it["ex"].cleanup()
        it["f"].cleanup()
        it["mm"].cleanup()
    return {
        "workload": "%d cubic float32 16-blob fields rescaled to [-1,1] (edges %s), isovalue 0, loss "
                    "sum_v (|p-c| - 0.3 min dim)^2; forward = min/max + filter + extraction with voxel "
                    "positions; adjoint = position copy + loss gradient (torch) + mtx_adjoint"
                    % (len(specs), ",".join(str(d[0]) for d, _ in specs)),
        "voxels": voxels, "vertices": verts,
        "forward_ms": round(f_ms, 4), "adjoint_ms": round(a_ms, 4),
        "fwd_plus_adjoint_over_fwd": round((f_ms + a_ms) / f_ms, 3), "target_ratio": 2.5,
        "forward_Gvoxels_per_s": round(voxels / (f_ms * 1e-3) / 1e9, 2),
        "fwd_adjoint_Gvoxels_per_s": round(voxels / ((f_ms + a_ms) * 1e-3) / 1e9, 2),
        "grad_check": {"voxels": int(pick.size), "volume_edge": dims[0], "h": hstep,
                       "max_rel_err_where_grad_ge_1e-3_max": float("%.3g" % raw),
                       "max_abs_err_over_max_grad": float("%.3g" % (absmax / gmax)),
                       "max_rel_err_above_floor": float("%.3g" % max(worst, 0.0)),
                       "floor": "1e-6 of the largest |dL/dvoxel| (entries whose contributions cancel)",
                       "tolerance": 1e-4},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-planes", type=int, default=192,
                    help="z-planes of the headline volume the CPU baseline processes")
    ap.add_argument("--skip-small", action="store_true",
                    help="leave out the 256^3 blob and 512^3 vessel workloads")
    ap.add_argument("--skip-plain", action="store_true",
                    help="profiling aid: leave out the headline workload (value is then null)")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-ingest", action="store_true")
    ap.add_argument("--batched-volumes", type=int, default=48)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--batch-block", type=int, default=0, help="0: autotuned")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--adjoint-volumes", type=int, default=16)
    ap.add_argument("--adjoint-steps", type=int, default=5)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the extraction path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt
    from meshtrex_amd.managers import Context, DeviceBuffer

    stream = hiprt.stream_create()
    if args.copy_calibration:
        nbytes = 1 << 30
        a, b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        for _ in range(3):
            assert L.synth().mtx_bench_copy(b.ptr, a.ptr, nbytes, stream) == 0
        hiprt.stream_sync(stream)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))
        return

    ctx = Context(0)
    if args.skip_plain:
        line = {"metric": METRIC, "value": None, "unit": "Gvoxels/s", "n_gpus": 1, "extra": {}}
    else:
        line = headline(args, ctx, stream)
    if not args.skip_batched:
        log("batched")
        line["extra"]["batched"] = bench_batched(args, ctx, stream)
        log("batched: %s" % json.dumps({k: line["extra"]["batched"][k] for k in
                                        ("speedup_vs_sequential_preallocated", "speedup_vs_sequential_api")}))
    if not args.skip_adjoint:
        log("forward + adjoint")
        line["extra"]["fwd_adjoint"] = bench_adjoint(args, ctx)
        log("fwd+adjoint: ratio %.3f" % line["extra"]["fwd_adjoint"]["fwd_plus_adjoint_over_fwd"])
    print(json.dumps(line), flush=True)
    ctx.close()


def headline(args, ctx, stream):
    import numpy as np
    from meshtrex_amd.managers import read_device
    log("plain %d^3 f32: %d warmup + %d timed steps" % (args.size, args.warmup, args.steps))
    p = bench_plain(args, ctx, stream)
    ms = p["ms"]
    line = {
        "metric": METRIC,
        "value": round(p["nvox"] / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-generated volumes, seeds 0x6D747278 + k)",
        "config": p["config"],
        "roofline": p["roofline"],
        "cpu_baseline": None,
        "extra": p["extra"],
    }
    log("plain: %.4f ms/step (median %.4f), %.1f Gvox/s" % (ms, p["extra"]["step_ms"]["median_ms"], line["value"]))
    vol, path = p.pop("_keep")
    path.close()
    T = cpu_threads()
    S = args.size
    if not args.no_cpu_baseline:
        P = min(args.cpu_planes, S)
        log("cpu baseline: %d z-planes of the %d^3 volume, 1 and %d threads" % (P, S, T))
        host = read_device(vol.ptr, np.float32, S * S * P)
        c = cpu_both(host, (S, S, P), 0.0, T)
        del host
        line["cpu_baseline"] = {
            "value": c["T"]["Gvoxels_per_s"], "unit": "Gvoxels/s", "cores": c["T"]["threads"], "kind": "port",
            "sample": "z-slab %dx%dx%d (%d voxels) of the same %d^3 f32 volume, iso 0: oracle min-max + filter "
                      "+ extraction with normals, split over %d threads by 8-plane block rows, in %.2f s "
                      "(%d triangles)" % (S, S, P, S * S * P, S, c["T"]["threads"], c["T"]["total_s"],
                                           c["T"]["triangles"]),
            "one_thread": c["1"], "threads_T": c["T"], "cpu_model": cpu_model()}
    line["extra"]["cpu"] = {"T": T, "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
                            "T_rule": "the CPUs this process may use (sched_getaffinity), at most 16"}
    if not args.skip_small:
        flush = make_flush(stream)
        small = {}
        o, buf0 = bench_small(args, ctx, stream, "u8", (256, 256, 256), SEED - 2, 128.0, flush, "blobs 256^3 u8",
                              "256^3 uint8, 16 Gaussian blobs + integer noise [-2,2], isovalue 128, 8^3 blocks; "
                              "min-max + pyramid + filter + extraction with normals and meshlets")
        small["blobs256_u8"] = o
        o, buf1 = bench_small(args, ctx, stream, "u8vessels", (512, 512, 512), SEED - 1, 80.0, flush,
                              "vessels 512^3 u8",
                              "512^3 uint8, 40 polylines of 8 segments (length 20-80, radius 2-6) over a "
                              "0-20 background, isovalue 80, 8^3 blocks; same stages")
        small["vessels512_u8"] = o
        small["reference_minmax_blobs256"] = bench_reference_minmax(ctx, stream, buf0, (256, 256, 256), flush)
        if not args.no_cpu_baseline:
            log("cpu: blobs 256^3 and vessels 512^3, 1 and %d threads" % T)
            h0 = read_device(buf0.ptr, np.uint8, 256 ** 3)
            small["blobs256_u8"]["cpu"] = cpu_both(h0, (256, 256, 256), 128.0, T)
            del h0
            h1 = read_device(buf1.ptr, np.uint8, 512 ** 3)
            small["vessels512_u8"]["cpu"] = cpu_both(h1, (512, 512, 512), 80.0, T)
            del h1
            for k in ("blobs256_u8", "vessels512_u8"):
                g = small[k]["Gvoxels_per_s"]
                small[k]["speedup_vs_cpu_1_thread"] = round(g / small[k]["cpu"]["1"]["Gvoxels_per_s"], 1)
                small[k]["speedup_vs_cpu_T_threads"] = round(g / small[k]["cpu"]["T"]["Gvoxels_per_s"], 1)
        buf0.free()
        buf1.free()
        line["extra"]["small_volumes"] = small
    if not args.skip_ingest:
        log("ingest from host memory")
        line["extra"]["ingest"] = bench_ingest(args, ctx, vol.ptr, p["dims"], p["extra"]["stage_ms"]["minmax_leaf"])
    vol.free()
    return line


if __name__ == "__main__":
    main()
