"""Benchmark of the isosurface-extraction path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--skip-batched] [--skip-adjoint]

Headline (`value`): a step is one pass of the path over a resident 1024^3
float32 SDF (64 spheres + 32 tori, N(0, 0.05^2) noise, isovalue 0, 8^3-cell
blocks): block min/max with the fused activity mask, the min/max pyramid,
ordered active-block compaction and PMB marching cubes (count, scan, emit of
vertices with normals, index buffer, meshlet descriptors), all on one stream
with no host synchronisation, into buffers sized by a warm-up pass.  `value`
is Gvoxels/s of the whole step.  `roofline` describes the block min/max
kernel (the dominant, HBM-bound kernel) from HIP events around its launch on
its own stream; `cpu_baseline` times the CPU oracle (scalar C, 1 thread) on a
z-slab of the same volume on this host.

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


def cpu_baseline(vol_ptr, dims, iso, planes):
    """The CPU oracle (scalar C, 1 thread) on the first `planes` z-planes."""
    import numpy as np
    import oracle
    from meshtrex_amd.managers import read_device
    X, Y = dims[0], dims[1]
    n = X * Y * planes
    host = read_device(vol_ptr, np.float32, n)
    sd = (X, Y, planes)
    t0 = time.perf_counter()
    pairs = oracle.minmax(host, sd, 8, 1)
    ids = oracle.filter_ids(pairs, oracle.F32, iso)
    ref = oracle.extract(host, sd, 8, iso, normals=True)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e9, 5), "unit": "Gvoxels/s", "cores": 1, "kind": "port",
            "sample": "z-slab %dx%dx%d (%d voxels) of the same 1024^3 f32 volume, iso %g: oracle "
                      "min-max + filter + full extraction with normals (%d active blocks, %d "
                      "triangles) in %.2f s" % (X, Y, planes, n, iso, ids.size, len(ref["tris"]), dt),
            "triangles_per_s": round(len(ref["tris"]) / dt, 1)}


def bench_plain(args, ctx, stream):
    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt
    from meshtrex_amd.managers import (DeviceBuffer, ExtractionOutput, FilteringOutput, MinMaxOutput, _check,
                                       makeOptions, makePushConstants)
    S = args.size
    dims = (S, S, S)
    nvox = S ** 3
    vol = DeviceBuffer(4 * nvox)
    d3 = (L.u32 * 3)(*dims)
    t0 = time.perf_counter()
    rc = L.synth().mtx_synth_sdf_f32(vol.ptr, d3, SEED, 64, 32, S / 1024.0, 0.05, stream)
    assert rc == 0, rc
    gen_s = time.perf_counter() - t0
    iso = 0.0
    pc = makePushConstants(dims, 8, iso)
    cvol = L.CVolume()
    cvol.data, cvol.dims[:], cvol.dtype, cvol.on_device = vol.ptr, dims, L.MTX_F32, 1
    o_mm = makeOptions(build_pyramid=0, reuse=1)
    o_ex = makeOptions(reuse=1)
    mm, filt, ex = MinMaxOutput(ctx), FilteringOutput(ctx), ExtractionOutput(ctx)
    lib = L.lib()
    h = ctx.handle

    # sizing pass: allocating extraction (one host sync inside)
    _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(cvol), ctypes.byref(pc), ctypes.byref(o_mm),
                                        stream, ctypes.byref(mm.c)))
    _check(ctx, lib.mtx_build_pyramid(h, ctypes.byref(mm.c), stream))
    _check(ctx, lib.mtx_filter_active_blocks(h, ctypes.byref(mm.c), ctypes.byref(pc), None, stream,
                                             ctypes.byref(filt.c)))
    _check(ctx, lib.mtx_extract(h, ctypes.byref(mm.c), ctypes.byref(filt.c), ctypes.byref(pc),
                                ctypes.byref(makeOptions()), stream, ctypes.byref(ex.c)))
    hiprt.stream_sync(stream)

    def step(events=None):
        if events:
            events[0].record(stream)
        _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(cvol), ctypes.byref(pc), ctypes.byref(o_mm),
                                            stream, ctypes.byref(mm.c)))
        if events:
            events[1].record(stream)
        _check(ctx, lib.mtx_build_pyramid(h, ctypes.byref(mm.c), stream))
        if events:
            events[2].record(stream)
        _check(ctx, lib.mtx_filter_active_blocks(h, ctypes.byref(mm.c), ctypes.byref(pc), None, stream,
                                                 ctypes.byref(filt.c)))
        if events:
            events[3].record(stream)
        _check(ctx, lib.mtx_extract(h, ctypes.byref(mm.c), ctypes.byref(filt.c), ctypes.byref(pc),
                                    ctypes.byref(o_ex), stream, ctypes.byref(ex.c)))
        if events:
            events[4].record(stream)

    for _ in range(args.warmup):
        step()
    hiprt.stream_sync(stream)

    per_step = [[hiprt.Event() for _ in range(5)] for _ in range(args.steps)]
    start, stop = hiprt.Event(), hiprt.Event()
    wall0 = time.perf_counter()
    start.record(stream)
    for k in range(args.steps):
        step(per_step[k])
    stop.record(stream)
    stop.sync()
    wall = time.perf_counter() - wall0
    total_ms = stop.ms_since(start)
    ms = total_ms / args.steps
    stages = {"minmax_leaf": 0.0, "pyramid": 0.0, "filter": 0.0, "extract": 0.0}
    for e in per_step:
        for i, name in enumerate(stages):
            stages[name] += e[i + 1].ms_since(e[i]) / args.steps

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

    nv, ni, nm = ex.read_counts()
    nmv = ex.meshletVertexCount
    na = int(filt.read_ids().size)
    nb = ((S + 7) // 8) ** 3
    ntri = ni // 3
    leaf_bytes = 4 * nvox + 8 * nb
    mmf_bytes = 4 * nvox + 8 * nb + 4 * na + 4
    ext_bytes = na * 8 ** 3 * 4 + 32 * nv + 12 * ntri + 16 * nm
    # everything the extraction writes: the shared mesh, the meshlet descriptors
    # and the renderer layout (meshlet vertices, their map, meshlet indices)
    ext_out_bytes = 32 * nv + 12 * ntri + 16 * nm + (32 + 4) * nmv + 12 * ntri
    leaf_gbs = leaf_bytes / (stages["minmax_leaf"] * 1e-3) / 1e9
    mmf_gbs = mmf_bytes / ((stages["minmax_leaf"] + stages["filter"]) * 1e-3) / 1e9
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
                     "algorithmic_bytes_per_launch": leaf_bytes,
                     "achievable_read_GBps": round(read_gbs, 1),
                     "frac_of_achievable_read": round(leaf_gbs / read_gbs, 4),
                     "achievable_copy_GBps": round(copy_gbs, 1)},
        "extra": {
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "minmax_plus_filter": {"bytes": mmf_bytes, "ms": round(stages["minmax_leaf"] + stages["filter"], 4),
                                   "GBps": round(mmf_gbs, 1), "frac_of_peak": round(mmf_gbs / HBM_PEAK_GBS, 4),
                                   "target_frac": 0.70},
            "extraction": {"algorithmic_bytes": ext_bytes, "ms": round(stages["extract"], 4),
                           "GBps": round(ext_bytes / (stages["extract"] * 1e-3) / 1e9, 1),
                           "bytes_in_plus_all_outputs": na * 8 ** 3 * 4 + ext_out_bytes,
                           "GBps_all_outputs": round((na * 8 ** 3 * 4 + ext_out_bytes) /
                                                     (stages["extract"] * 1e-3) / 1e9, 1)},
            "triangles_per_s": round(ntri / (ms * 1e-3), 1),
            "wall_s_timed": round(wall, 4),
            "generate_s": round(gen_s, 3),
        },
        "_keep": (vol, mm, filt, ex),
    }
    return res


def bench_ingest(args, ctx, vol_ptr, dims, minmax_ms):
    """Host volume -> min/max: slabs uploaded on a copy stream, each slab's
    min/max run as soon as it is resident; pageable and pinned host memory,
    beside a bare pinned copy and the device-resident min/max."""
    import numpy as np
    from meshtrex_amd import _lib as L
    from meshtrex_amd.managers import (MinMaxOutput, Volume, _check, makeOptions, makePushConstants,
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
        for _ in range(3):
            t0 = time.perf_counter()
            _check(ctx, lib.mtx_compute_min_max(h, ctypes.byref(v), ctypes.byref(pc), ctypes.byref(o), None,
                                                ctypes.byref(mm.c)))
            times.append(time.perf_counter() - t0)
        mm.cleanup()
        best = min(times[1:])   # the first call also allocates the device copy
        res[name] = {"ms": round(best * 1e3, 2), "GBps": round(nbytes / best / 1e9, 2)}
    from meshtrex_amd.managers import DeviceBuffer
    d = DeviceBuffer(nbytes)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        assert lib.mtx_memcpy_h2d(d.ptr, pin.ctypes.data, nbytes) == 0
        times.append(time.perf_counter() - t0)
    d.free()
    copy_s = min(times)
    res["pinned_copy_only"] = {"ms": round(copy_s * 1e3, 2), "GBps": round(nbytes / copy_s / 1e9, 2)}
    res["device_resident_minmax_ms"] = round(minmax_ms, 3)
    serial = copy_s * 1e3 + minmax_ms
    res["pinned_overlap"] = {"serial_copy_then_minmax_ms": round(serial, 2),
                             "slabbed_ms": res["pinned"]["ms"],
                             "hidden_fraction_of_minmax": round(max(0.0, min(1.0, (serial - res["pinned"]["ms"])
                                                                             / minmax_ms)), 3)}
    return res


def bench_batched(args, ctx, stream):
    import numpy as np
    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt, workloads
    from meshtrex_amd.batch import Batch
    from meshtrex_amd.managers import (ExtractionOutput, FilteringOutput, MinMaxOutput, _check,
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
    e0.record(stream)
    for _ in range(args.batched_steps):
        b.run(stream)
    e1.record(stream)
    e1.sync()
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
    ap.add_argument("--cpu-planes", type=int, default=320,
                    help="z-planes of the volume the CPU baseline processes")
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
    log("plain: %.4f ms/step, %.1f Gvox/s" % (ms, line["value"]))
    vol, mm, filt, ex = p.pop("_keep")
    ex.cleanup()
    filt.cleanup()
    mm.cleanup()
    if not args.no_cpu_baseline:
        log("cpu baseline: %d z-planes" % args.cpu_planes)
        line["cpu_baseline"] = cpu_baseline(vol.ptr, p["dims"], 0.0, min(args.cpu_planes, args.size))
    if not args.skip_ingest:
        log("ingest from host memory")
        line["extra"]["ingest"] = bench_ingest(args, ctx, vol.ptr, p["dims"], p["extra"]["stage_ms"]["minmax_leaf"])
    vol.free()
    return line


if __name__ == "__main__":
    main()
