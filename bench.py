"""Benchmark: isosurface extraction of a 1024^3 float32 SDF on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is one pass of the path over the resident volume: block min/max (with
the fused activity mask), the min/max pyramid, ordered active-block
compaction and PMB marching cubes (count, scan, emit of vertices with normals
and the index buffer), all on one stream with no host synchronisation, into
buffers sized by a warm-up pass.  The volume is generated on the GPU
(64 spheres + 32 tori, radii as the workload defines, N(0, 0.05^2) noise),
isovalue 0, 8^3-cell blocks.

Prints one JSON line.  `value` is Gvoxels/s of the whole step; `roofline`
describes the block min/max kernel (the dominant kernel, HBM-bound) from HIP
events around its launch on the same stream; `cpu_baseline` times the CPU
oracle on a z-slab of the same volume on this host (1 thread).
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


def parse_traffic():
    """HBM bytes per launch of the min/max kernel, from the committed PMC summary."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def cpu_baseline(vol_ptr, dims, iso, planes=40):
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
                      "triangles) in %.2f s" % (X, Y, planes, n, iso, ids.size, len(ref["tris"]), dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the extraction path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    from meshtrex_amd import _lib as L
    from meshtrex_amd import hiprt
    from meshtrex_amd.managers import (Context, DeviceBuffer, ExtractionOutput, FilteringOutput,
                                       MinMaxOutput, _check, makeOptions, makePushConstants)

    stream = hiprt.stream_create()
    if args.copy_calibration:
        nbytes = 1 << 30
        a, b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        for _ in range(3):
            assert L.synth().mtx_bench_copy(b.ptr, a.ptr, nbytes, stream) == 0
        hiprt.stream_sync(stream)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))
        return

    S = args.size
    dims = (S, S, S)
    nvox = S ** 3
    ctx = Context(0)
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

    ev = [hiprt.Event() for _ in range(5)]

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
    hiprt.stream_sync(stream)
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

    nv, ni, _ = ex.read_counts()
    na = int(filt.read_ids().size)
    nb = (S // 8) ** 3 if S % 8 == 0 else ((S + 7) // 8) ** 3
    ntri = ni // 3
    leaf_bytes = 4 * nvox + 8 * nb
    mmf_bytes = 4 * nvox + 8 * nb + 4 * na + 4
    ext_bytes = na * 8 ** 3 * 4 + 32 * nv + 12 * ntri
    leaf_gbs = leaf_bytes / (stages["minmax_leaf"] * 1e-3) / 1e9
    mmf_gbs = mmf_bytes / ((stages["minmax_leaf"] + stages["filter"]) * 1e-3) / 1e9
    traffic = parse_traffic()

    line = {
        "metric": METRIC,
        "value": round(nvox / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-generated SDF, seed 0x%x)" % SEED,
        "config": {"workload": "plain single extraction, %d^3 float32 SDF (64 spheres + 32 tori + "
                               "N(0,0.05^2)), isovalue 0, 8^3-cell blocks; min-max + pyramid + "
                               "filter + PMB extraction with normals" % S,
                   "volume_bytes": 4 * nvox, "blocks": nb, "active_blocks": na,
                   "vertices": nv, "triangles": ntri},
        "roofline": {"bound": "hbm", "achieved": round(leaf_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(leaf_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_minmax_b8<float> (block min/max + fused activity mask)",
                     "algorithmic_bytes_per_launch": leaf_bytes},
        "cpu_baseline": None,
        "extra": {
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "minmax_plus_filter": {"bytes": mmf_bytes, "ms": round(stages["minmax_leaf"] + stages["filter"], 4),
                                   "GBps": round(mmf_gbs, 1), "frac_of_peak": round(mmf_gbs / HBM_PEAK_GBS, 4),
                                   "target_frac": 0.70},
            "extraction": {"algorithmic_bytes": ext_bytes, "ms": round(stages["extract"], 4),
                           "GBps": round(ext_bytes / (stages["extract"] * 1e-3) / 1e9, 1)},
            "triangles_per_s": round(ntri / (ms * 1e-3), 1),
            "wall_s_timed": round(wall, 4),
            "generate_s": round(gen_s, 3),
        },
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(vol.ptr, dims, iso)
    print(json.dumps(line))
    ex.cleanup()
    filt.cleanup()
    mm.cleanup()
    vol.free()
    ctx.close()


if __name__ == "__main__":
    main()
