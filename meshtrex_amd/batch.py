"""Batched extraction over the C ABI (mtx_batch_*).

    b = Batch(ctx, [(vol0, (iso_a, iso_b)), (vol1, (iso_c,))])   # device volumes
    b.run()                    # replays one captured hipGraph: no allocation, no host sync
    r = b.read_ranges()        # per (volume, isovalue) unit: its slice of the outputs
    verts, idx = b.unit_mesh(0)

A volume is a Volume whose data is device-resident (a DeviceBuffer or a raw
device address).  Units are numbered volume by volume, isovalues in order.
Each unit's slice equals what the single-volume path (extractMeshletDescriptors)
gives for that volume and isovalue at the unit's block edge, with indices
relative to the unit's first vertex.  No reference counterpart: the reference
extracts one volume at one isovalue per call (src/meshtrex.cpp:107-132).
"""
import ctypes

import numpy as np

from . import _lib as L
from .managers import _check, Volume, read_device

RANGE_FIELDS = [f for f, _ in L.BatchRange._fields_]


class Batch:
    def __init__(self, context, volumes, block_dim=0, minmax_rows=0, rule=L.MTX_RULE_CPU, normals=True,
                 edge_ids=False, graph=True, tune_iters=3, headroom_pct=0, schedule=0, wg_per_cu=0):
        self.ctx = context
        self.volumes = list(volumes)
        vols = (L.CBatchVolume * len(self.volumes))()
        self._isos = []
        for i, (vol, isos) in enumerate(self.volumes):
            if not isinstance(vol, Volume):
                raise TypeError("batch volumes are Volume objects")
            cv = vol._c()
            if not cv.on_device:
                raise ValueError("batch volume %d is not device-resident" % i)
            a = (ctypes.c_float * len(isos))(*[float(x) for x in isos])
            self._isos.append(a)
            vols[i].volume = cv
            vols[i].isovalues = ctypes.cast(a, ctypes.POINTER(ctypes.c_float))
            vols[i].num_isovalues = len(isos)
        o = L.BatchOptions()
        L.lib().mtx_batch_options_default(ctypes.byref(o))
        o.block_dim, o.minmax_rows, o.rule = block_dim, minmax_rows, rule
        o.normals, o.edge_ids, o.graph = int(bool(normals)), int(bool(edge_ids)), int(bool(graph))
        o.tune_iters, o.headroom_pct = tune_iters, headroom_pct
        o.schedule, o.wg_per_cu = schedule, wg_per_cu
        h = ctypes.c_void_p()
        _check(context, L.lib().mtx_batch_create(context.handle, vols, len(self.volumes), ctypes.byref(o),
                                                 ctypes.byref(h)))
        self.handle = h
        self.out = L.CBatchOutput()
        _check(context, L.lib().mtx_batch_output_get(self.handle, ctypes.byref(self.out)))
        self.num_units = self.out.num_units
        self.edge_ids = bool(edge_ids)

    def run(self, stream=None):
        _check(self.ctx, L.lib().mtx_batch_run(self.handle, None if stream is None else int(stream)))

    def read_ranges(self):
        """List of dicts (mtx_batch_range) of the last run, one per unit."""
        n = self.num_units
        buf = (L.BatchRange * n)()
        _check(self.ctx, L.lib().mtx_batch_read_ranges(self.handle, buf, n))
        return [{f: getattr(r, f) for f in RANGE_FIELDS} for r in buf]

    def class_info(self):
        out = []
        for k in range(self.out.num_classes):
            ci = L.BatchClassInfo()
            _check(self.ctx, L.lib().mtx_batch_class_info_get(self.handle, k, ctypes.byref(ci)))
            d = {f: getattr(ci, f) for f, _ in L.BatchClassInfo._fields_}
            d["tuned_ms"] = list(ci.tuned_ms)
            d["tuned_rows_ms"] = list(ci.tuned_rows_ms)
            d["tuned_sched_ms"] = list(ci.tuned_sched_ms)
            d["dtype"] = "float32" if ci.dtype == L.MTX_F32 else "uint8"
            out.append(d)
        return out

    def unit_mesh(self, unit, ranges=None):
        """(vertices float32 [n, 8], indices uint32 [m]) of one unit, read to host."""
        r = (ranges or self.read_ranges())[unit]
        v = read_device(self.out.vertices + 32 * r["vertex_offset"], np.float32, 8 * r["vertex_count"])
        i = read_device(self.out.indices + 4 * r["index_offset"], np.uint32, r["index_count"])
        return v.reshape(-1, 8), i

    def unit_meshlets(self, unit, ranges=None):
        r = (ranges or self.read_ranges())[unit]
        m = read_device(self.out.meshlets + 16 * r["meshlet_offset"], np.uint32, 4 * r["meshlet_count"])
        return m.reshape(-1, 4)

    def unit_edge_ids(self, unit, ranges=None):
        if not self.edge_ids:
            raise ValueError("batch made without edge_ids")
        r = (ranges or self.read_ranges())[unit]
        return read_device(self.out.vertex_edge_ids + 8 * r["vertex_offset"], np.uint64, r["vertex_count"])

    def unit_active_ids(self, unit, ranges=None):
        r = (ranges or self.read_ranges())[unit]
        return read_device(self.out.active_block_ids + 4 * r["active_offset"], np.uint32, r["active_count"])

    def close(self):
        if getattr(self, "handle", None):
            L.lib().mtx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


__all__ = ["Batch", "RANGE_FIELDS"]
