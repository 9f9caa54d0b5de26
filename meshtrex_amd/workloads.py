"""Synthetic workloads, generated on the GPU by libmtx_synth.

    batched_specs(n)        many volumes of uneven dimensions with 1-8 isovalues each
    generate(spec)          one volume in device memory (DeviceBuffer)
    differentiable_specs(n) cubic float32 blob fields for the adjoint

Volume recipes (shared with the CPU oracle through csrc/synth_recipes.h):
uint8 blobs -- 16 Gaussian blobs, quantised to 0..255 plus integer noise in
[-2, 2]; uint8 vessels -- 40 polylines of 8 segments (length 20-80, radius
2-6), 255 exp(-d^2 / 2r^2) over a uniform 0..20 background; float32 SDF -- min over 64 spheres and 32 tori with radii scaled by
min(dim)/1024, plus N(0, 0.05^2); float32 blobs -- the blob field rescaled to
[-1, 1].  Seeds are 0x6D747278 + k (k = 3 for the batched set, 4 for the
differentiable set) mixed with the volume's index.
"""
import numpy as np

from . import _lib as L
from .managers import DeviceBuffer, Volume, read_device

SEED_BASE = 0x6D747278


def _dim(rng, lo, hi):
    # uniform in [lo, hi], redrawn while a multiple of 4 (so never a multiple of 4, 8 or 16)
    d = int(rng.integers(lo, hi + 1))
    while d % 4 == 0:
        d = int(rng.integers(lo, hi + 1))
    return d


def batched_specs(n=48, lo=64, hi=384):
    """[(dims, kind, seed, iso_fractions)]: even index uint8 blobs, odd float32 SDF;
    1-8 isovalues each as fractions of the volume's [min, max] range."""
    rng = np.random.default_rng(SEED_BASE + 3)
    out = []
    for j in range(n):
        dims = tuple(_dim(rng, lo, hi) for _ in range(3))
        niso = int(rng.integers(1, 9))
        fr = rng.random(niso)
        out.append((dims, "u8" if j % 2 == 0 else "f32", SEED_BASE + 3 + (j << 32), fr))
    return out


def differentiable_specs(n=16, lo=128, hi=256):
    """[(dims, seed)]: cubic float32 blob fields rescaled to [-1, 1], isovalue 0."""
    rng = np.random.default_rng(SEED_BASE + 4)
    out = []
    for j in range(n):
        e = int(rng.integers(lo, hi + 1))
        out.append(((e, e, e), SEED_BASE + 4 + (j << 32)))
    return out


def generate(dims, kind, seed, stream=None):
    """One volume on the device: (DeviceBuffer, Volume)."""
    n = int(np.prod(dims))
    d3 = (L.u32 * 3)(*dims)
    syn = L.synth()
    if kind == "u8":
        buf = DeviceBuffer(n)
        rc = syn.mtx_synth_blobs_u8(buf.ptr, d3, seed, 16, stream)
        vt = "uint8"
    elif kind == "f32":
        buf = DeviceBuffer(4 * n)
        rc = syn.mtx_synth_sdf_f32(buf.ptr, d3, seed, 64, 32, min(dims) / 1024.0, 0.05, stream)
        vt = "float32"
    elif kind == "u8vessels":
        buf = DeviceBuffer(n)
        rc = syn.mtx_synth_vessels_u8(buf.ptr, d3, seed, 40, 8, stream)
        vt = "uint8"
    elif kind == "f32blobs":
        buf = DeviceBuffer(4 * n)
        rc = syn.mtx_synth_blobs_f32(buf.ptr, d3, seed, 16, stream)
        vt = "float32"
    else:
        raise ValueError(kind)
    if rc != 0:
        raise RuntimeError("volume synthesis failed (%d)" % rc)
    return buf, Volume(dims, vt, buf)


def isovalues(buf, dims, kind, fractions):
    """Isovalues at the given fractions of the volume's [min, max] (read back once)."""
    host = read_device(buf.ptr, np.uint8 if kind == "u8" else np.float32, int(np.prod(dims)))
    lo, hi = float(host.min()), float(host.max())
    return [float(np.float32(lo + f * (hi - lo))) for f in fractions]
