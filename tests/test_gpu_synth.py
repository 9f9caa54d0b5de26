"""The GPU volume generator (libmtx_synth) against the oracle's generators.

Both evaluate the same recipes (csrc/synth_recipes.h) in FP64; the GPU's exp
may differ from the host's in the last bits, so quantised uint8 volumes may
differ by 1 where a value sits on a rounding boundary and float32 volumes by
an ulp.  Dimensions include sizes where the launch has more threads than
voxels and 64-bit seeds.
"""
import numpy as np
import pytest

import oracle
from meshtrex_amd import DeviceBuffer, read_device
from meshtrex_amd import workloads

pytestmark = pytest.mark.gpu

SEEDS = [oracle.SEED_BASE, oracle.SEED_BASE + 3 + (5 << 32)]
DIMS = [(37, 45, 29), (338, 63, 17), (5, 7, 3), (64, 64, 64)]


def gpu(kind, dims, seed):
    buf, _ = workloads.generate(dims, kind, seed)
    n = int(np.prod(dims))
    out = read_device(buf.ptr, np.uint8 if kind.startswith("u8") else np.float32, n)
    buf.free()
    return out


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("seed", SEEDS)
def test_blobs_u8_match_oracle(dims, seed):
    g = gpu("u8", dims, seed)
    c = oracle.gen_blobs_u8(dims, seed, 16)
    d = np.abs(g.astype(np.int32) - c.astype(np.int32))
    assert d.max() <= 1 and np.mean(d != 0) <= 1e-4


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("seed", SEEDS)
def test_blobs_f32_and_sdf_match_oracle(dims, seed):
    g = gpu("f32blobs", dims, seed)
    c = oracle.gen_blobs_f32(dims, seed, 16)
    assert np.max(np.abs(g - c)) <= 1e-6
    g = gpu("f32", dims, seed)
    c = oracle.gen_sdf_f32(dims, seed, 64, 32, min(dims) / 1024.0, 0.05)
    assert np.max(np.abs(g - c) / np.maximum(np.abs(c), 1.0)) <= 1e-6


@pytest.mark.parametrize("dims", DIMS)
def test_vessels_u8_match_oracle(dims):
    seed = oracle.SEED_BASE + 1
    g = gpu("u8vessels", dims, seed)
    c = oracle.gen_vessels_u8(dims, seed, 40, 8)
    d = np.abs(g.astype(np.int32) - c.astype(np.int32))
    assert d.max() <= 1 and np.mean(d != 0) <= 1e-4
