"""GPU adjoint (mtx_adjoint) and the torch.autograd binding.

Checked against (a) the FP64 CPU adjoint of the oracle for the same upstream
gradient, (b) FP64 central finite differences of the loss with the topology
frozen, h = 1e-3, on 256 random voxels that are endpoints of active edges with
|v - iso| > 10h.  The difference quotient is Richardson-extrapolated from
steps h and h/2 (its truncation error is O((h/dv)^4) instead of O((h/dv)^2),
dv the edge's value difference; on gently varying fields the plain quotient's
own error would exceed the 1e-4 tolerance).  Tolerance: 1e-4 relative.
"""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def blob_volume(dims, seed):
    return oracle.gen_blobs_f32(dims, seed, 8)


def sphere_loss_np(p, c, R):
    return float(np.sum((np.linalg.norm(p - c, axis=1) - R) ** 2))


@pytest.mark.parametrize("dims,seed", [((48, 40, 44), 1), ((33, 65, 20), 2)])
def test_adjoint_matches_fp64_oracle(dims, seed):
    from meshtrex_amd.autograd import Extraction
    v = blob_volume(dims, seed)
    vol = torch.from_numpy(v.reshape(dims[::-1])).cuda()
    ext = Extraction(vol, 0.0, 8, edge_ids=True)
    e = ext.edge_ids().cpu().numpy().astype(np.uint64)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((e.size, 3)).astype(np.float32)
    dvol, diso = ext.adjoint(torch.from_numpy(w).cuda())
    torch.cuda.synchronize()
    ref_dv, ref_di = oracle.adjoint(v, dims, 0.0, e, w.astype(np.float64))
    got = dvol.cpu().numpy().reshape(-1).astype(np.float64)
    scale = np.abs(ref_dv).max()
    assert scale > 0
    assert np.all(np.abs(got - ref_dv) <= 1e-4 * np.abs(ref_dv) + 1e-6 * scale)
    assert abs(float(diso) - ref_di) <= 1e-4 * abs(ref_di) + 1e-6
    # voxels no vertex depends on get exactly zero
    touched = np.zeros(v.size, bool)
    vox = (e // 3).astype(np.int64)
    st = np.array([1, dims[0], dims[0] * dims[1]])[(e % 3).astype(np.int64)]
    touched[vox] = True
    touched[vox + st] = True
    assert np.all(got[~touched] == 0.0)


def test_adjoint_is_deterministic():
    from meshtrex_amd.autograd import Extraction
    dims = (40, 40, 40)
    vol = torch.from_numpy(blob_volume(dims, 3).reshape(dims[::-1])).cuda()
    ext = Extraction(vol, 0.0, 8)
    g = torch.randn(ext.positions.shape, device="cuda")
    a1, i1 = ext.adjoint(g)
    a2, i2 = ext.adjoint(g)
    assert torch.equal(a1, a2) and torch.equal(i1, i2)


def test_autograd_sphere_loss_vs_central_differences():
    from meshtrex_amd.autograd import Extraction, marching_cubes
    dims = (64, 64, 64)
    v = blob_volume(dims, 4)
    iso = 0.0
    vol = torch.from_numpy(v.reshape(dims[::-1])).cuda().requires_grad_(True)
    iso_t = torch.tensor(iso, device="cuda", requires_grad=True)
    c = torch.tensor([(d - 1) / 2.0 for d in dims], device="cuda", dtype=torch.float64)
    R = 0.3 * min(dims)
    pos, faces = marching_cubes(vol, iso_t)
    # the loss in FP64: |p - c| - R cancels near the target sphere
    loss = ((pos.double() - c).norm(dim=1) - R).pow(2).sum()
    loss.backward()
    g_vol = vol.grad.cpu().numpy().reshape(-1).astype(np.float64)
    g_iso = float(iso_t.grad)

    # frozen topology: the forward pass's owning edges
    e = Extraction(vol.detach(), iso, 8, edge_ids=True).edge_ids().cpu().numpy().astype(np.uint64)
    cc = np.array([(d - 1) / 2.0 for d in dims])

    # the FP64 loss of FP64 voxel values, positions restated from the owning edges
    ax = (e % 3).astype(np.int64)
    vox = (e // 3).astype(np.int64)
    st = np.array([1, dims[0], dims[0] * dims[1]])[ax]
    owner = np.stack([vox % dims[0], (vox // dims[0]) % dims[1], vox // (dims[0] * dims[1])], 1).astype(np.float64)

    def L64(vals, iv=iso):
        t = (iv - vals[vox]) / (vals[vox + st] - vals[vox])
        p = owner.copy()
        p[np.arange(e.size), ax] += t
        return sphere_loss_np(p, cc, R)

    base = v.astype(np.float64)
    h = 1e-3
    cand = np.unique(np.concatenate([vox, vox + st]))
    cand = cand[np.abs(base[cand] - iso) > 10 * h]
    rng = np.random.default_rng(7)
    pick = rng.choice(cand, min(256, cand.size), replace=False)

    def fd(fun, h):
        return (fun(+h) - fun(-h)) / (2 * h)

    # relative 1e-4, with a floor of 1e-6 of the largest gradient for entries
    # whose contributions cancel to ~0
    floor = 1e-6 * np.abs(g_vol).max()
    worst = 0.0
    for i in pick.tolist():
        def f(dh, i=i):
            vv = base.copy()
            vv[i] += dh
            return L64(vv)
        d = (4 * fd(f, h / 2) - fd(f, h)) / 3   # Richardson extrapolation of central differences
        worst = max(worst, (abs(d - g_vol[i]) - floor) / max(abs(d), 1e-300))
    assert worst <= 1e-4, worst
    d_iso = (4 * fd(lambda dh: L64(base, iso + dh), h / 2) - fd(lambda dh: L64(base, iso + dh), h)) / 3
    assert abs(d_iso - g_iso) <= 1e-4 * abs(d_iso)

    # and against the FP64 CPU gradient of the same loss
    p64 = owner.copy()
    p64[np.arange(e.size), ax] += (iso - base[vox]) / (base[vox + st] - base[vox])
    r = np.linalg.norm(p64 - cc, axis=1)
    dLdp = (2 * (r - R) / r)[:, None] * (p64 - cc)
    ref_dv, ref_di = oracle.adjoint(v, dims, iso, e, dLdp)
    scale = np.abs(ref_dv).max()
    assert np.all(np.abs(g_vol - ref_dv) <= 1e-4 * np.abs(ref_dv) + 1e-5 * scale)
    assert abs(g_iso - ref_di) <= 1e-4 * abs(ref_di)


def test_autograd_through_uint8_volume_has_no_volume_grad():
    from meshtrex_amd.autograd import marching_cubes
    dims = (24, 24, 24)
    v = oracle.gen_blobs_u8(dims, 5, 4).reshape(dims[::-1])
    vol = torch.from_numpy(v).cuda()
    iso = torch.tensor(100.5, device="cuda", requires_grad=True)
    pos, faces = marching_cubes(vol, iso)
    pos.sum().backward()
    assert iso.grad is not None and torch.isfinite(iso.grad)
