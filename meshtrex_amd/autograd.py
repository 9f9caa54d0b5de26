"""Differentiable isosurface extraction for PyTorch-ROCm.

    pos, faces = marching_cubes(volume, isovalue)      # volume: float32 CUDA tensor [Z, Y, X]
    loss = f(pos); loss.backward()                     # -> volume.grad, isovalue.grad

The forward pass is the library's extraction (block min/max, filter, PMB
marching cubes) on torch's current stream; `pos` holds the vertex positions in
voxel units (x, y, z), one per active edge, `faces` the triangles (int64).
The backward pass is the hand-written adjoint kernel (mtx_adjoint): gradients
flow to the voxel values and the isovalue through t = (iso - v0)/(v1 - v0)
with the mesh topology frozen.  torch is the harness only: every kernel is the
library's.
"""
import ctypes

import torch

from . import _lib as L
from .managers import (Context, ExtractionOutput, FilteringOutput, MinMaxOutput, _check, makeOptions,
                       makePushConstants)

_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(torch.cuda.current_device())
    return _default_ctx


class Extraction:
    """One forward pass on a torch volume; keeps the device outputs the adjoint needs."""

    def __init__(self, volume, isovalue, block_dim=8, ctx=None, edge_ids=False, normals=False):
        if not (volume.is_cuda and volume.dim() == 3):
            raise ValueError("volume must be a 3-D CUDA tensor [Z, Y, X]")
        if volume.dtype not in (torch.float32, torch.uint8):
            raise ValueError("volume must be float32 or uint8")
        self.ctx = ctx or default_context()
        self.volume = volume.contiguous()
        Z, Y, X = self.volume.shape
        self.dims = (X, Y, Z)
        self.isovalue = float(isovalue)
        self.block_dim = block_dim
        self.stream = torch.cuda.current_stream().cuda_stream
        lib, h = L.lib(), self.ctx.handle
        self.pc = makePushConstants(self.dims, block_dim, self.isovalue)
        cvol = L.CVolume()
        cvol.data = self.volume.data_ptr()
        cvol.dims[:] = self.dims
        cvol.dtype = L.MTX_F32 if self.volume.dtype == torch.float32 else L.MTX_U8
        cvol.on_device = 1
        self.mm, self.filt, self.ex = MinMaxOutput(self.ctx), FilteringOutput(self.ctx), ExtractionOutput(self.ctx)
        s = self.stream
        _check(self.ctx, lib.mtx_compute_min_max(h, ctypes.byref(cvol), ctypes.byref(self.pc),
                                                 ctypes.byref(makeOptions(build_pyramid=0)), s,
                                                 ctypes.byref(self.mm.c)))
        _check(self.ctx, lib.mtx_filter_active_blocks(h, ctypes.byref(self.mm.c), ctypes.byref(self.pc), None, s,
                                                      ctypes.byref(self.filt.c)))
        opts = makeOptions(voxel_positions=1, edge_ids=int(edge_ids), normals=int(normals), meshlets=0)
        _check(self.ctx, lib.mtx_extract(h, ctypes.byref(self.mm.c), ctypes.byref(self.filt.c),
                                         ctypes.byref(self.pc), ctypes.byref(opts), s, ctypes.byref(self.ex.c)))
        nv, ni = self.ex.c.vertexCount, self.ex.c.indexCount
        dev = self.volume.device
        self.positions = torch.empty((nv, 3), dtype=torch.float32, device=dev)
        faces32 = torch.empty((ni // 3, 3), dtype=torch.int32, device=dev)
        if nv:
            _check(self.ctx, lib.mtx_memcpy_d2d_async(self.positions.data_ptr(), self.ex.c.vertex_voxel_pos,
                                                      12 * nv, s))
        if ni:
            _check(self.ctx, lib.mtx_memcpy_d2d_async(faces32.data_ptr(), self.ex.c.indices, 4 * ni, s))
        self.faces = faces32.long()

    def edge_ids(self):
        t = torch.empty(self.ex.c.vertexCount, dtype=torch.int64, device=self.volume.device)
        if t.numel():
            _check(self.ctx, L.lib().mtx_memcpy_d2d_async(t.data_ptr(), self.ex.c.vertex_edge_ids, 8 * t.numel(),
                                                          self.stream))
        return t

    def adjoint(self, grad_positions):
        """(dL/dvolume, dL/disovalue) for dL/dpositions (n, 3) in voxel units."""
        g = grad_positions.contiguous().to(torch.float32)
        stream = torch.cuda.current_stream().cuda_stream
        dvol = torch.empty(self.volume.shape, dtype=torch.float32, device=self.volume.device)
        diso = torch.empty((), dtype=torch.float32, device=self.volume.device)
        _check(self.ctx, L.lib().mtx_adjoint(self.ctx.handle, ctypes.byref(self.mm.c), ctypes.byref(self.filt.c),
                                             ctypes.byref(self.ex.c), ctypes.byref(self.pc),
                                             g.data_ptr() if g.numel() else None, dvol.data_ptr(),
                                             diso.data_ptr(), stream))
        return dvol, diso


class _MarchingCubes(torch.autograd.Function):
    @staticmethod
    def forward(fctx, volume, isovalue, block_dim, ctx):
        ext = Extraction(volume.detach(), float(isovalue), block_dim, ctx)
        fctx.ext = ext
        fctx.iso_shape = isovalue.shape if torch.is_tensor(isovalue) else None
        # the adjoint recomputes t from the volume: saving it lets torch's
        # version counter reject an in-place change between forward and backward
        fctx.save_for_backward(volume)
        fctx.mark_non_differentiable(ext.faces)
        return ext.positions, ext.faces

    @staticmethod
    def backward(fctx, grad_pos, grad_faces):
        fctx.saved_tensors  # raises if the volume was modified in place since the forward pass
        dvol, diso = fctx.ext.adjoint(grad_pos)
        if fctx.iso_shape is not None:
            diso = diso.reshape(fctx.iso_shape)
        else:
            diso = None
        return dvol.to(fctx.ext.volume.dtype) if fctx.ext.volume.dtype == torch.float32 else None, diso, None, None


def marching_cubes(volume, isovalue, block_dim=8, ctx=None):
    """Differentiable extraction: (positions [n, 3] float32 voxel units, faces [m, 3] int64)."""
    if not torch.is_tensor(isovalue):
        isovalue = torch.tensor(float(isovalue), dtype=torch.float32, device=volume.device)
    return _MarchingCubes.apply(volume, isovalue, block_dim, ctx)
