"""Sparse marching cubes: drop-in for the reference's ``sparse_marching_cubes``
(cubvh/api.py:173-187 over src/api_gpu.cu:154-211).

Arguments are handled as the reference wrapper handles them (coords cast to int32,
corners to float32, both made contiguous and moved to the GPU) and checked as its
binding checks them (shapes [N, 3] and [N, 8], equal N: ``RuntimeError``).  The
extraction runs in libcubvh_amd.so on torch's current stream.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def sparse_marching_cubes(coords, corners, iso, ensure_consistency=False):
    """Triangulate the iso level through sparse cells.

    coords: int [N, 3], the lowest lattice point of each cell; corners: float [N, 8],
    the values at its corners (corner k at offset ((k ^ k >> 1) & 1, k >> 1 & 1, k >> 2));
    iso: level (values below it are inside); ensure_consistency: average the values of
    lattice points shared by several cells first.
    Returns (vertices float32 [M, 3] in lattice units, triangles int32 [T, 3]) on the GPU.
    """
    coords = coords.int().contiguous()
    corners = corners.float().contiguous()
    if not coords.is_cuda:
        coords = coords.cuda()
    if not corners.is_cuda:
        corners = corners.cuda()
    # src/api_gpu.cu:160-169
    if coords.dim() != 2 or coords.size(1) != 3:
        raise RuntimeError("coords must be of shape [N,3]")
    if corners.dim() != 2 or corners.size(1) != 8:
        raise RuntimeError("corners must be of shape [N,8]")
    if coords.size(0) != corners.size(0):
        raise RuntimeError("coords and corners must have the same first-dim (N)")
    if corners.device != coords.device:
        corners = corners.to(coords.device)
    dev = coords.device
    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    h = ctypes.c_void_p()
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    _lib.check(lib.cubvh_sparse_marching_cubes(dev.index, coords.data_ptr(), corners.data_ptr(), coords.size(0),
                                               float(iso), int(bool(ensure_consistency)), ctypes.byref(h),
                                               ctypes.byref(nv), ctypes.byref(nt), stream))
    try:
        verts = torch.empty(nv.value, 3, dtype=torch.float32, device=dev)
        tris = torch.empty(nt.value, 3, dtype=torch.int32, device=dev)
        _lib.check(lib.cubvh_smc_write(h, verts.data_ptr(), tris.data_ptr()))
    finally:
        lib.cubvh_smc_destroy(h)
    return verts, tris
