"""Differentiable unsigned distance (no counterpart in the reference).

``udf(bvh, points, vertices)`` returns the same distances as
``bvh.unsigned_distance`` and supports backward to ``points`` [N, 3] and to
``vertices`` [V, 3] through a hand-written HIP kernel (cubvh_udf_backward): with the
stored nearest face f and barycentrics w of the closest point c,

    d dist / d p   = n,              n = (p - c) / |p - c|
    d dist / d v_k = -w_k n          for the k-th vertex of f

(the envelope theorem on the minimum over the face).  The forward already stores
face_id and uvw, so the backward is a single pass with no traversal.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .api import cuBVH, _query_tensor, _stream_of


def vertex_corner_table(owner, faces, n_vertices):
    """The vertex -> corners table of a mesh's device faces (cubvh_vertex_corners),
    built once and kept on `owner` (the BVH or forest whose faces never change) per
    device and vertex count: (corners [3F], begin [V], end [V]) uint32 tensors, as int32."""
    key = (faces.device, int(faces.data_ptr()), int(faces.shape[0]), int(n_vertices))
    table = getattr(owner, "_corner_table", None)
    if table is not None and table[0] == key:
        return table[1]
    dev = faces.device
    corners = torch.empty(3 * faces.shape[0], dtype=torch.int32, device=dev)
    begin = torch.empty(n_vertices, dtype=torch.int32, device=dev)
    end = torch.empty(n_vertices, dtype=torch.int32, device=dev)
    lib = _lib.load()
    _lib.check(lib.cubvh_vertex_corners(dev.index, faces.data_ptr(), faces.shape[0], n_vertices, corners.data_ptr(),
                                        begin.data_ptr(), end.data_ptr(), _stream_of(faces)))
    owner._corner_table = (key, (corners, begin, end))
    return corners, begin, end


def udf_backward(points, face_id, uvw, vertices, faces, g, need_p, need_v, table=None):
    """Point and vertex gradients of sum(g * distance) (cubvh_udf_backward, or
    cubvh_udf_backward_with_corners when the mesh's vertex -> corners table is given)."""
    dev = points.device
    gp = torch.empty_like(points) if need_p else None
    gv = torch.zeros_like(vertices) if need_v else None
    lib = _lib.load()
    args = [dev.index, points.shape[0], g.data_ptr(), points.data_ptr(), face_id.data_ptr(), uvw.data_ptr(),
            vertices.data_ptr(), vertices.shape[0], faces.data_ptr(), faces.shape[0],
            gp.data_ptr() if gp is not None else None, gv.data_ptr() if gv is not None else None]
    if table is not None and need_v:
        _lib.check(lib.cubvh_udf_backward_with_corners(*args, table[0].data_ptr(), table[1].data_ptr(),
                                                       table[2].data_ptr(), _stream_of(points)))
    else:
        _lib.check(lib.cubvh_udf_backward(*args, _stream_of(points)))
    return gp, gv


class _UnsignedDistance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, vertices, bvh, faces):
        dist, face_id, uvw = bvh.unsigned_distance(points, return_uvw=True)
        ctx.save_for_backward(points, face_id, uvw, vertices, faces)
        ctx.bvh = bvh
        ctx.mark_non_differentiable(face_id, uvw)
        return dist, face_id, uvw

    @staticmethod
    def backward(ctx, g_dist, g_face, g_uvw):
        points, face_id, uvw, vertices, faces = ctx.saved_tensors
        need_p, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_p or need_v):
            return None, None, None, None
        g = g_dist.float().contiguous()
        table = vertex_corner_table(ctx.bvh, faces, vertices.shape[0]) if need_v else None
        gp, gv = udf_backward(points, face_id, uvw, vertices, faces, g, need_p, need_v, table)
        return gp, gv, None, None


def udf(bvh: cuBVH, points, vertices=None):
    """Differentiable unsigned distance.

    points: [..., 3]; vertices: the mesh's [V, 3] vertex tensor (same values the BVH
    was built from) when gradients to the vertices are wanted, else None.
    Returns (distances [...], face_id [...], uvw [..., 3]); only distances carry a
    gradient.
    """
    p = points if points.is_cuda else points.cuda()
    if p.dtype != torch.float32:
        p = p.float()
    prefix = p.shape[:-1]
    p = p.reshape(-1, 3).contiguous()
    dev = p.device
    if vertices is None:
        v = torch.from_numpy(bvh._vertices).to(dev)
    else:
        v = vertices if vertices.device == dev else vertices.to(dev)
        if v.dtype != torch.float32:
            v = v.float()
        v = v.contiguous()
    f = getattr(bvh, "_faces_dev", None)
    if f is None or f.device != dev:
        f = torch.from_numpy(bvh._faces).to(dev)
        bvh._faces_dev = f
    if v.dim() != 2 or v.shape[1] != 3 or v.shape[0] <= int(bvh._faces.max()):
        raise ValueError("vertices do not match the mesh of this BVH")
    dist, face_id, uvw = _UnsignedDistance.apply(p, v, bvh, f)
    return dist.view(*prefix), face_id.view(*prefix), uvw.view(*prefix, 3)
