"""Batched distance queries over many meshes (no counterpart in the reference).

``cuBVHForest([bvh_0, ..., bvh_{M-1}])`` concatenates the trees of several
:class:`cuBVH` objects; ``unsigned_distance(points, offsets)`` answers queries for
all meshes in one launch: ``points`` [N, 3] holds the queries of mesh 0, then mesh 1,
..., and ``offsets`` [M + 1] (CSR) delimits them.  Results per query are identical to
``bvh_m.unsigned_distance`` on the same points; face ids are per-mesh.

``ForestGraph`` captures one such call (fixed shapes and buffers) into a HIP graph
and replays it; the workspace is reserved before capture, so the captured sequence
is kernels only.  ``udf_batched`` adds the backward pass.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .api import cuBVH, _query_tensor, _stream_of

# Morton-prefix bits, persistent grid, waves per CU, waves per workgroup, XCD chunk
DEFAULT_TUNING = (12, 1, 8, 4, 0)


def _offsets_tensor(offsets, n, device, check=True):
    """offsets as an int64 tensor on `device`.  check: verify on the host that they are
    a CSR of the n queries (offsets[0] == 0, non-decreasing, offsets[-1] == n); the
    kernels read them on the device only.  Offsets already on a GPU are copied back for
    the check (a synchronization), so callers capturing a graph check before capture."""
    if not torch.is_tensor(offsets):
        offsets = torch.as_tensor(offsets)
    if offsets.dim() != 1 or offsets.numel() < 2:
        raise ValueError("offsets must be a 1-D tensor of n_meshes + 1 entries")
    if offsets.dtype.is_floating_point or offsets.dtype == torch.bool:
        raise ValueError(f"offsets must be integers, got {offsets.dtype}")
    if check:
        h = offsets.detach().to("cpu", torch.int64)
        if int(h[0]) != 0:
            raise ValueError(f"offsets[0] must be 0, got {int(h[0])}")
        if int(h[-1]) != n:
            raise ValueError(f"offsets[-1] must equal the number of queries {n}, got {int(h[-1])}")
        if bool((h[1:] < h[:-1]).any()):
            raise ValueError("offsets must be non-decreasing")
    return offsets.to(device=device, dtype=torch.int64).contiguous()


class cuBVHForest:
    def __init__(self, bvhs):
        bvhs = [b if isinstance(b, cuBVH) else cuBVH(*b) for b in bvhs]
        if not bvhs:
            raise ValueError("need at least one mesh")
        lib = _lib.load()
        arr = (ctypes.c_void_p * len(bvhs))(*[b._h.value for b in bvhs])
        h = ctypes.c_void_p()
        _lib.check(lib.cubvh_forest_create(arr, len(bvhs), ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self.bvhs = bvhs
        self.n_meshes = len(bvhs)
        self._face_base = None
        self._cat = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.cubvh_forest_destroy(h)
            self._h = None

    def info(self):
        arr = (ctypes.c_int64 * 4)()
        _lib.check(self._lib.cubvh_forest_info(self._h, arr))
        return {"n_meshes": arr[0], "n_nodes": arr[1], "n_faces": arr[2], "stack_need": arr[3]}

    def reserve(self, max_queries: int, max_order_bits: int = 16, device=None):
        """Upload the forest and allocate the workspace for up to max_queries queries
        (required before capturing a call in a graph)."""
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        s = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(self._lib.cubvh_forest_reserve(self._h, dev.index, int(max_queries), int(max_order_bits), s))

    def unsigned_distance(self, points, offsets, return_uvw=False, tuning=None, out=None, check=True):
        """Distances of points [N, 3] (concatenated per mesh, split by offsets [M + 1])
        to their own mesh.  Returns (distances [N], face_id [N] int64, uvw [N, 3] or None).
        ``out`` may pass preallocated (distances, face_id, uvw) buffers.  check=False
        skips the host check of the offsets (for calls captured in a graph, whose
        offsets were checked before capture)."""
        points = _query_tensor(points)
        points = points.view(-1, 3)
        n = points.shape[0]
        dev = points.device
        offsets = _offsets_tensor(offsets, n, dev, check)
        if offsets.numel() != self.n_meshes + 1:
            raise ValueError(f"offsets must have {self.n_meshes + 1} entries")
        if out is None:
            dist = torch.empty(n, dtype=torch.float32, device=dev)
            face = torch.empty(n, dtype=torch.int64, device=dev)
            uvw = torch.empty(n, 3, dtype=torch.float32, device=dev) if return_uvw else None
        else:
            dist, face, uvw = out
        t = tuple(tuning) if tuning is not None else DEFAULT_TUNING
        if not 1 <= len(t) <= 5:
            raise ValueError("tuning holds 1 to 5 entries (bits, persistent, waves per CU, waves per workgroup, "
                             "XCD chunk)")
        tarr = (ctypes.c_int32 * len(t))(*[int(x) for x in t])
        _lib.check(self._lib.cubvh_forest_unsigned_distance(
            self._h, dev.index, n, points.data_ptr(), offsets.data_ptr(), dist.data_ptr(), face.data_ptr(),
            uvw.data_ptr() if uvw is not None else None, tarr, len(t), _stream_of(points)))
        return dist, face, uvw

    # ------------------------------------------------------------- helpers
    def face_base(self):
        """Offset of each mesh's faces in the concatenated face list [M]."""
        if self._face_base is None:
            counts = torch.tensor([b._faces.shape[0] for b in self.bvhs], dtype=torch.int64)
            self._face_base = torch.cumsum(counts, 0) - counts
        return self._face_base

    def concatenated_mesh(self, device):
        """Vertices [sum V, 3] and faces [sum F, 3] (global vertex indices) of all meshes."""
        vs, fs, base = [], [], 0
        for b in self.bvhs:
            vs.append(torch.from_numpy(b._vertices))
            fs.append(torch.from_numpy(b._faces).to(torch.int64) + base)
            base += b._vertices.shape[0]
        return torch.cat(vs).to(device), torch.cat(fs).to(torch.int32).to(device)


class ForestGraph:
    """One forest query captured in a HIP graph (static buffers; copy new points into
    ``points`` and call :meth:`replay`)."""

    def __init__(self, forest: cuBVHForest, points, offsets, return_uvw=False, tuning=None, warmup=1):
        self.forest = forest
        self.points = _query_tensor(points).view(-1, 3).clone()
        dev = self.points.device
        n = self.points.shape[0]
        self.offsets = _offsets_tensor(offsets, n, dev).clone()
        t = tuning if tuning is not None else DEFAULT_TUNING
        forest.reserve(n, int(t[0]), dev)
        self.dist = torch.empty(n, dtype=torch.float32, device=dev)
        self.face = torch.empty(n, dtype=torch.int64, device=dev)
        self.uvw = torch.empty(n, 3, dtype=torch.float32, device=dev) if return_uvw else None
        out = (self.dist, self.face, self.uvw)
        s = torch.cuda.Stream(device=dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            for _ in range(warmup):
                forest.unsigned_distance(self.points, self.offsets, return_uvw, t, out, check=False)
        torch.cuda.current_stream(dev).wait_stream(s)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            forest.unsigned_distance(self.points, self.offsets, return_uvw, t, out, check=False)

    def replay(self):
        self.graph.replay()
        return self.dist, self.face, self.uvw


class _BatchedUDF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, vertices, forest, offsets, faces, face_base_per_query, tuning):
        dist, face, uvw = forest.unsigned_distance(points, offsets, True, tuning, check=False)
        gface = face + face_base_per_query
        ctx.save_for_backward(points, gface, uvw, vertices, faces)
        ctx.forest = forest
        ctx.mark_non_differentiable(face, uvw)
        return dist, face, uvw

    @staticmethod
    def backward(ctx, g_dist, g_face, g_uvw):
        from .autograd import udf_backward, vertex_corner_table

        points, gface, uvw, vertices, faces = ctx.saved_tensors
        need_p, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_p or need_v):
            return None, None, None, None, None, None, None
        g = g_dist.float().contiguous()
        # the concatenated faces never change: their vertex -> corners table is built once
        table = vertex_corner_table(ctx.forest, faces, vertices.shape[0]) if need_v else None
        gp, gv = udf_backward(points, gface, uvw, vertices, faces, g, need_p, need_v, table)
        return gp, gv, None, None, None, None, None


def udf_batched(forest: cuBVHForest, points, offsets, vertices=None, tuning=None):
    """Differentiable batched unsigned distance.  vertices: the concatenated vertex
    tensor of all meshes ([sum V, 3], as returned by forest.concatenated_mesh) when
    vertex gradients are wanted."""
    p = points if points.is_cuda else points.cuda()
    if p.dtype != torch.float32:
        p = p.float()
    p = p.reshape(-1, 3).contiguous()
    dev = p.device
    offsets = _offsets_tensor(offsets, p.shape[0], dev)
    cat = getattr(forest, "_cat", None)
    if cat is None or cat[0].device != dev:
        cat = forest.concatenated_mesh(dev)
        forest._cat = cat
    vcat, fcat = cat
    v = vertices if vertices is not None else vcat
    if v.dtype != torch.float32 or v.device != dev:
        v = v.to(device=dev, dtype=torch.float32)
    v = v.contiguous()
    counts = (offsets[1:] - offsets[:-1]).to(torch.int64)
    # output_size: the total is known, so no device synchronization
    fb = torch.repeat_interleave(forest.face_base().to(dev), counts, output_size=p.shape[0])
    return _BatchedUDF.apply(p, v, forest, offsets, fcat, fb, tuning if tuning is not None else DEFAULT_TUNING)
