"""Python API of cubvh_amd: a drop-in for ``cubvh/api.py`` of ashawkey/cubvh on its BVH
query path (``cuBVH``, ``ray_trace``, ``unsigned_distance``, ``signed_distance``,
``state_dict`` / ``from_state_dict``, ``floodfill``).

Argument handling mirrors the reference wrapper (cubvh/api.py:12-140): queries are
cast to fp32, made contiguous, moved to the GPU, flattened to [N, 3]; outputs are
allocated here and reshaped to the input's leading dimensions.  Every query runs
in libcubvh_amd.so (HIP, gfx950) on torch's current stream.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

# cubvh/api.py:7-10
_sdf_mode_to_id = {
    "watertight": 0,
    "raystab": 1,
}


def _as_mesh_arrays(vertices, triangles):
    if torch.is_tensor(vertices):
        vertices = vertices.detach().cpu().numpy()
    if torch.is_tensor(triangles):
        triangles = triangles.detach().cpu().numpy()
    # The reference's pybind11 caster force-casts to float32 / uint32
    # (patch/eigen/matrix.h:490-524); geometry is fp32 from here on.
    v = np.ascontiguousarray(np.asarray(vertices), dtype=np.float32)
    f = np.asarray(triangles)
    if v.ndim != 2 or v.shape[1] != 3:
        raise ValueError(f"vertices must have shape [N, 3], got {v.shape}")
    if f.ndim != 2 or f.shape[1] != 3:
        raise ValueError(f"triangles must have shape [M, 3], got {f.shape}")
    if f.size and (f.min() < 0 or f.max() >= v.shape[0]):
        raise ValueError("triangles reference vertices outside [0, N)")
    f = np.ascontiguousarray(f, dtype=np.int32)
    return v, f


def _query_tensor(x):
    # cubvh/api.py:70-72: fp32, contiguous, on the GPU
    x = x.float().contiguous()
    if not x.is_cuda:
        x = x.cuda()
    return x


def _stream_of(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


class cuBVH:
    """BVH over a triangle mesh (cubvh/api.py:12-126).

    vertices: [N, 3] float array or tensor (any device); triangles: [M, 3] integer.
    build="host" (default): binned-SAH build on the host, no GPU touched; the tree is
    copied to a GPU on first use.  build="gpu": built on `device` (default the current
    CUDA device) by object-median splits on each run's longest centroid axis, level by
    level; build="gpu_sah": every run split where the surface-area cost over the sorted
    face lists is least, the slowest GPU build and the fastest to query; build="gpu_lbvh":
    the Morton radix tree (linear BVH), the fastest build and the slowest to query.
    Query results are the same whatever the build.  leaf_size: faces per leaf, 1..16
    (0 = 16, the fastest for distance queries; 8 suits ray tracing better).
    """

    def __init__(self, vertices, triangles, leaf_size: int = 0, build: str = "host", device=None):
        v, f = _as_mesh_arrays(vertices, triangles)
        # cubvh/api.py:21
        assert f.shape[0] > 8, "BVH needs at least 8 triangles."
        lib = _lib.load()
        h = ctypes.c_void_p()
        if build == "host":
            _lib.check(lib.cubvh_bvh_create(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0], int(leaf_size),
                                            ctypes.byref(h)))
        elif build in ("gpu", "gpu_sah", "gpu_lbvh"):
            dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
            vt = torch.from_numpy(v).to(dev)
            ft = torch.from_numpy(f).to(dev)
            flags = int(leaf_size) | {"gpu": 0, "gpu_lbvh": 0x100, "gpu_sah": 0x200}[build]
            _lib.check(lib.cubvh_bvh_create_gpu(dev.index, vt.data_ptr(), v.shape[0], ft.data_ptr(), f.shape[0],
                                                flags, ctypes.byref(h), _stream_of(vt)))
        else:
            raise ValueError(f"build must be 'host', 'gpu', 'gpu_sah' or 'gpu_lbvh', not {build!r}")
        self._h = h
        self._lib = lib
        self._faces = f
        self._vertices = v

    @staticmethod
    def build_many(meshes, build: str = "gpu", leaf_size: int = 0, device=None, streams: int = 4):
        """Trees for many meshes built on the GPU concurrently: `streams` host threads,
        each with its own stream (a GPU build is a few dozen short launches and one or
        more synchronizations, so one mesh alone leaves the GPU mostly idle).  meshes:
        (vertices, triangles) pairs as accepted by the constructor (host arrays are
        copied to `device`; device tensors are read where they are, after the current
        stream's pending work).  build: "gpu", "gpu_sah" or "gpu_lbvh".  Returns the
        cuBVH objects in input order, as `cuBVH(v, f, build=build)` would."""
        from concurrent.futures import ThreadPoolExecutor

        flags = {"gpu": 0, "gpu_lbvh": 0x100, "gpu_sah": 0x200}.get(build)
        if flags is None:
            raise ValueError(f"build must be 'gpu', 'gpu_sah' or 'gpu_lbvh', not {build!r}")
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        lib = _lib.load()
        producer = torch.cuda.current_stream(dev)
        pool = [torch.cuda.Stream(dev) for _ in range(max(1, int(streams)))]
        for st in pool:
            st.wait_stream(producer)

        def one(job):
            i, (vertices, triangles) = job
            st = pool[i % len(pool)]
            with torch.cuda.device(dev), torch.cuda.stream(st):
                if torch.is_tensor(vertices) and vertices.device == dev and torch.is_tensor(triangles) \
                        and triangles.device == dev:
                    vt = vertices.detach().to(torch.float32).contiguous()
                    ft = triangles.detach().to(torch.int32).contiguous()
                    v = f = None
                else:
                    v, f = _as_mesh_arrays(vertices, triangles)
                    vt = torch.from_numpy(v).to(dev)
                    ft = torch.from_numpy(f).to(dev)
                assert ft.shape[0] > 8, "BVH needs at least 8 triangles."
                h = ctypes.c_void_p()
                _lib.check(lib.cubvh_bvh_create_gpu(dev.index, vt.data_ptr(), vt.shape[0], ft.data_ptr(), ft.shape[0],
                                                    int(leaf_size) | flags, ctypes.byref(h),
                                                    ctypes.c_void_p(st.cuda_stream)))
                obj = cuBVH.__new__(cuBVH)
                obj._h = h
                obj._lib = lib
                if v is None:
                    v = vt.cpu().numpy()
                    f = ft.cpu().numpy()
                obj._vertices = v
                obj._faces = f
                return obj

        with ThreadPoolExecutor(max_workers=len(pool)) as ex:
            return list(ex.map(one, enumerate(meshes)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.cubvh_bvh_destroy(h)
            self._h = None

    # ------------------------------------------------------------------ info
    def info(self) -> dict:
        arr = (ctypes.c_int64 * 5)()
        _lib.check(self._lib.cubvh_bvh_info(self._h, arr))
        return {"n_nodes": arr[0], "n_faces": arr[1], "depth": arr[2], "stack_need": arr[3], "leaf_size": arr[4]}

    # ------------------------------------------------------- serialization
    def state_dict(self, layout: str = "own"):
        """dict of str -> CPU int32 tensor: 'nodes' [n_nodes, 32], 'triangles' [M, 12]
        (cubvh/api.py:26-28; bvh.cuh:84-94).  The triangle rows use the reference's
        48-byte record (a, b, c fp32, 4 zero bytes, int64 face id).  layout="reference"
        writes the reference's own node layout (nodes [n, 9]), which the reference
        package loads with its from_state_dict."""
        if layout == "reference":
            n9 = ctypes.c_int64(0)
            _lib.check(self._lib.cubvh_bvh_export_reference(self._h, None, ctypes.byref(n9), None))
            nodes = torch.empty(n9.value, 9, dtype=torch.int32)
            tris = torch.empty(self.info()["n_faces"], 12, dtype=torch.int32)
            _lib.check(self._lib.cubvh_bvh_export_reference(self._h, nodes.data_ptr(), ctypes.byref(n9),
                                                            tris.data_ptr()))
            return {"nodes": nodes, "triangles": tris}
        if layout != "own":
            raise ValueError(f"layout must be 'own' or 'reference', not {layout!r}")
        info = self.info()
        nodes = torch.empty(info["n_nodes"], 32, dtype=torch.int32)
        tris = torch.empty(info["n_faces"], 12, dtype=torch.int32)
        _lib.check(self._lib.cubvh_bvh_export(self._h, nodes.data_ptr(), tris.data_ptr()))
        return {"nodes": nodes, "triangles": tris}

    @staticmethod
    def from_state_dict(state_dict):
        """Rebuild from :meth:`state_dict` output, or from a state_dict written by the
        reference (nodes [n, 9]), whose triangle records are rebuilt into this layout
        (cubvh/api.py:30-36; src/api_gpu.cu:47-68)."""
        for key in ("nodes", "triangles"):
            if key not in state_dict:
                raise RuntimeError(f"state_dict missing '{key}' tensor")
        nodes = state_dict["nodes"]
        tris = state_dict["triangles"]
        if not (torch.is_tensor(nodes) and torch.is_tensor(tris)):
            raise RuntimeError("state_dict entries must be tensors")
        if not tris.device.type == "cpu" or not nodes.device.type == "cpu":
            raise RuntimeError("triangles must be on CPU")
        nodes = nodes.to(torch.int32).contiguous()
        tris = tris.to(torch.int32).contiguous()
        if tris.dim() != 2 or tris.shape[1] != 12:
            raise RuntimeError(f"'triangles' must be int32 [M, 12], got {tuple(tris.shape)}")
        lib = _lib.load()
        h = ctypes.c_void_p()
        if nodes.dim() == 2 and nodes.shape[1] == 32:
            rc = lib.cubvh_bvh_import(nodes.data_ptr(), nodes.shape[0], tris.data_ptr(), tris.shape[0],
                                      ctypes.byref(h))
        elif nodes.dim() == 2 and nodes.shape[1] == 9:
            rc = lib.cubvh_bvh_import_reference(nodes.data_ptr(), nodes.shape[0], tris.data_ptr(), tris.shape[0], 0,
                                                ctypes.byref(h))
        else:
            raise RuntimeError(f"'nodes' must be int32 [n, 32] or [n, 9], got {tuple(nodes.shape)}")
        _lib.check(rc)
        obj = cuBVH.__new__(cuBVH)
        obj._h = h
        obj._lib = lib
        m = tris.shape[0]
        tv = np.empty((m, 3, 3), dtype=np.float32)
        _lib.check(lib.cubvh_bvh_triangles(h, tv.ctypes.data))
        # the mesh is recovered as a triangle soup (3 vertices per face)
        obj._vertices = np.ascontiguousarray(tv.reshape(-1, 3))
        obj._faces = np.arange(3 * m, dtype=np.int32).reshape(m, 3)
        return obj

    # ------------------------------------------------------------ queries
    def ray_trace(self, rays_o, rays_d):
        """Closest hit per ray (cubvh/api.py:38-65).  Returns positions [..., 3],
        face_id int64 [...] (-1 on a miss), depth [...] (1000 on a miss)."""
        rays_o = _query_tensor(rays_o)
        rays_d = _query_tensor(rays_d).to(rays_o.device)
        prefix = rays_o.shape[:-1]
        rays_o = rays_o.view(-1, 3)
        rays_d = rays_d.view(-1, 3)
        if rays_d.shape != rays_o.shape:
            raise ValueError("rays_o and rays_d must have the same shape")
        n = rays_o.shape[0]
        positions = torch.empty(n, 3, dtype=torch.float32, device=rays_o.device)
        face_id = torch.empty(n, dtype=torch.int64, device=rays_o.device)
        depth = torch.empty(n, dtype=torch.float32, device=rays_o.device)
        _lib.check(self._lib.cubvh_ray_trace(self._h, rays_o.device.index, n, rays_o.data_ptr(), rays_d.data_ptr(),
                                             positions.data_ptr(), face_id.data_ptr(), depth.data_ptr(),
                                             _stream_of(rays_o)))
        return positions.view(*prefix, 3), face_id.view(*prefix), depth.view(*prefix)

    def unsigned_distance(self, positions, return_uvw=False):
        """Distance to the nearest face (cubvh/api.py:67-95).  Returns distances [...],
        face_id int64 [...] and, with return_uvw, barycentrics [..., 3] of the closest
        point (else None)."""
        return self._distance(positions, return_uvw, None)

    def signed_distance(self, positions, return_uvw=False, mode="watertight"):
        """Signed distance, negative inside (cubvh/api.py:98-126); mode 'watertight'
        (normal averaging at the closest point) or 'raystab' (64 stab rays)."""
        return self._distance(positions, return_uvw, _sdf_mode_to_id[mode])

    def distance_ordered(self, positions, order, return_uvw=False, signed=False):
        """unsigned_distance (or watertight signed_distance with signed=True) over a
        query order from morton_order(positions), which a caller querying the same
        points repeatedly computes once.  positions [n, 3] on the GPU."""
        positions = _query_tensor(positions).view(-1, 3)
        n = positions.shape[0]
        dev = positions.device
        if order.dtype != torch.int32 or order.shape != (n,) or order.device != dev:
            raise ValueError("order must be an int32 tensor [n] on the device of positions (from morton_order)")
        order = order.contiguous()
        distances = torch.empty(n, dtype=torch.float32, device=dev)
        face_id = torch.empty(n, dtype=torch.int64, device=dev)
        uvw = torch.empty(n, 3, dtype=torch.float32, device=dev) if return_uvw else None
        _lib.check(self._lib.cubvh_distance_ordered(self._h, dev.index, n, 1 if signed else 0, positions.data_ptr(),
                                                    order.data_ptr(), distances.data_ptr(), face_id.data_ptr(),
                                                    uvw.data_ptr() if uvw is not None else None,
                                                    _stream_of(positions)))
        return distances, face_id, uvw

    def _distance(self, positions, return_uvw, mode):
        positions = _query_tensor(positions)
        prefix = positions.shape[:-1]
        positions = positions.view(-1, 3)
        n = positions.shape[0]
        dev = positions.device
        distances = torch.empty(n, dtype=torch.float32, device=dev)
        face_id = torch.empty(n, dtype=torch.int64, device=dev)
        uvw = torch.empty(n, 3, dtype=torch.float32, device=dev) if return_uvw else None
        uptr = uvw.data_ptr() if uvw is not None else None
        if mode is None:
            rc = self._lib.cubvh_unsigned_distance(self._h, dev.index, n, positions.data_ptr(), distances.data_ptr(),
                                                   face_id.data_ptr(), uptr, _stream_of(positions))
        else:
            rc = self._lib.cubvh_signed_distance(self._h, dev.index, n, mode, positions.data_ptr(),
                                                 distances.data_ptr(), face_id.data_ptr(), uptr,
                                                 _stream_of(positions))
        _lib.check(rc)
        distances = distances.view(*prefix)
        face_id = face_id.view(*prefix)
        if uvw is not None:
            uvw = uvw.view(*prefix, 3)
        return distances, face_id, uvw


def morton_order(positions):
    """Permutation (int32 [n]) that puts query points [n, 3] in Morton order of their
    bounding box, for cuBVH.distance_ordered."""
    positions = _query_tensor(positions).view(-1, 3)
    n = positions.shape[0]
    order = torch.empty(n, dtype=torch.int32, device=positions.device)
    lib = _lib.load()
    _lib.check(lib.cubvh_morton_order(positions.device.index, n, positions.data_ptr(), order.data_ptr(),
                                      _stream_of(positions)))
    return order


def floodfill(grid):
    """Label the 6-connected free (False) voxels of a bool grid [B, H, W, D] or
    [H, W, D] (cubvh/api.py:128-140).  Returns int32 of the same shape: -1 for
    occupied voxels, otherwise the minimum linear index of the voxel's component
    (global across the batch)."""
    if not torch.is_tensor(grid):
        grid = torch.as_tensor(np.asarray(grid))
    if grid.dtype != torch.bool:
        # src/api_gpu.cu:139 asserts a bool grid
        raise TypeError(f"floodfill expects a bool grid, got {grid.dtype}")
    grid = grid.contiguous()
    if not grid.is_cuda:
        grid = grid.cuda()
    squeeze = grid.dim() == 3
    if squeeze:
        grid = grid.unsqueeze(0)
    if grid.dim() != 4:
        raise ValueError(f"floodfill expects [B, H, W, D] or [H, W, D], got {tuple(grid.shape)}")
    B, H, W, D = grid.shape
    labels = torch.empty(B, H, W, D, dtype=torch.int32, device=grid.device)
    lib = _lib.load()
    _lib.check(lib.cubvh_floodfill(grid.device.index, grid.data_ptr(), B, H, W, D, labels.data_ptr(),
                                   _stream_of(grid)))
    return labels.squeeze(0) if squeeze else labels


# ---------------------------------------------------------------------------
# Names exported by the reference package (cubvh/__init__.py:1-11) that are not
# part of the BVH query path.  They exist so that `import cubvh_amd as cubvh`
# keeps working; calling them raises.
# ---------------------------------------------------------------------------
def _out_of_scope(name):
    def fn(*args, **kwargs):
        raise NotImplementedError(f"cubvh_amd does not implement {name} (outside the BVH query path)")

    fn.__name__ = name
    return fn


class _OutOfScopeClass:
    _name = ""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError(f"cubvh_amd does not implement {self._name} (outside the BVH query path)")


class HashTable(_OutOfScopeClass):
    _name = "HashTable"


sparse_marching_cubes_cpu = _out_of_scope("sparse_marching_cubes_cpu")
fill_holes = _out_of_scope("fill_holes")
merge_vertices = _out_of_scope("merge_vertices")
decimate = _out_of_scope("decimate")
parallel_decimate = _out_of_scope("parallel_decimate")
