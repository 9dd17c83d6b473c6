"""Sparse-voxel hash table: drop-in for the reference's ``cuHashTable``
(cubvh/api.py:143-171 over the pybind11 class of src/bindings.cpp:34-44).

``create_cuHashTable()`` returns the object the reference's ``_backend`` factory
returns, with the same methods (``set_num_dims``, ``get_num_dims``, ``resize``,
``prepare``, ``insert``, ``build``, ``search``) and the same argument checks
(src/api_gpu.cu:239-274: int32 tensors on the GPU, 2-D).  ``cuHashTable`` is the
reference's Python wrapper.  Every call runs the HIP kernels of libcubvh_amd.so on
torch's current stream; the table lives on the device of the first tensor it is
sized for.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _stream_of(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check_keys(t, name):
    # src/api_gpu.cu:240-242, 252-254, 264-266
    if not torch.is_tensor(t) or not t.is_cuda:
        raise RuntimeError(f"{name} must reside on CUDA")
    if t.dtype != torch.int32:
        raise RuntimeError(f"{name} must be int32")
    if t.dim() != 2:
        raise RuntimeError(f"{name} must be 2D [N,D]")
    return t.contiguous()


class _HashTableImpl:
    """The table object of ``create_cuHashTable()`` (include/gpu/api_gpu.h:37-49)."""

    def __init__(self, num_dims: int = 3):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.cubvh_hashtable_create(int(num_dims), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.cubvh_hashtable_destroy(h)
            self._h = None

    def _info(self):
        info = (ctypes.c_int64 * 5)()
        _lib.check(self._lib.cubvh_hashtable_info(self._h, info))
        return list(info)

    def set_num_dims(self, d: int) -> None:
        _lib.check(self._lib.cubvh_hashtable_set_num_dims(self._h, int(d)))

    def get_num_dims(self) -> int:
        return int(self._info()[0])

    @property
    def capacity(self) -> int:
        return int(self._info()[1])

    def resize(self, capacity: int) -> None:
        """capacity empty slots on the current CUDA device."""
        dev = torch.device("cuda", torch.cuda.current_device())
        _lib.check(self._lib.cubvh_hashtable_resize(self._h, dev.index, int(capacity), _stream_of(dev)))

    def prepare(self) -> None:
        """Mark every slot empty."""
        info = self._info()
        dev = torch.device("cuda", int(info[4]) if info[4] >= 0 else torch.cuda.current_device())
        _lib.check(self._lib.cubvh_hashtable_prepare(self._h, _stream_of(dev)))

    def insert(self, coords) -> None:
        """Insert the rows of coords int32 [N, D] (row i under index i)."""
        coords = _check_keys(coords, "coords")
        n, d = coords.shape
        _lib.check(self._lib.cubvh_hashtable_insert(self._h, coords.device.index, coords.data_ptr(), n, d,
                                                    _stream_of(coords.device)))

    def build(self, coords) -> None:
        """capacity = max(16, 2N), clear, insert."""
        coords = _check_keys(coords, "coords")
        n, d = coords.shape
        _lib.check(self._lib.cubvh_hashtable_build(self._h, coords.device.index, coords.data_ptr(), n, d,
                                                   _stream_of(coords.device)))

    def search(self, queries) -> torch.Tensor:
        """int32 [M]: the lowest row index stored with each query key, or -1."""
        queries = _check_keys(queries, "queries")
        m, d = queries.shape
        out = torch.empty(m, dtype=torch.int32, device=queries.device)
        _lib.check(self._lib.cubvh_hashtable_search(self._h, queries.device.index, queries.data_ptr(), m, d,
                                                    out.data_ptr(), _stream_of(queries.device)))
        return out


def create_cuHashTable() -> _HashTableImpl:
    """src/api_gpu.cu:278-280."""
    return _HashTableImpl()


class cuHashTable:
    """GPU hash table of integer N-D coordinates (cubvh/api.py:143-171).

    - Default dimensionality is 3; change it with num_dims or impl.set_num_dims.
    - Static table: build() once, then search().
    """

    def __init__(self, num_dims: int = 3):
        self.impl = create_cuHashTable()
        self.impl.set_num_dims(int(num_dims))

    @property
    def num_dims(self) -> int:
        return int(self.impl.get_num_dims())

    def build(self, coords):
        """Build the table from coords [N, D] int32 on the GPU; capacity max(16, 2N)."""
        if coords.shape[1] != self.num_dims:
            self.impl.set_num_dims(int(coords.size(1)))
        self.impl.build(coords)

    def search(self, queries) -> torch.Tensor:
        """queries [M, D] -> indices [M] int32 on the GPU; -1 where absent."""
        assert queries.shape[1] == self.num_dims, f"queries must be {self.num_dims}D"
        return self.impl.search(queries)
