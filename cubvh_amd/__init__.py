"""cubvh_amd: the BVH query path of ashawkey/cubvh, MI355X-native (HIP, gfx950).

Drop-in for the reference package's exports (cubvh/__init__.py:1-11): use
``import cubvh_amd as cubvh``.  The BVH query path (cuBVH, floodfill) and the
sparse-voxel hash table (cuHashTable) and sparse marching cubes run in
libcubvh_amd.so; the other reference names (the CPU-only ones, mesh repair and
decimation) raise NotImplementedError.
Additions: :func:`udf` (differentiable unsigned distance), :class:`cuBVHForest`
(one launch over many meshes), :class:`ForestGraph` (HIP-graph replay of a forest
query), :func:`udf_batched`, :func:`tune_forest` (launch autotuner),
:func:`morton_order` + ``cuBVH.distance_ordered`` (order a point set once, query it
many times).
"""
from .api import (
    cuBVH,
    HashTable,
    sparse_marching_cubes_cpu,
    floodfill,
    fill_holes,
    merge_vertices,
    decimate,
    parallel_decimate,
    morton_order,
)
from .autograd import udf
from .hashtable import cuHashTable, create_cuHashTable
from .marching_cubes import sparse_marching_cubes
from .forest import cuBVHForest, ForestGraph, udf_batched
from .autotune import tune_forest

__all__ = [
    "cuBVH",
    "HashTable",
    "cuHashTable",
    "sparse_marching_cubes",
    "sparse_marching_cubes_cpu",
    "floodfill",
    "fill_holes",
    "merge_vertices",
    "decimate",
    "parallel_decimate",
    "udf",
    "cuBVHForest",
    "ForestGraph",
    "udf_batched",
    "tune_forest",
    "morton_order",
    "create_cuHashTable",
]
