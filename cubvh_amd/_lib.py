"""ctypes binding of libcubvh_amd.so (C ABI declared in include/cubvh_amd.h).

The library must be built (``python -m cubvh_amd.build``); there is no fallback:
every query runs in the HIP kernels of that library or raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_int64, c_uint8, c_uint32, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcubvh_amd.so"

CUBVH_OK = 0

# Every C entry point and its (restype, argtypes), as declared in include/cubvh_amd.h.
SIGNATURES = {
    "cubvh_last_error": (c_char_p, []),
    "cubvh_abi_version": (c_int, []),
    "cubvh_bvh_create": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int32, POINTER(c_void_p)]),
    "cubvh_bvh_create_gpu": (
        c_int,
        [c_int, c_void_p, c_int64, c_void_p, c_int64, c_int32, POINTER(c_void_p), c_void_p],
    ),
    "cubvh_bvh_export_reference": (c_int, [c_void_p, c_void_p, POINTER(c_int64), c_void_p]),
    "cubvh_bvh_destroy": (None, [c_void_p]),
    "cubvh_bvh_info": (c_int, [c_void_p, c_void_p]),
    "cubvh_bvh_export": (c_int, [c_void_p, c_void_p, c_void_p]),
    "cubvh_bvh_import": (c_int, [c_void_p, c_int64, c_void_p, c_int64, POINTER(c_void_p)]),
    "cubvh_bvh_import_reference": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int32, POINTER(c_void_p)]),
    "cubvh_bvh_triangles": (c_int, [c_void_p, c_void_p]),
    "cubvh_bvh_upload": (c_int, [c_void_p, c_int, c_void_p]),
    "cubvh_unsigned_distance": (c_int, [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "cubvh_unsigned_distance_noalloc": (
        c_int, [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    ),
    "cubvh_morton_order": (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p]),
    "cubvh_valu_probe": (c_int, [c_int, c_int32, c_int32, c_void_p, c_void_p]),
    "cubvh_distance_ordered": (
        c_int,
        [c_void_p, c_int, c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "cubvh_signed_distance": (
        c_int,
        [c_void_p, c_int, c_int64, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "cubvh_ray_trace": (
        c_int,
        [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "cubvh_floodfill": (c_int, [c_int, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    "cubvh_forest_create": (c_int, [c_void_p, c_int32, POINTER(c_void_p)]),
    "cubvh_forest_destroy": (None, [c_void_p]),
    "cubvh_forest_info": (c_int, [c_void_p, c_void_p]),
    "cubvh_forest_reserve": (c_int, [c_void_p, c_int, c_int64, c_int32, c_void_p]),
    "cubvh_forest_unsigned_distance": (
        c_int,
        [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p],
    ),
    "cubvh_udf_backward": (
        c_int,
        [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
         c_void_p, c_void_p],
    ),
    "cubvh_vertex_corners": (c_int, [c_int, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "cubvh_udf_backward_with_corners": (
        c_int,
        [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "cubvh_hashtable_create": (c_int, [c_int32, POINTER(c_void_p)]),
    "cubvh_hashtable_destroy": (None, [c_void_p]),
    "cubvh_hashtable_set_num_dims": (c_int, [c_void_p, c_int32]),
    "cubvh_hashtable_info": (c_int, [c_void_p, c_void_p]),
    "cubvh_hashtable_resize": (c_int, [c_void_p, c_int, c_int64, c_void_p]),
    "cubvh_hashtable_prepare": (c_int, [c_void_p, c_void_p]),
    "cubvh_hashtable_insert": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_int32, c_void_p]),
    "cubvh_hashtable_build": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_int32, c_void_p]),
    "cubvh_hashtable_search": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "cubvh_sparse_marching_cubes": (
        c_int,
        [c_int, c_void_p, c_void_p, c_int64, c_float, c_int32, POINTER(c_void_p), POINTER(c_int64), POINTER(c_int64),
         c_void_p],
    ),
    "cubvh_smc_write": (c_int, [c_void_p, c_void_p, c_void_p]),
    "cubvh_smc_destroy": (None, [c_void_p]),
}

_lib = None


class CubvhError(RuntimeError):
    """A failure reported by libcubvh_amd (status code plus message)."""


def load():
    """Load libcubvh_amd.so once.  Raises if it has not been built.

    torch is imported first so that the library binds to the HIP runtime torch
    already loaded (both use the soname libamdhip64.so.7) and therefore shares its
    streams and allocations.
    """
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (see docstring)

    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m cubvh_amd.build` "
            "(cubvh_amd has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != CUBVH_OK:
        msg = _lib.cubvh_last_error().decode("utf-8", "replace")
        raise CubvhError(f"libcubvh_amd error {rc}: {msg}")
