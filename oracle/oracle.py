"""numpy wrappers over the CPU oracle (libcubvh_oracle*.so, built from cubvh_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product (cubvh_amd/) never
imports it.  See cubvh_oracle.cpp for what each function restates and what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_uint32, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None


REFERENCE = Path("/root/reference")
REF_HASH_LIB = HERE / "_ref" / "libref_hashtable.so"
REF_SMC_LIB = HERE / "_ref" / "libref_smc.so"
REF_PCG_LIB = HERE / "_ref" / "libref_pcg32.so"


def build(verbose: bool = False) -> None:
    """Compile both oracle variants with the committed Makefile (g++), and the
    reference's own CPU hash table, CPU marching cubes and PCG32 generator into _ref/
    when the reference sources are present."""
    r = subprocess.run(["make", "-B", "-C", str(HERE)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    if verbose:
        print(r.stdout)
    if all((REFERENCE / "include" / h).exists() for h in ("cpu/hashtable.h", "cpu/spmc.h", "gpu/pcg32.h")):
        r = subprocess.run(["make", "-B", "-C", str(HERE), "ref", f"REF={REFERENCE}"], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference CPU hash table / marching cubes / PCG32 build failed:\n" + r.stdout)


def _cpu_has_fma() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " fma " in (line.rstrip() + " ")
    except OSError:
        pass
    return False


def load():
    global _lib
    if _lib is not None:
        return _lib
    name = "libcubvh_oracle_fma.so" if _cpu_has_fma() else "libcubvh_oracle.so"
    path = HERE / name
    if not path.exists():
        build()
    lib = ctypes.CDLL(str(path))
    vp, i64 = c_void_p, c_int64
    sig = {
        "orc_num_threads": (c_int, []),
        "orc_tri_dist2": (None, [vp, vp, i64, vp]),
        "orc_tri_closest": (None, [vp, vp, i64, vp, vp]),
        "orc_tri_ray": (None, [vp, vp, vp, i64, vp, c_int]),
        "orc_safe_divide": (c_float, [c_float, c_float]),
        "orc_pcg32_uints": (None, [i64, i64, vp]),
        "orc_pcg32_offsets": (None, [i64, vp]),
        "orc_fibonacci_dirs32": (None, [c_float, c_float, vp]),
        "orc_udf_brute": (c_int, [vp, i64, vp, i64, vp, i64, vp, vp, vp, c_int]),
        "orc_ray_brute": (c_int, [vp, i64, vp, i64, vp, vp, i64, vp, vp, vp, c_int]),
        "orc_sdf_brute": (c_int, [vp, i64, vp, i64, vp, i64, c_uint32, vp, vp, vp, c_int]),
        "orc_udf_f64": (c_int, [vp, i64, vp, i64, vp, i64, vp, vp, vp, vp, c_int]),
        "orc_ray_f64": (c_int, [vp, i64, vp, i64, vp, vp, i64, c_double, vp, vp, vp, c_int]),
        "orc_winding_f64": (c_int, [vp, i64, vp, i64, vp, i64, vp, c_int]),
        "orc_udf_backward": (c_int, [vp, i64, vp, i64, vp, vp, vp, vp, i64, vp, vp]),
        "orc_refbvh_build": (vp, [vp, i64, vp, i64]),
        "orc_refbvh_import": (vp, [vp, i64, vp, i64]),
        "orc_refbvh_free": (None, [vp]),
        "orc_refbvh_num_nodes": (i64, [vp]),
        "orc_refbvh_export": (None, [vp, vp, vp]),
        "orc_refbvh_udf": (c_int, [vp, vp, i64, vp, vp, vp, c_int, vp, c_int]),
        "orc_refbvh_ray": (c_int, [vp, vp, vp, i64, vp, vp, vp, c_int, vp, c_int]),
        "orc_refbvh_sdf": (c_int, [vp, vp, i64, c_uint32, vp, vp, vp, vp, c_int]),
        "orc_refbvh_sdf_idx": (c_int, [vp, vp, vp, i64, c_uint32, c_int, vp, vp, vp, c_int]),
        "orc_floodfill_ref": (c_int, [vp, c_int, c_int, c_int, c_int, vp]),
        "orc_floodfill_bfs": (c_int, [vp, c_int, c_int, c_int, c_int, vp]),
        "orc_hash_build": (i64, [vp, i64, c_int, c_int, vp]),
        "orc_hash_search": (None, [vp, c_int, vp, c_int, vp, i64, vp]),
        "orc_smc": (c_int, [vp, vp, i64, c_float, c_int, vp, vp, vp, vp]),
    }
    for k, (res, args) in sig.items():
        fn = getattr(lib, k)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data if a is not None else None


def num_threads() -> int:
    return load().orc_num_threads()


# ---------------------------------------------------------------- primitives
def tri_dist2(tri, pts):
    tri, pts = _f32(tri).reshape(9), _f32(pts).reshape(-1, 3)
    out = np.empty(len(pts), np.float32)
    load().orc_tri_dist2(_p(tri), _p(pts), len(pts), _p(out))
    return out


def tri_closest(tri, pts):
    tri, pts = _f32(tri).reshape(9), _f32(pts).reshape(-1, 3)
    cp = np.empty_like(pts)
    uvw = np.empty_like(pts)
    load().orc_tri_closest(_p(tri), _p(pts), len(pts), _p(cp), _p(uvw))
    return cp, uvw


def tri_ray(tri, ro, rd, exact_div=True):
    tri, ro, rd = _f32(tri).reshape(9), _f32(ro).reshape(-1, 3), _f32(rd).reshape(-1, 3)
    out = np.empty(len(ro), np.float32)
    load().orc_tri_ray(_p(tri), _p(ro), _p(rd), len(ro), _p(out), int(exact_div))
    return out


def safe_divide(a, b):
    return load().orc_safe_divide(float(a), float(b))


def pcg32_uints(advance, n):
    out = np.empty(n, np.uint32)
    load().orc_pcg32_uints(int(advance), int(n), _p(out))
    return out


def pcg32_offsets(n):
    out = np.empty((n, 2), np.float32)
    load().orc_pcg32_offsets(int(n), _p(out))
    return out


def fibonacci_dirs32(ox, oy):
    out = np.empty((32, 3), np.float32)
    load().orc_fibonacci_dirs32(float(ox), float(oy), _p(out))
    return out


# ---------------------------------------------------------------- brute force
def udf_brute(V, F, pts, return_uvw=True, nthreads=0):
    V, F, pts = _f32(V), _i32(F), _f32(pts).reshape(-1, 3)
    n = len(pts)
    d = np.empty(n, np.float32)
    f = np.empty(n, np.int64)
    uvw = np.empty((n, 3), np.float32) if return_uvw else None
    rc = load().orc_udf_brute(_p(V), len(V), _p(F), len(F), _p(pts), n, _p(d), _p(f), _p(uvw), nthreads)
    assert rc == 0
    return d, f, uvw


def ray_brute(V, F, ro, rd, nthreads=0):
    V, F, ro, rd = _f32(V), _i32(F), _f32(ro).reshape(-1, 3), _f32(rd).reshape(-1, 3)
    n = len(ro)
    pos = np.empty((n, 3), np.float32)
    f = np.empty(n, np.int64)
    t = np.empty(n, np.float32)
    rc = load().orc_ray_brute(_p(V), len(V), _p(F), len(F), _p(ro), _p(rd), n, _p(pos), _p(f), _p(t), nthreads)
    assert rc == 0
    return pos, f, t


def sdf_brute(V, F, pts, mode, return_uvw=True, nthreads=0):
    V, F, pts = _f32(V), _i32(F), _f32(pts).reshape(-1, 3)
    n = len(pts)
    d = np.empty(n, np.float32)
    f = np.empty(n, np.int64)
    uvw = np.empty((n, 3), np.float32) if return_uvw else None
    rc = load().orc_sdf_brute(_p(V), len(V), _p(F), len(F), _p(pts), n, int(mode), _p(d), _p(f), _p(uvw), nthreads)
    assert rc == 0
    return d, f, uvw


def udf_f64(V, F, pts, nthreads=0):
    V, F, pts = _f32(V), _i32(F), _f32(pts).reshape(-1, 3)
    n = len(pts)
    d = np.empty(n, np.float64)
    f = np.empty(n, np.int64)
    cp = np.empty((n, 3), np.float64)
    gap = np.empty(n, np.float64)
    rc = load().orc_udf_f64(_p(V), len(V), _p(F), len(F), _p(pts), n, _p(d), _p(f), _p(cp), _p(gap), nthreads)
    assert rc == 0
    return d, f, cp, gap


def ray_f64(V, F, ro, rd, tol=1e-6, nthreads=0):
    V, F, ro, rd = _f32(V), _i32(F), _f32(ro).reshape(-1, 3), _f32(rd).reshape(-1, 3)
    n = len(ro)
    t = np.empty(n, np.float64)
    f = np.empty(n, np.int64)
    amb = np.empty(n, np.uint8)
    rc = load().orc_ray_f64(_p(V), len(V), _p(F), len(F), _p(ro), _p(rd), n, float(tol), _p(t), _p(f), _p(amb),
                            nthreads)
    assert rc == 0
    return t, f, amb.astype(bool)


def udf_backward(V, F, pts, g, face, uvw):
    """Gradients of sum_i g_i d_i (cubvh_oracle.cpp (9)): grad_points [n,3] and
    grad_vertices [V,3] in fp32, with the product's operations and summation order."""
    V, F, pts = _f32(V), _i32(F), _f32(pts).reshape(-1, 3)
    g = _f32(g).reshape(-1)
    face = np.ascontiguousarray(face, dtype=np.int64).reshape(-1)
    uvw = _f32(uvw).reshape(-1, 3)
    n = len(pts)
    gp = np.empty((n, 3), np.float32)
    gv = np.zeros((len(V), 3), np.float32)
    rc = load().orc_udf_backward(_p(V), len(V), _p(F), len(F), _p(pts), _p(g), _p(face), _p(uvw), n, _p(gp), _p(gv))
    assert rc == 0
    return gp, gv


def winding_f64(V, F, pts, nthreads=0):
    V, F, pts = _f32(V), _i32(F), _f32(pts).reshape(-1, 3)
    w = np.empty(len(pts), np.float64)
    rc = load().orc_winding_f64(_p(V), len(V), _p(F), len(F), _p(pts), len(pts), _p(w), nthreads)
    assert rc == 0
    return w


# ---------------------------------------------------------------- reference BVH
class RefBvh:
    """The reference's BVH4 (src/bvh.cu:515-620) and its traversals."""

    def __init__(self, V, F):
        self.V, self.F = _f32(V), _i32(F)
        self.h = load().orc_refbvh_build(_p(self.V), len(self.V), _p(self.F), len(self.F))
        if not self.h:
            raise ValueError("bad mesh")

    @classmethod
    def from_arrays(cls, nodes9, tris12):
        """A tree in the reference's state_dict layout (nodes int32 [n, 9], triangles
        int32 [M, 12]), walked by the restated reference traversals."""
        self = cls.__new__(cls)
        nodes9 = np.ascontiguousarray(nodes9, np.int32)
        tris12 = np.ascontiguousarray(tris12, np.int32)
        self.F = np.zeros((len(tris12), 3), np.int32)  # only its length is used
        self.h = load().orc_refbvh_import(_p(nodes9), len(nodes9), _p(tris12), len(tris12))
        if not self.h:
            raise ValueError("empty tree")
        return self

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "h", None) and load is not None:
            load().orc_refbvh_free(self.h)
            self.h = None

    @property
    def n_nodes(self):
        return load().orc_refbvh_num_nodes(self.h)

    def export(self):
        """Reference state_dict arrays: nodes int32 [n, 9], triangles int32 [M, 12]."""
        nodes = np.empty((self.n_nodes, 9), np.int32)
        tris = np.empty((len(self.F), 12), np.int32)
        load().orc_refbvh_export(self.h, _p(nodes), _p(tris))
        return nodes, tris

    def udf(self, pts, return_uvw=True, tie_mode=0, nthreads=0):
        """tie_mode 0: the reference's rule; 1: canonical ties; +2: escape links only."""
        pts = _f32(pts).reshape(-1, 3)
        n = len(pts)
        d = np.empty(n, np.float32)
        f = np.empty(n, np.int64)
        uvw = np.empty((n, 3), np.float32) if return_uvw else None
        counts = np.zeros(2, np.int64)
        load().orc_refbvh_udf(self.h, _p(pts), n, _p(d), _p(f), _p(uvw), tie_mode, _p(counts), nthreads)
        return d, f, uvw, counts

    def ray(self, ro, rd, tie_mode=0, nthreads=0):
        ro, rd = _f32(ro).reshape(-1, 3), _f32(rd).reshape(-1, 3)
        n = len(ro)
        pos = np.empty((n, 3), np.float32)
        f = np.empty(n, np.int64)
        t = np.empty(n, np.float32)
        counts = np.zeros(2, np.int64)
        load().orc_refbvh_ray(self.h, _p(ro), _p(rd), n, _p(pos), _p(f), _p(t), tie_mode, _p(counts), nthreads)
        return pos, f, t, counts

    def sdf(self, pts, mode, return_uvw=True, nthreads=0):
        pts = _f32(pts).reshape(-1, 3)
        n = len(pts)
        d = np.empty(n, np.float32)
        f = np.empty(n, np.int64)
        uvw = np.empty((n, 3), np.float32) if return_uvw else None
        counts = np.zeros(4, np.int64)
        load().orc_refbvh_sdf(self.h, _p(pts), n, int(mode), _p(d), _p(f), _p(uvw), _p(counts), nthreads)
        return d, f, uvw, counts

    def sdf_idx(self, pts, mode, qidx=None, rule=1, return_uvw=True, nthreads=0):
        """Signed distance of pts (mode 0 watertight, else raystab) where query i has
        raystab RNG index qidx[i] within the caller's whole call (None: 0..n-1).
        rule 1: canonical (the product's bit-exact target); rule 0: the reference's
        own rules (safe_divide intersector, last-visited ties, fp32 normal sum)."""
        pts = _f32(pts).reshape(-1, 3)
        n = len(pts)
        q = None if qidx is None else np.ascontiguousarray(qidx, dtype=np.int64)
        if q is not None and q.shape != (n,):
            raise ValueError("qidx must have one entry per point")
        d = np.empty(n, np.float32)
        f = np.empty(n, np.int64)
        uvw = np.empty((n, 3), np.float32) if return_uvw else None
        load().orc_refbvh_sdf_idx(self.h, _p(pts), _p(q), n, int(mode), int(rule), _p(d), _p(f), _p(uvw), nthreads)
        return d, f, uvw


# ---------------------------------------------------------------- floodfill
def floodfill_ref(grid):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    if g.ndim == 3:
        g = g[None]
    B, H, W, D = g.shape
    out = np.empty(g.shape, np.int32)
    it = load().orc_floodfill_ref(_p(g), B, H, W, D, _p(out))
    return out, it


def floodfill_bfs(grid):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    if g.ndim == 3:
        g = g[None]
    B, H, W, D = g.shape
    out = np.empty(g.shape, np.int32)
    load().orc_floodfill_bfs(_p(g), B, H, W, D, _p(out))
    return out


# ---------------------------------------------------------------- hash table
def hash_build(coords, capacity=None):
    """The reference's sequential table over coords int32 [n, d]: (marker, row) int32
    [capacity, 2] and the number of rows that found no slot.  capacity None: the
    reference's build() size max(16, 2n)."""
    c = _i32(coords)
    n, d = c.shape
    cap = max(16, 2 * n) if capacity is None else int(capacity)
    table = np.empty((cap, 2), np.int32)
    dropped = load().orc_hash_build(_p(c), n, d, cap, _p(table))
    return table, int(dropped)


def hash_search(table, coords, queries):
    """int32 [m]: the first stored row equal to each query along its probe run, or -1."""
    c = _i32(coords)
    q = _i32(queries)
    t = np.ascontiguousarray(table, dtype=np.int32)
    out = np.empty(q.shape[0], np.int32)
    load().orc_hash_search(_p(t), t.shape[0], _p(c), q.shape[1], _p(q), q.shape[0], _p(out))
    return out


def hash_lookup(coords, queries, capacity=None):
    """build + search in one call."""
    table, _ = hash_build(coords, capacity)
    return hash_search(table, coords, queries)


_ref_hash = None


def ref_hash_available() -> bool:
    return REF_HASH_LIB.exists()


def ref_hash_build_search(coords, queries, capacity=None):
    """The reference's own CPU table (include/cpu/hashtable.h, built into _ref/):
    (search results int32 [m], table int32 [capacity, 2])."""
    global _ref_hash
    if _ref_hash is None:
        lib = ctypes.CDLL(str(REF_HASH_LIB))
        lib.ref_hash_build_search.restype = c_int
        lib.ref_hash_build_search.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]
        _ref_hash = lib
    c = _i32(coords)
    q = _i32(queries)
    n, d = c.shape
    cap = max(16, 2 * n) if capacity is None else int(capacity)
    table = np.empty((cap, 2), np.int32)
    out = np.empty(q.shape[0], np.int32)
    got = _ref_hash.ref_hash_build_search(_p(c), n, d, 0 if capacity is None else cap, _p(q), q.shape[0], _p(out),
                                          _p(table))
    assert got == cap
    return out, table


# ---------------------------------------------------------------- sparse marching cubes
def smc(coords, corners, iso, consistent=False):
    """The reference GPU extraction's output, restated: (vertices f32 [M, 3], triangles
    int32 [T, 3])."""
    c = _i32(coords).reshape(-1, 3)
    v = _f32(corners).reshape(-1, 8)
    n = c.shape[0]
    verts = np.empty((12 * n + 1, 3), np.float32)
    tris = np.empty((5 * n + 1, 3), np.int32)
    nv = c_int64()
    nt = c_int64()
    rc = load().orc_smc(_p(c), _p(v), n, float(iso), int(bool(consistent)), _p(verts), _p(tris), ctypes.byref(nv),
                        ctypes.byref(nt))
    if rc != 0:
        raise RuntimeError("case table uses an edge without a sign change")
    return verts[: nv.value].copy(), tris[: nt.value].copy()


_ref_smc = None


def ref_smc_available() -> bool:
    return REF_SMC_LIB.exists()


def _ref_smc_lib():
    global _ref_smc
    if _ref_smc is None:
        lib = ctypes.CDLL(str(REF_SMC_LIB))
        lib.ref_mc_tables.restype = None
        lib.ref_mc_tables.argtypes = [c_void_p, c_void_p]
        lib.ref_smc.restype = None
        lib.ref_smc.argtypes = [c_void_p, c_void_p, c_int, c_float, c_int, c_void_p, c_int64, c_void_p, c_int64,
                                c_void_p, c_void_p]
        _ref_smc = lib
    return _ref_smc


def ref_mc_tables():
    """The reference's case tables (include/cpu/spmc.h): edge masks int32 [256] and
    triangle lists int32 [256, 16]."""
    edges = np.empty(256, np.int32)
    tris = np.empty((256, 16), np.int32)
    _ref_smc_lib().ref_mc_tables(_p(edges), _p(tris))
    return edges, tris


def ref_smc_cpu(coords, corners, iso, consistent=False):
    """The reference's own CPU extraction (include/cpu/spmc.h, built into _ref/):
    (vertices in first-use order, triangles)."""
    c = _i32(coords).reshape(-1, 3)
    v = _f32(corners).reshape(-1, 8)
    n = c.shape[0]
    verts = np.empty((12 * n + 1, 3), np.float32)
    tris = np.empty((5 * n + 1, 3), np.int32)
    nv = c_int64()
    nt = c_int64()
    _ref_smc_lib().ref_smc(_p(c), _p(v), n, float(iso), int(bool(consistent)), _p(verts), len(verts), _p(tris),
                           len(tris), ctypes.byref(nv), ctypes.byref(nt))
    return verts[: nv.value].copy(), tris[: nt.value].copy()


# ---------------------------------------------------------------- reference PCG32
_ref_pcg = None


def ref_pcg32_available() -> bool:
    return REF_PCG_LIB.exists()


def _ref_pcg_lib():
    global _ref_pcg
    if _ref_pcg is None:
        lib = ctypes.CDLL(str(REF_PCG_LIB))
        lib.ref_pcg32_uints.restype = None
        lib.ref_pcg32_uints.argtypes = [c_int64, c_int64, c_void_p]
        lib.ref_pcg32_offsets.restype = None
        lib.ref_pcg32_offsets.argtypes = [c_int64, c_void_p]
        lib.ref_pcg32_state.restype = ctypes.c_uint64
        lib.ref_pcg32_state.argtypes = [c_int64]
        _ref_pcg = lib
    return _ref_pcg


def ref_pcg32_uints(advance, n):
    """next_uint() outputs of the reference's own generator (include/gpu/pcg32.h, built
    into _ref/), default state, advanced by `advance` first."""
    out = np.empty(n, np.uint32)
    _ref_pcg_lib().ref_pcg32_uints(int(advance), int(n), _p(out))
    return out


def ref_pcg32_offsets(n):
    """The reference generator's raystab offsets of queries 0..n-1: float32 [n, 2]."""
    out = np.empty((n, 2), np.float32)
    _ref_pcg_lib().ref_pcg32_offsets(int(n), _p(out))
    return out


def ref_pcg32_state(advance):
    """The reference generator's internal 64-bit state after advance(advance)."""
    return int(_ref_pcg_lib().ref_pcg32_state(int(advance)))
