"""The C ABI library without a GPU: symbols, host build, serialization, errors.

Nothing here launches a kernel; construction and state_dict are host-only, like the
reference's lazy GPU initialisation (test/lazy_gpu_initialization.py:94-105).
"""
import ctypes
import re

import numpy as np
import pytest
import torch

from conftest import ROOT


def _declared_symbols():
    text = (ROOT / "include" / "cubvh_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cubvh_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from cubvh_amd import _lib

    lib = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"
    assert lib.cubvh_abi_version() >> 16 == 1


def test_create_info_export_import_roundtrip(ico4):
    import cubvh_amd

    v, f = ico4
    bvh = cubvh_amd.cuBVH(v, f)
    info = bvh.info()
    assert info["n_faces"] == len(f)
    assert info["stack_need"] == 3 * info["depth"] + 1 <= 64
    sd = bvh.state_dict()
    assert sd["nodes"].dtype == torch.int32 and sd["nodes"].shape == (info["n_nodes"], 32)
    assert sd["triangles"].shape == (len(f), 12)
    # triangle rows carry the 48-byte record: a, b, c, 4 zero bytes, int64 face id
    rec = sd["triangles"].numpy()
    ids = rec.view(np.int64)[:, 5]
    assert sorted(ids.tolist()) == list(range(len(f)))
    assert (rec[:, 9] == 0).all()
    abc = rec[:, :9].view(np.float32).reshape(-1, 3, 3)
    np.testing.assert_array_equal(abc, v[f[ids]])
    b2 = cubvh_amd.cuBVH.from_state_dict(sd)
    sd2 = b2.state_dict()
    for k in sd:
        assert torch.equal(sd[k], sd2[k])


def test_state_dict_survives_torch_save(tmp_path, ico4):
    import cubvh_amd

    v, f = ico4
    sd = cubvh_amd.cuBVH(v, f).state_dict()
    torch.save(sd, tmp_path / "bvh.pt")
    sd2 = torch.load(tmp_path / "bvh.pt", weights_only=True)
    b = cubvh_amd.cuBVH.from_state_dict(sd2)
    assert b.info()["n_faces"] == len(f)


def test_import_reference_state_dict(oracle, ico4):
    """A state_dict in the reference's layout (nodes [n, 9], triangles [M, 12]) is
    accepted; the mesh recovered from it equals the original faces."""
    import cubvh_amd

    v, f = ico4
    nodes9, tris12 = oracle.RefBvh(v, f).export()
    sd = {"nodes": torch.from_numpy(nodes9), "triangles": torch.from_numpy(tris12)}
    b = cubvh_amd.cuBVH.from_state_dict(sd)
    tv = b._vertices.reshape(-1, 3, 3)
    np.testing.assert_array_equal(tv, v[f])


def test_reference_state_dict_validation(oracle, ico4):
    import cubvh_amd
    from cubvh_amd._lib import CubvhError

    v, f = ico4
    nodes9, tris12 = oracle.RefBvh(v, f).export()
    bad = tris12.copy()
    bad.view(np.int64)[3, 5] = bad.view(np.int64)[4, 5]  # duplicate face id
    with pytest.raises(CubvhError, match="permutation"):
        cubvh_amd.cuBVH.from_state_dict({"nodes": torch.from_numpy(nodes9), "triangles": torch.from_numpy(bad)})
    badn = nodes9.copy()
    badn[0, 7] = 10**6  # child range past the end
    with pytest.raises(CubvhError, match="child range"):
        cubvh_amd.cuBVH.from_state_dict({"nodes": torch.from_numpy(badn), "triangles": torch.from_numpy(tris12)})
    with pytest.raises(RuntimeError, match="missing"):
        cubvh_amd.cuBVH.from_state_dict({"nodes": torch.from_numpy(nodes9)})


def test_own_state_dict_validation(ico4):
    import cubvh_amd
    from cubvh_amd._lib import CubvhError

    v, f = ico4
    sd = cubvh_amd.cuBVH(v, f).state_dict()
    nodes = sd["nodes"].clone()
    nodes[0, 24] = 10**7  # child reference out of range
    with pytest.raises(CubvhError, match="invalid state"):
        cubvh_amd.cuBVH.from_state_dict({"nodes": nodes, "triangles": sd["triangles"]})


def test_imported_empty_slots_get_empty_boxes(ico4):
    """An imported node whose empty child slot carries a box (finite or NaN) has that box
    reset to +inf .. -inf: the walks never reach an empty slot through its box."""
    import cubvh_amd

    v, f = ico4
    sd = cubvh_amd.cuBVH(v, f).state_dict()
    nodes = sd["nodes"].clone()
    child = nodes[:, 24:28]
    empty = (child == -1).nonzero()
    assert len(empty) > 0  # the tree has nodes with fewer than 4 children
    boxes = nodes[:, :24].view(torch.float32).view(-1, 6, 4)  # lo_x, hi_x, lo_y, hi_y, lo_z, hi_z per slot
    ref = boxes.clone()
    for n, k in empty.tolist()[:8]:
        boxes[n, :, k] = torch.tensor([0.0, 0.0, float("nan"), 1.0, -1.0, 1.0])
    out = cubvh_amd.cuBVH.from_state_dict({"nodes": nodes, "triangles": sd["triangles"]}).state_dict()["nodes"]
    got = out[:, :24].view(torch.float32).view(-1, 6, 4)
    assert torch.equal(got, ref)
    e = child == -1
    assert torch.isinf(got[:, 0::2][e.unsqueeze(1).expand(-1, 3, -1)]).all()


def test_bad_meshes_raise(ico4):
    import cubvh_amd

    v, f = ico4
    with pytest.raises(AssertionError, match="at least 8"):
        cubvh_amd.cuBVH(v, f[:8])
    bad = f.copy()
    bad[10, 1] = len(v)
    with pytest.raises(ValueError, match="outside"):
        cubvh_amd.cuBVH(v, bad)
    with pytest.raises(ValueError):
        cubvh_amd.cuBVH(v[:, :2], f)


def test_c_abi_rejects_bad_face_index_with_message():
    from cubvh_amd import _lib

    lib = _lib.load()
    v = np.zeros((4, 3), np.float32)
    f = np.array([[0, 1, 5]] * 9, np.int32)
    h = ctypes.c_void_p()
    rc = lib.cubvh_bvh_create(v.ctypes.data, 4, f.ctypes.data, 9, 0, ctypes.byref(h))
    assert rc == -1
    assert b"references vertex 5" in lib.cubvh_last_error()
    rc = lib.cubvh_bvh_create(v.ctypes.data, 4, f.ctypes.data, 9, 9, ctypes.byref(h))
    assert rc == -1
    # non-finite coordinates are refused
    v[0, 0] = np.nan
    f = np.array([[0, 1, 2]] * 9, np.int32)
    rc = lib.cubvh_bvh_create(v.ctypes.data, 4, f.ctypes.data, 9, 0, ctypes.byref(h))
    assert rc == -1 and b"finite" in lib.cubvh_last_error()


def test_float64_and_int64_inputs_are_cast(ico4):
    """The reference's caster rounds geometry to fp32 (patch/eigen/matrix.h:490-524)."""
    import cubvh_amd

    v, f = ico4
    a = cubvh_amd.cuBVH(v.astype(np.float64), f.astype(np.int64)).state_dict()
    b = cubvh_amd.cuBVH(torch.from_numpy(v), torch.from_numpy(f.astype(np.int64))).state_dict()
    for k in a:
        assert torch.equal(a[k], b[k])


def test_tree_properties_large_mesh():
    """Level-6 icosphere: every face in exactly one leaf, boxes contain their faces."""
    import cubvh_amd
    from cubvh_amd import synthetic as S

    v, f = S.jittered_sphere(6)
    bvh = cubvh_amd.cuBVH(v, f)
    sd = bvh.state_dict()
    nodes = sd["nodes"].numpy()
    rec = sd["triangles"].numpy()
    abc = rec[:, :9].view(np.float32).reshape(-1, 3, 3)
    bounds = nodes[:, :24].view(np.float32).reshape(-1, 6, 4)  # lo_x, hi_x, lo_y, hi_y, lo_z, hi_z
    child = nodes[:, 24:28]
    seen = np.zeros(len(f), np.int64)
    for ni in range(len(nodes)):
        for k in range(4):
            r = child[ni, k]
            if r == -1 or r >= 0:
                continue
            u = np.uint32(r & 0x7FFFFFFF)
            first, cnt = int(u >> 4), int(u & 15) + 1
            seen[first:first + cnt] += 1
            tri = abc[first:first + cnt]
            lo = bounds[ni, 0::2, k]
            hi = bounds[ni, 1::2, k]
            assert (tri.min(axis=(0, 1)) >= lo).all() and (tri.max(axis=(0, 1)) <= hi).all()
    assert (seen == 1).all()


def test_reference_package_names_exist():
    import cubvh_amd

    for name in ("cuBVH", "HashTable", "cuHashTable", "sparse_marching_cubes", "sparse_marching_cubes_cpu",
                 "floodfill", "fill_holes", "merge_vertices", "decimate", "parallel_decimate"):
        assert hasattr(cubvh_amd, name)
    with pytest.raises(NotImplementedError):
        cubvh_amd.fill_holes(None, None)
    with pytest.raises(NotImplementedError):
        cubvh_amd.HashTable()


def test_hashtable_host_side_checks():
    """cuHashTable's host logic: dims, argument checks (src/api_gpu.cu:239-274), and a
    search before any build fails; nothing here touches a GPU."""
    import cubvh_amd

    ht = cubvh_amd.cuHashTable()
    assert ht.num_dims == 3
    ht.impl.set_num_dims(5)
    assert ht.num_dims == 5 and ht.impl.capacity == 0
    with pytest.raises(cubvh_amd._lib.CubvhError, match="1..63"):
        ht.impl.set_num_dims(0)
    with pytest.raises(cubvh_amd._lib.CubvhError, match="1..63"):
        cubvh_amd.cuHashTable(num_dims=64)
    cpu = torch.zeros(4, 3, dtype=torch.int32)
    with pytest.raises(RuntimeError, match="reside on CUDA"):
        ht.impl.build(cpu)
    with pytest.raises(RuntimeError, match="reside on CUDA"):
        ht.impl.search(cpu)
    lib = cubvh_amd._lib.load()
    h = ctypes.c_void_p()
    cubvh_amd._lib.check(lib.cubvh_hashtable_create(3, ctypes.byref(h)))
    try:
        assert lib.cubvh_hashtable_search(h, 0, None, 4, 3, None, None) == -1
        assert "not built" in lib.cubvh_last_error().decode()
        assert lib.cubvh_hashtable_resize(h, 0, -1, None) == -1
        assert lib.cubvh_hashtable_insert(h, 0, None, -1, 3, None) == -1
        # capacity 0: insert and prepare do nothing, as in the reference
        assert lib.cubvh_hashtable_insert(h, 0, None, 10, 3, None) == 0
        assert lib.cubvh_hashtable_prepare(h, None) == 0
        assert lib.cubvh_hashtable_build(h, 0, None, 1 << 30, 3, None) == -1
    finally:
        lib.cubvh_hashtable_destroy(h)


def test_floodfill_requires_bool_grid():
    import cubvh_amd

    with pytest.raises(TypeError, match="bool"):
        cubvh_amd.floodfill(torch.zeros(4, 4, 4, dtype=torch.uint8))


def test_forest_host_concatenation():
    import cubvh_amd
    from cubvh_amd import synthetic as S

    bs = [cubvh_amd.cuBVH(*S.icosphere(l)) for l in (2, 3, 1)]
    f = cubvh_amd.cuBVHForest(bs)
    info = f.info()
    assert info["n_meshes"] == 3
    assert info["n_faces"] == sum(b.info()["n_faces"] for b in bs)
    assert info["n_nodes"] == sum(b.info()["n_nodes"] for b in bs)
    assert info["stack_need"] == max(b.info()["stack_need"] for b in bs)
    vcat, fcat = f.concatenated_mesh("cpu")
    assert vcat.shape[0] == sum(b._vertices.shape[0] for b in bs)
    assert fcat.shape[0] == info["n_faces"] and int(fcat.max()) == vcat.shape[0] - 1
    assert f.face_base().tolist() == [0, 320, 1600]


def test_forest_offsets_checked_on_the_host():
    """The offsets must be a CSR of the queries: offsets[0] == 0, non-decreasing,
    offsets[-1] == n; other offsets are refused before anything reaches the device."""
    from cubvh_amd.forest import _offsets_tensor

    assert _offsets_tensor([0, 3, 3, 7], 7, "cpu").tolist() == [0, 3, 3, 7]
    for bad, msg in (([1, 3, 7], "offsets\\[0\\]"), ([0, 3, 6], "offsets\\[-1\\]"), ([0, 5, 3, 7], "non-decreasing"),
                     ([0], "n_meshes \\+ 1"), ([[0, 7]], "n_meshes \\+ 1"), ([0.0, 7.0], "integers")):
        with pytest.raises(ValueError, match=msg):
            _offsets_tensor(torch.tensor(bad), 7, "cpu")
    assert _offsets_tensor(torch.tensor([5, 1]), 7, "cpu", check=False).tolist() == [5, 1]


def test_forest_morton_bins_limit_refused_before_device_work():
    """A forest call's Morton bins (n_meshes << order bits) must stay within 2^30: the
    bin index is 32-bit and the histogram holds a word per bin.  16,385 meshes with 16
    order bits exceed it and are refused by both entry points before any device work."""
    import cubvh_amd
    from cubvh_amd import _lib
    from cubvh_amd import synthetic as S

    lib = _lib.load()
    v, f = S.icosphere(0)
    b = cubvh_amd.cuBVH(v, f)
    n_meshes = (1 << 14) + 1
    arr = (ctypes.c_void_p * n_meshes)(*([b._h.value] * n_meshes))
    h = ctypes.c_void_p()
    _lib.check(lib.cubvh_forest_create(arr, n_meshes, ctypes.byref(h)))
    try:
        assert lib.cubvh_forest_reserve(h, 0, 1000, 16, None) == -1
        assert b"Morton bins" in lib.cubvh_last_error()
        tun = (ctypes.c_int32 * 4)(16, 1, 8, 4)
        assert lib.cubvh_forest_unsigned_distance(h, 0, 1000, 1, 1, 1, 1, None, tun, 4, None) == -1
        assert b"Morton bins" in lib.cubvh_last_error()
        # the tuning array's length is checked, and so is the workgroup shape
        assert lib.cubvh_forest_unsigned_distance(h, 0, 1000, 1, 1, 1, 1, None, tun, 6, None) == -1
        assert b"0 to 5 entries" in lib.cubvh_last_error()
        bad_xcd = (ctypes.c_int32 * 5)(0, 0, 8, 4, -1)
        assert lib.cubvh_forest_unsigned_distance(h, 0, 1000, 1, 1, 1, 1, None, bad_xcd, 5, None) == -1
        assert b"XCD chunk" in lib.cubvh_last_error()
        assert lib.cubvh_forest_unsigned_distance(h, 0, 1000, 1, 1, 1, 1, None, None, 2, None) == -1
        bad = (ctypes.c_int32 * 4)(0, 1, 8, 3)
        assert lib.cubvh_forest_unsigned_distance(h, 0, 1000, 1, 1, 1, 1, None, bad, 4, None) == -1
        assert b"waves per block" in lib.cubvh_last_error()
    finally:
        lib.cubvh_forest_destroy(h)


def test_build_options_and_depth_fallback():
    """Leaf size and split rule flags; a mesh whose SAH tree would be too deep for
    the traversal stack (geometric cluster sizes) still builds (median fallback)."""
    import cubvh_amd
    from cubvh_amd import synthetic as S

    v, f = S.icosphere(4)
    for flags, leaf in ((0, 16), (8, 8), (4, 4), (0x104, 4), (0x101, 1), (16, 16)):
        info = cubvh_amd.cuBVH(v, f, leaf_size=flags).info()
        assert info["leaf_size"] <= leaf and info["stack_need"] <= 64
    # triangles at geometrically shrinking scales: very unbalanced SAH splits
    tris = []
    for k in range(3000):
        s = 0.7 ** (k % 60)
        c = np.array([s, 0.0, 0.0])
        tris.append([c, c + [s * 0.1, 0, 0], c + [0, s * 0.1, 0]])
    V = np.asarray(tris, np.float32).reshape(-1, 3)
    F = np.arange(len(V), dtype=np.int32).reshape(-1, 3)
    info = cubvh_amd.cuBVH(V, F).info()
    assert info["stack_need"] <= 64


def test_host_trees_reference_work(oracle):
    """Tree quality as the reference's walk counts it: at the reference's leaf size (8)
    the host SAH tree, exported to the reference layout and walked by the oracle's
    restated reference traversal (src/bvh.cu:300-353), takes fewer box and triangle
    tests per query than the reference's own build (src/bvh.cu:515-620) on a level-6
    sphere; the object-median tree does not beat the SAH tree; every tree gives the
    same distances."""
    import cubvh_amd
    from cubvh_amd import synthetic as S

    v, f, _ = S.perturbed_sphere(6, seed=5)
    pts = S.uniform_points(1 << 14, seed=13)
    ref = oracle.RefBvh(v, f)
    d_ref, _, _, c = ref.udf(pts, return_uvw=False, tie_mode=0)
    cost = {"reference": 24 * c[0] + 36 * c[1]}
    for name, flags in (("sah", 8), ("median", 0x108)):
        sd = cubvh_amd.cuBVH(v, f, leaf_size=flags).state_dict(layout="reference")
        tree = oracle.RefBvh.from_arrays(sd["nodes"].numpy(), sd["triangles"].numpy())
        d, _, _, c = tree.udf(pts, return_uvw=False, tie_mode=0)
        np.testing.assert_array_equal(d, d_ref)
        cost[name] = 24 * c[0] + 36 * c[1]
    assert cost["sah"] < cost["reference"], cost
    assert cost["sah"] <= cost["median"], cost


def test_build_many_rejects_unknown_build():
    import cubvh_amd

    with pytest.raises(ValueError, match="gpu_sah"):
        cubvh_amd.cuBVH.build_many([], build="host")
    with pytest.raises(ValueError, match="gpu_lbvh"):
        cubvh_amd.cuBVH([[0, 0, 0]] * 30, [[0, 1, 2]] * 10, build="cpu")


def test_gpu_build_flags_checked_before_device_work():
    """cubvh_bvh_create_gpu: leaf_size bits 8 (linear BVH) and 9 (SAH) exclude each
    other, and higher bits are unknown; both are refused before any device access."""
    import ctypes

    from cubvh_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    for flags, msg in ((0x300, b"exclude"), (0x400, b"unknown bits"), (0x211, b"0..16")):
        assert lib.cubvh_bvh_create_gpu(0, 1, 3, 1, 100, flags, ctypes.byref(h), None) == -1
        assert msg in lib.cubvh_last_error(), lib.cubvh_last_error()
        assert not h.value


def test_ordered_entry_points_validate_arguments(ico4):
    """Argument errors are reported before any device work."""
    import cubvh_amd
    from cubvh_amd import _lib

    lib = _lib.load()
    v, f = ico4
    b = cubvh_amd.cuBVH(v, f)
    assert lib.cubvh_distance_ordered(b._h, 0, 10, 2, 1, 1, 1, 1, None, None) == -1
    assert b"kind" in lib.cubvh_last_error()
    assert lib.cubvh_distance_ordered(b._h, 0, -1, 0, 1, 1, 1, 1, None, None) == -1
    assert lib.cubvh_morton_order(0, 1 << 31, 1, 1, None) == -1
    assert b"at most" in lib.cubvh_last_error()
    assert lib.cubvh_morton_order(0, 0, None, None, None) == 0


def test_export_reference_layout_walked_by_reference_traversal(oracle):
    """state_dict(layout="reference") is a valid tree for the reference's own walks:
    the oracle's restatement of closest_triangle / ray_intersect (stack walk and the
    escape-link walk) over the exported arrays gives the brute-force answers."""
    import cubvh_amd
    from cubvh_amd import synthetic as S

    v, f, _ = S.perturbed_sphere(4, seed=7)
    for leaf in (0, 3):
        b = cubvh_amd.cuBVH(v, f, leaf_size=leaf)
        sd = b.state_dict(layout="reference")
        nodes, tris = sd["nodes"].numpy(), sd["triangles"].numpy()
        assert nodes.shape == (1 + 4 * b.info()["n_nodes"], 9) and tris.shape == (len(f), 12)
        R = oracle.RefBvh.from_arrays(nodes, tris)
        p = S.uniform_points(3000, seed=leaf)
        bd, bf, bu = oracle.udf_brute(v, f, p)
        for mode in (1, 3):  # canonical ties; stack walk, then escape links only
            d, fi, u, _ = R.udf(p, tie_mode=mode)
            np.testing.assert_array_equal(d, bd)
            np.testing.assert_array_equal(fi, bf)
        ro, rd = S.sphere_rays(3000, seed=leaf)
        bp, bfr, bt = oracle.ray_brute(v, f, ro, rd)
        for mode in (1, 3):
            pos, fr, t, _ = R.ray(ro, rd, tie_mode=mode)
            np.testing.assert_array_equal(t, bt)
            np.testing.assert_array_equal(fr, bfr)
        # and it imports back (the reference layout importer rebuilds this layout)
        b2 = cubvh_amd.cuBVH.from_state_dict(sd)
        np.testing.assert_array_equal(b2._vertices.reshape(-1, 3, 3), v[f])


def test_valu_probe_argument_checks():
    """cubvh_valu_probe refuses null outputs, forms other than 0 and 1, iteration counts
    outside [1, 2^20] and bad device ordinals before it touches a device."""
    from cubvh_amd import _lib

    lib = _lib.load()
    wi = ctypes.c_double(0.0)
    ms = ctypes.c_float(0.0)
    assert lib.cubvh_valu_probe(0, 0, 100, None, ctypes.byref(ms)) == -1
    assert b"null output" in lib.cubvh_last_error()
    for form in (-1, 2):
        assert lib.cubvh_valu_probe(0, form, 100, ctypes.byref(wi), ctypes.byref(ms)) == -1
        assert b"form" in lib.cubvh_last_error()
    for iters in (0, -5, (1 << 20) + 1):
        assert lib.cubvh_valu_probe(0, 0, iters, ctypes.byref(wi), ctypes.byref(ms)) == -1
        assert b"iters" in lib.cubvh_last_error()
    assert lib.cubvh_valu_probe(-1, 0, 100, ctypes.byref(wi), ctypes.byref(ms)) == -1
    assert b"device ordinal" in lib.cubvh_last_error()


def test_vertex_corner_entry_points_argument_checks():
    """cubvh_vertex_corners and cubvh_udf_backward_with_corners refuse null buffers and
    empty meshes before they touch a device."""
    from cubvh_amd import _lib

    lib = _lib.load()
    buf = ctypes.c_void_p(1)
    assert lib.cubvh_vertex_corners(0, None, 10, 10, buf, buf, buf, None) == -1
    assert b"null buffer" in lib.cubvh_last_error()
    assert lib.cubvh_vertex_corners(0, buf, 0, 10, buf, buf, buf, None) == -1
    assert b"empty mesh" in lib.cubvh_last_error()
    assert lib.cubvh_vertex_corners(0, buf, 1 << 30, 10, buf, buf, buf, None) == -1
    assert b"too large" in lib.cubvh_last_error()
    assert lib.cubvh_udf_backward_with_corners(0, 4, buf, buf, buf, buf, buf, 3, buf, 1, None, buf, None, buf, buf,
                                               None) == -1
    assert b"corner table" in lib.cubvh_last_error()
