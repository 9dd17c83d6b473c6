"""Tree built on the GPU (cubvh_bvh_create_gpu): structure, and query results
identical to the host-built tree and to the oracle.

Query results do not depend on the tree (canonical ties to the lowest face index and
conservative box tests), so a GPU-built tree must give bit-identical answers.
"""
import ctypes

import numpy as np
import pytest
import torch

import cubvh_amd
from cubvh_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _leaves_cover_faces(bvh, v, f):
    sd = bvh.state_dict()
    nodes = sd["nodes"].numpy()
    rec = sd["triangles"].numpy()
    abc = rec[:, :9].view(np.float32).reshape(-1, 3, 3)
    ids = rec.view(np.int64)[:, 5]
    assert sorted(ids.tolist()) == list(range(len(f)))
    np.testing.assert_array_equal(abc, v[f[ids]])
    bounds = nodes[:, :24].view(np.float32).reshape(-1, 6, 4)
    seen = np.zeros(len(f), np.int64)
    for ni in range(len(nodes)):
        for k in range(4):
            r = nodes[ni, 24 + k]
            if r == -1 or r >= 0:
                continue
            u = np.uint32(r & 0x7FFFFFFF)
            first, cnt = int(u >> 4), int(u & 15) + 1
            seen[first:first + cnt] += 1
            tri = abc[first:first + cnt]
            assert (tri.min(axis=(0, 1)) >= bounds[ni, 0::2, k]).all()
            assert (tri.max(axis=(0, 1)) <= bounds[ni, 1::2, k]).all()
    assert (seen == 1).all()


@pytest.mark.parametrize("build", ["gpu", "gpu_sah", "gpu_lbvh"])
@pytest.mark.parametrize("leaf", [0, 1, 4, 13, 16])
def test_gpu_build_structure(gpu, leaf, build):
    v, f, _ = S.perturbed_sphere(5, seed=2)
    bvh = cubvh_amd.cuBVH(v, f, leaf_size=leaf, build=build)
    info = bvh.info()
    assert info["n_faces"] == len(f) and info["stack_need"] == 3 * info["depth"] + 1 <= 64
    assert info["leaf_size"] == (leaf or 16)
    _leaves_cover_faces(bvh, v, f)


def _leaf_sets(bvh):
    sd = bvh.state_dict()
    nodes = sd["nodes"].numpy()
    ids = sd["triangles"].numpy().view(np.int64)[:, 5]
    out = set()
    for r in nodes[:, 24:28].reshape(-1):
        if r == -1 or r >= 0:
            continue
        u = np.uint32(r & 0x7FFFFFFF)
        first, cnt = int(u >> 4), int(u & 15) + 1
        out.add(frozenset(ids[first:first + cnt].tolist()))
    return out


@pytest.mark.parametrize("leaf", [0, 3])
def test_gpu_median_build_has_the_host_median_leaves(gpu, leaf):
    """The GPU median build halves every run at the median centroid of its longest axis,
    the host's object-median rule: both trees have exactly the same leaves."""
    v, f, _ = S.perturbed_sphere(6, seed=7)
    g = cubvh_amd.cuBVH(v, f, leaf_size=leaf, build="gpu")
    h = cubvh_amd.cuBVH(v, f, leaf_size=leaf | 0x100)  # host build, object median
    assert _leaf_sets(g) == _leaf_sets(h)


@pytest.mark.parametrize("build", ["gpu", "gpu_sah", "gpu_lbvh"])
def test_gpu_build_queries_equal_host_tree_and_oracle(gpu, oracle, ico4, build):
    v, f = ico4
    p = torch.from_numpy(S.uniform_points(30000, seed=5)).to(gpu)
    g = cubvh_amd.cuBVH(v, f, build=build)
    h = cubvh_amd.cuBVH(v, f)
    for args in ((False,), (True,)):
        a = g.unsigned_distance(p, return_uvw=args[0])
        b = h.unsigned_distance(p, return_uvw=args[0])
        for x, y in zip(a, b):
            assert (x is None and y is None) or torch.equal(x, y)
    od, of, ou = oracle.udf_brute(v, f, _np(p)[:3000])
    d, fi, u = g.unsigned_distance(p[:3000], return_uvw=True)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(u), ou)
    for mode in ("watertight", "raystab"):
        a = g.signed_distance(p[:8192], mode=mode)
        b = h.signed_distance(p[:8192], mode=mode)
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    o, dirs = (torch.from_numpy(x).to(gpu) for x in S.sphere_rays(20000, seed=6))
    a = g.ray_trace(o, dirs)
    b = h.ray_trace(o, dirs)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_gpu_sah_build_on_duplicate_faces(gpu):
    """Thousands of copies of one triangle tie every split cost; the SAH levels stop
    after a bounded number of levels (median splits below), and a tree too deep for
    the walk's stack is rebuilt with median splits.  Results equal the host tree's."""
    v, f, _ = S.perturbed_sphere(4, seed=3)
    f = np.concatenate([f, np.repeat(f[:1], 3000, axis=0)]).astype(f.dtype)
    g = cubvh_amd.cuBVH(v, f, build="gpu_sah")
    h = cubvh_amd.cuBVH(v, f)
    info = g.info()
    assert info["stack_need"] == 3 * info["depth"] + 1 <= 64
    _leaves_cover_faces(g, v, f)
    p = torch.from_numpy(S.uniform_points(20000, seed=9)).to(gpu)
    a = g.unsigned_distance(p, return_uvw=True)
    b = h.unsigned_distance(p, return_uvw=True)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_gpu_sah_build_on_geometric_clusters(gpu):
    """Triangles at geometrically shrinking scales make very unbalanced SAH splits; the
    GPU SAH build still yields a tree the walk accepts (median splits past 40 levels,
    or the median rebuild), with the host tree's results."""
    tris = []
    for k in range(3000):
        s = 0.7 ** (k % 60)
        c = np.array([s, 0.0, 0.0])
        tris.append([c, c + [s * 0.1, 0, 0], c + [0, s * 0.1, 0]])
    v = np.asarray(tris, np.float32).reshape(-1, 3)
    f = np.arange(len(v), dtype=np.int32).reshape(-1, 3)
    g = cubvh_amd.cuBVH(v, f, build="gpu_sah")
    assert g.info()["stack_need"] <= 64
    _leaves_cover_faces(g, v, f)
    h = cubvh_amd.cuBVH(v, f)
    p = torch.from_numpy(S.uniform_points(20000, seed=11)).to(gpu)
    a = g.unsigned_distance(p, return_uvw=True)
    b = h.unsigned_distance(p, return_uvw=True)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_gpu_sah_tree_is_shallower_in_cost_than_median(gpu):
    """The SAH tree's expected work (sum over nodes of child box area x faces below)
    is below the median tree's on a non-uniform mesh."""
    v, f, _ = S.perturbed_sphere(6, seed=4)
    v = v * np.array([4.0, 1.0, 0.25], np.float32)  # stretched: median splits cut badly

    def sah_cost(bvh):
        nodes = bvh.state_dict()["nodes"].numpy()
        b = nodes[:, :24].view(np.float32).reshape(-1, 6, 4)
        d = np.clip(b[:, 1::2, :] - b[:, 0::2, :], 0, None)
        area = 2 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
        return float(area.sum())

    assert sah_cost(cubvh_amd.cuBVH(v, f, build="gpu_sah")) < sah_cost(cubvh_amd.cuBVH(v, f, build="gpu"))


def test_gpu_build_level6_and_reuse(gpu):
    """A larger mesh; the GPU-built tree's state_dict imports, and it joins a forest
    (whose host arrays it recomputes)."""
    v, f, _ = S.perturbed_sphere(6, seed=1)
    g = cubvh_amd.cuBVH(v, f, build="gpu")
    h = cubvh_amd.cuBVH(v, f)
    p = torch.from_numpy(S.grid_points(48)).to(gpu)
    dg, fg, _ = g.unsigned_distance(p)
    dh, fh, _ = h.unsigned_distance(p)
    assert torch.equal(dg, dh) and torch.equal(fg, fh)
    r = cubvh_amd.cuBVH.from_state_dict(g.state_dict())
    dr, fr, _ = r.unsigned_distance(p)
    assert torch.equal(dr, dh) and torch.equal(fr, fh)
    forest = cubvh_amd.cuBVHForest([g, h])
    pts = torch.cat([p, p])
    offs = torch.tensor([0, p.shape[0], 2 * p.shape[0]], dtype=torch.int64, device=gpu)
    d2, f2, _ = forest.unsigned_distance(pts, offs)
    assert torch.equal(d2[: p.shape[0]], dh) and torch.equal(d2[p.shape[0]:], dh)
    assert torch.equal(f2[: p.shape[0]], fh) and torch.equal(f2[p.shape[0]:], fh)


@pytest.mark.parametrize("build", ["gpu", "gpu_sah"])
def test_build_many_equals_one_by_one(gpu, build):
    """cuBVH.build_many: concurrent builds on 4 streams, host and device inputs; each
    tree answers as the tree built alone."""
    # 320 faces (one short run: the lane-group path from the root), 1,280 and 5,120 (bins at the top)
    meshes = [S.perturbed_sphere(2 + (i % 3), seed=20 + i)[:2] for i in range(9)]
    inputs = [(v, f) if i % 2 else (torch.from_numpy(v).to(gpu), torch.from_numpy(f).to(gpu))
              for i, (v, f) in enumerate(meshes)]
    many = cubvh_amd.cuBVH.build_many(inputs, build=build, device=gpu, streams=4)
    p = torch.from_numpy(S.uniform_points(5000, seed=12)).to(gpu)
    for (v, f), b in zip(meshes, many):
        one = cubvh_amd.cuBVH(v, f, build=build)
        assert b.info() == one.info()
        for x, y in zip(b.unsigned_distance(p, return_uvw=True), one.unsigned_distance(p, return_uvw=True)):
            assert torch.equal(x, y)
    forest = cubvh_amd.cuBVHForest(many)
    assert forest.info()["n_faces"] == sum(len(f) for _, f in meshes)


def test_sah_trees_do_less_reference_work_than_the_reference_tree(gpu, oracle):
    """Tree quality as the reference's walk counts it (N_box, N_tri per query): at the
    reference's leaf size (8) the host and GPU SAH trees, exported to the reference
    layout and walked by the
    oracle's restated reference traversal, cost fewer bytes per query than the
    reference's own build; the distances equal."""
    v, f, _ = S.perturbed_sphere(6, seed=5)
    pts = S.uniform_points(1 << 14, seed=13)
    ref = oracle.RefBvh(v, f)
    d_ref, _, _, c = ref.udf(pts, return_uvw=False, tie_mode=0)
    ref_bytes = 24 * c[0] + 36 * c[1]
    for kw in ({"leaf_size": 8}, {"build": "gpu_sah", "leaf_size": 8}):
        sd = cubvh_amd.cuBVH(v, f, **kw).state_dict(layout="reference")
        tree = oracle.RefBvh.from_arrays(sd["nodes"].numpy(), sd["triangles"].numpy())
        d, _, _, c = tree.udf(pts, return_uvw=False, tie_mode=0)
        np.testing.assert_array_equal(d, d_ref)
        assert 24 * c[0] + 36 * c[1] < ref_bytes, (kw, c, ref_bytes)


def test_results_do_not_depend_on_leaf_size(gpu, ico4):
    """Every query form gives the same answers on host trees with leaves of 1, 8, 13 and
    16 faces (conservative boxes and disks, ties to the lowest face id)."""
    v, f = ico4
    p = torch.from_numpy(S.uniform_points(20000, seed=21)).to(gpu)
    o, dirs = (torch.from_numpy(x).to(gpu) for x in S.sphere_rays(20000, seed=22))
    ref = None
    for leaf in (1, 8, 13, 16):
        b = cubvh_amd.cuBVH(v, f, leaf_size=leaf)
        out = list(b.unsigned_distance(p, return_uvw=True))
        for mode in ("watertight", "raystab"):
            out += list(b.signed_distance(p[:8192], mode=mode, return_uvw=True))
        out += list(b.ray_trace(o, dirs))
        if ref is None:
            ref = out
        else:
            for x, y in zip(out, ref):
                assert torch.equal(x, y), leaf


def test_gpu_build_checks_input(gpu):
    from cubvh_amd import _lib

    lib = _lib.load()
    v = torch.rand(40, 3, device=gpu)
    f = torch.randint(0, 40, (30, 3), dtype=torch.int32, device=gpu)
    h = ctypes.c_void_p()
    s = torch.cuda.current_stream(gpu).cuda_stream
    f_bad = f.clone()
    f_bad[17, 2] = 40
    assert lib.cubvh_bvh_create_gpu(gpu.index, v.data_ptr(), 40, f_bad.data_ptr(), 30, 0, ctypes.byref(h), s) == -1
    assert b"outside" in lib.cubvh_last_error()
    v_bad = v.clone()
    v_bad[int(f[4, 1]), 1] = float("inf")  # a vertex that a face uses (unused vertices are not read)
    assert lib.cubvh_bvh_create_gpu(gpu.index, v_bad.data_ptr(), 40, f.data_ptr(), 30, 0, ctypes.byref(h), s) == -1
    assert b"finite" in lib.cubvh_last_error()
    # a mesh of at most leaf_size faces is one leaf (built on the host after a copy back)
    assert lib.cubvh_bvh_create_gpu(gpu.index, v.data_ptr(), 40, f.data_ptr(), 5, 0, ctypes.byref(h), s) == 0
    info = (ctypes.c_int64 * 5)()
    assert lib.cubvh_bvh_info(h, info) == 0 and info[0] == 1 and info[1] == 5
    lib.cubvh_bvh_destroy(h)
    assert lib.cubvh_bvh_create_gpu(gpu.index, v.data_ptr(), 40, f.data_ptr(), 30, 0, ctypes.byref(h), s) == 0
    lib.cubvh_bvh_destroy(h)


def test_forest_of_gpu_built_trees_assembled_on_device(gpu):
    """Trees that live only on the device form a forest there (no host copy): results
    equal the per-mesh calls, and the same forest from host-built trees."""
    meshes = [S.perturbed_sphere(l, seed=s)[:2] for l, s in ((3, 1), (5, 2), (4, 3))]
    g = [cubvh_amd.cuBVH(v, f, build="gpu") for v, f in meshes]
    h = [cubvh_amd.cuBVH(v, f) for v, f in meshes]
    fg = cubvh_amd.cuBVHForest(g)
    fh = cubvh_amd.cuBVHForest(h)
    assert fg.info()["n_faces"] == sum(len(f) for _, f in meshes)
    assert fg.info()["n_nodes"] == sum(b.info()["n_nodes"] for b in g)
    pts = [torch.from_numpy(S.uniform_points(n, seed=n)).to(gpu) for n in (3000, 9000, 100)]
    P = torch.cat(pts)
    offs = torch.tensor([0, 3000, 12000, 12100], dtype=torch.int64, device=gpu)
    dg, fgi, ug = fg.unsigned_distance(P, offs, return_uvw=True)
    dh, fhi, uh = fh.unsigned_distance(P, offs, return_uvw=True)
    assert torch.equal(dg, dh) and torch.equal(fgi, fhi) and torch.equal(ug, uh)
    for m, (b, p) in enumerate(zip(g, pts)):
        d, fi, _ = b.unsigned_distance(p)
        assert torch.equal(dg[offs[m]:offs[m + 1]], d) and torch.equal(fgi[offs[m]:offs[m + 1]], fi)
