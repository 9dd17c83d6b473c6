"""GPU parity of the BVH queries (HIP kernels in libcubvh_amd.so) against the oracle.

Bar (tolerances written in each test):
  * against the oracle's fp32 brute force (same formulas, canonical ties to the
    lowest face index): distances, face ids, uvw, depths, hit positions and signs
    are bit-exact;
  * against FP64: |d - d64| <= 1e-5 * max(d64, 1); hit/miss and clear-winner face
    ids exact; closest point rebuilt from face_id + uvw within 1e-5.
"""
import numpy as np
import pytest
import torch

import cubvh_amd
from cubvh_amd import synthetic as S
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_udf_small_workload_bit_exact(gpu, oracle, ico4):
    v, f = ico4
    p = S.uniform_points(65536, seed=0)
    bvh = cubvh_amd.cuBVH(v, f)
    d, fi, uvw = bvh.unsigned_distance(torch.from_numpy(p).to(gpu), return_uvw=True)
    assert fi.dtype == torch.int64 and d.dtype == torch.float32 and uvw.shape == (65536, 3)
    od, of, ou = oracle.udf_brute(v, f, p)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(uvw), ou)
    # committed golden answers (first 4096 points), so a changed oracle cannot hide a change
    g = np.load(GOLDEN / "config1_subset.npz")
    np.testing.assert_array_equal(_np(d)[:4096], g["dist"])
    np.testing.assert_array_equal(_np(fi)[:4096], g["face"])
    np.testing.assert_array_equal(_np(uvw)[:4096], g["uvw"])


def test_udf_vs_fp64(gpu, oracle, ico4):
    v, f = ico4
    p = S.uniform_points(16384, seed=11)
    bvh = cubvh_amd.cuBVH(v, f)
    d, fi, uvw = bvh.unsigned_distance(torch.from_numpy(p).to(gpu), return_uvw=True)
    d, fi, uvw = _np(d).astype(np.float64), _np(fi), _np(uvw).astype(np.float64)
    d64, f64, cp64, gap = oracle.udf_f64(v, f, p)
    err = np.abs(d - d64) / np.maximum(d64, 1.0)
    assert err.max() <= 1e-5, err.max()
    print(f"max |d - d64| / max(d64, 1) = {err.max():.3g}")
    clear = gap > 1e-5 * np.maximum(d64, 1.0)
    assert (fi[clear] == f64[clear]).all()
    vd = v.astype(np.float64)
    cp = sum(vd[f[fi, k]] * uvw[:, [k]] for k in range(3))
    assert np.abs(cp - cp64)[clear].max() <= 1e-5
    assert np.abs(np.linalg.norm(p - cp, axis=1) - d64).max() <= 1e-5


def test_udf_reference_fixture_dodecahedron(gpu, oracle, dodecahedron):
    """The reference's own check (test/unsigned_distance.py:93-139): 1000 randn points
    on the dodecahedron, distance and rebuilt closest point within atol 1e-5."""
    v, f = dodecahedron
    p = np.random.default_rng(0).normal(size=(1000, 3)).astype(np.float32)
    bvh = cubvh_amd.cuBVH(v, f)
    d, fi, uvw = bvh.unsigned_distance(torch.from_numpy(p), return_uvw=True)  # CPU input is moved to the GPU
    assert d.is_cuda
    d64, _, cp64, gap = oracle.udf_f64(v, f, p)
    np.testing.assert_allclose(_np(d), d64, atol=1e-5)
    fi, uvw = _np(fi), _np(uvw).astype(np.float64)
    vd = v.astype(np.float64)
    cp = sum(vd[f[fi, k]] * uvw[:, [k]] for k in range(3))
    clear = gap > 1e-5
    np.testing.assert_allclose(cp[clear], cp64[clear], atol=1e-5)


def test_udf_leading_dims_and_far_points(gpu, oracle, ico4):
    v, f = ico4
    bvh = cubvh_amd.cuBVH(v, f)
    p = torch.from_numpy(S.uniform_points(2 * 3 * 50, seed=5).reshape(2, 3, 50, 3)).double()
    d, fi, uvw = bvh.unsigned_distance(p.to(gpu), return_uvw=True)
    assert d.shape == (2, 3, 50) and fi.shape == (2, 3, 50) and uvw.shape == (2, 3, 50, 3)
    d2, _, _ = bvh.unsigned_distance(p.reshape(-1, 3).to(gpu))
    assert torch.equal(d.reshape(-1), d2)
    # nothing within 1000: distance 0 and face 0 (src/bvh.cu:346-350 returns slot 0)
    far = torch.tensor([[5000.0, 0, 0], [0, -3000.0, 0]], device=gpu)
    d, fi, _ = bvh.unsigned_distance(far)
    assert _np(d).tolist() == [0.0, 0.0] and _np(fi).tolist() == [0, 0]
    # empty query set
    d, fi, uvw = bvh.unsigned_distance(torch.zeros(0, 3, device=gpu), return_uvw=True)
    assert d.shape == (0,) and uvw.shape == (0, 3)


def test_udf_perturbed_level6_bit_exact(gpu, oracle):
    """A deeper tree (81,920 faces) against the oracle's reference-BVH walk with
    canonical ties (itself equal to the brute force, see test_oracle_bvh)."""
    v, f = S.jittered_sphere(6)
    p = S.normal_points(20000, 0.4, seed=2)
    bvh = cubvh_amd.cuBVH(v, f)
    d, fi, uvw = bvh.unsigned_distance(torch.from_numpy(p).to(gpu), return_uvw=True)
    R = oracle.RefBvh(v, f)
    od, of, ou, _ = R.udf(p, tie_mode=1)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(uvw), ou)
    bd, bf, bu = oracle.udf_brute(v, f, p[:2000])
    np.testing.assert_array_equal(_np(d)[:2000], bd)
    np.testing.assert_array_equal(_np(fi)[:2000], bf)


def test_rays_bit_exact_and_vs_fp64(gpu, oracle, ico4):
    v, f = ico4
    ro, rd = S.sphere_rays(65536, seed=0)
    bvh = cubvh_amd.cuBVH(v, f)
    pos, fi, t = bvh.ray_trace(torch.from_numpy(ro).to(gpu), torch.from_numpy(rd).to(gpu))
    opos, of, ot = oracle.ray_brute(v, f, ro, rd)
    np.testing.assert_array_equal(_np(t), ot)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(pos), opos)
    g = np.load(GOLDEN / "config1_subset.npz")
    np.testing.assert_array_equal(_np(t)[:4096], g["ray_t"])
    np.testing.assert_array_equal(_np(fi)[:4096], g["ray_face"])
    t64, f64, amb = oracle.ray_f64(v, f, ro, rd, tol=1e-5)
    ok = ~amb
    fi, t = _np(fi), _np(t)
    assert ((fi >= 0) == (f64 >= 0))[ok].all()
    hit = ok & (f64 >= 0)
    assert (fi[hit] == f64[hit]).all()
    np.testing.assert_allclose(t[hit], t64[hit], rtol=1e-5)
    assert (t[fi < 0] == 1000.0).all()


def test_rays_miss_and_shapes(gpu, ico4):
    v, f = ico4
    bvh = cubvh_amd.cuBVH(v, f)
    o = torch.tensor([[[0.0, 0.0, 3.0], [0.0, 0.0, 3.0]]], device=gpu)
    d = torch.tensor([[[0.0, 0.0, -1.0], [0.0, 1.0, 0.0]]], device=gpu)
    pos, fi, t = bvh.ray_trace(o, d)
    assert pos.shape == (1, 2, 3) and fi.shape == (1, 2)
    assert _np(fi)[0, 1] == -1 and _np(t)[0, 1] == 1000.0
    np.testing.assert_allclose(_np(pos)[0, 1], [0, 1000.0, 3.0])
    assert _np(fi)[0, 0] >= 0 and abs(_np(t)[0, 0] - 2.5) < 2e-3  # sphere of radius 0.5
    # axis-parallel rays (zero direction components) are handled
    o = torch.tensor([[0.01, 0.02, -2.0], [2.0, 0.0, 0.0]], device=gpu)
    d = torch.tensor([[0.0, 0.0, 1.0], [-1.0, 0.0, 0.0]], device=gpu)
    _, fi, t = bvh.ray_trace(o, d)
    assert (_np(fi) >= 0).all()
    np.testing.assert_allclose(_np(t), [1.5, 1.5], atol=5e-3)


def test_sdf_watertight_bit_exact(gpu, oracle, ico4):
    v, f = ico4
    p = S.uniform_points(32768, seed=6)
    bvh = cubvh_amd.cuBVH(v, f)
    d, fi, uvw = bvh.signed_distance(torch.from_numpy(p).to(gpu), return_uvw=True, mode="watertight")
    od, of, ou = oracle.sdf_brute(v, f, p, 0)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(uvw), ou)
    w = oracle.winding_f64(v, f, p)
    far = np.abs(od) > 1e-4
    assert ((_np(d) < 0) == (w > 0.5))[far].all()


def test_sdf_raystab_bit_exact_closed_and_soup(gpu, oracle, ico4):
    v, f = ico4
    p = S.uniform_points(8192, seed=8)
    bvh = cubvh_amd.cuBVH(v, f)
    d, fi, _ = bvh.signed_distance(torch.from_numpy(p).to(gpu), mode="raystab")
    od, of, _ = oracle.sdf_brute(v, f, p, 1, return_uvw=False)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    w = oracle.winding_f64(v, f, p)
    assert ((_np(d) < 0) == (w > 0.5))[np.abs(od) > 1e-4].all()
    # below 2,048 queries the per-lane kernel runs instead of the ordered pair of kernels;
    # the stab rays are keyed by the query index, so a prefix gives the same answers
    ds, fs, _ = bvh.signed_distance(torch.from_numpy(p[:1000]).to(gpu), mode="raystab")
    np.testing.assert_array_equal(_np(ds), od[:1000])
    np.testing.assert_array_equal(_np(fs), of[:1000])
    # non-watertight soup: level-4 sphere minus 1% of its faces plus small random
    # triangles (with 1% holes about half of the inside points still see 64 hits)
    rng = np.random.default_rng(3)
    keep = f[np.sort(rng.permutation(len(f))[: int(0.99 * len(f))])]
    c = rng.uniform(-0.6, 0.6, size=(500, 3))
    tv = (c[:, None, :] + rng.normal(0, 0.02, size=(500, 3, 3))).reshape(-1, 3)
    V2 = np.concatenate([v, tv]).astype(np.float32)
    F2 = np.concatenate([keep, len(v) + np.arange(1500).reshape(-1, 3)]).astype(np.int32)
    bvh2 = cubvh_amd.cuBVH(V2, F2)
    d, fi, uvw = bvh2.signed_distance(torch.from_numpy(p).to(gpu), return_uvw=True, mode="raystab")
    od, of, ou = oracle.sdf_brute(V2, F2, p, 1)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(uvw), ou)
    assert (od < 0).any() and (od > 0).any()
    # an odd count: the refilled raystab grid's last chunk is partial
    p3 = S.uniform_points(10037, seed=9)
    d, fi, _ = bvh2.signed_distance(torch.from_numpy(p3).to(gpu), mode="raystab")
    od, of, _ = oracle.sdf_brute(V2, F2, p3, 1, return_uvw=False)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)


def test_sdf_raystab_threshold_sizes_and_empty_lists(gpu, oracle, ico4):
    """The raystab passes at the packet path's threshold (2,048 and 2,049 queries: lists
    shorter than a workgroup) and with every query outside the mesh (line 0 lists no
    query, so every line batch and compaction runs on an empty list): bit-exact with the
    oracle."""
    v, f = ico4
    bvh = cubvh_amd.cuBVH(v, f)
    for n, seed in ((2048, 21), (2049, 22)):
        p = S.uniform_points(n, seed=seed, lo=-0.7, hi=0.7)
        d, fi, _ = bvh.signed_distance(torch.from_numpy(p).to(gpu), mode="raystab")
        od, of, _ = oracle.sdf_brute(v, f, p, 1, return_uvw=False)
        np.testing.assert_array_equal(_np(d), od)
        np.testing.assert_array_equal(_np(fi), of)
        assert (od < 0).any() and (od > 0).any()
    far = S.uniform_points(4096, seed=23, lo=2.0, hi=3.0)
    d, fi, _ = bvh.signed_distance(torch.from_numpy(far).to(gpu), mode="raystab")
    od, of, _ = oracle.sdf_brute(v, f, far, 1, return_uvw=False)
    np.testing.assert_array_equal(_np(d), od)
    np.testing.assert_array_equal(_np(fi), of)
    assert (od > 0).all()


def test_state_dict_roundtrip_on_gpu(gpu, oracle, ico4):
    """test/state_dict.py:96-112: save, load, from_state_dict give identical results;
    the reference's own layout loads too."""
    v, f = ico4
    p = torch.from_numpy(S.uniform_points(4096, seed=9)).to(gpu)
    bvh = cubvh_amd.cuBVH(v, f)
    a = bvh.unsigned_distance(p, return_uvw=True)
    b = cubvh_amd.cuBVH.from_state_dict(bvh.state_dict()).unsigned_distance(p, return_uvw=True)
    nodes9, tris12 = oracle.RefBvh(v, f).export()
    ref = cubvh_amd.cuBVH.from_state_dict({"nodes": torch.from_numpy(nodes9), "triangles": torch.from_numpy(tris12)})
    c = ref.unsigned_distance(p, return_uvw=True)
    for x, y, z in zip(a, b, c):
        assert torch.equal(x, y) and torch.equal(x, z)


def test_ordered_distance_equals_plain_calls(gpu, oracle):
    """morton_order + distance_ordered give the plain calls' results for the Morton
    permutation, for an arbitrary permutation, and skip order entries >= n."""
    v, f, _ = S.perturbed_sphere(5, seed=3)
    p = torch.from_numpy(S.uniform_points(40000, seed=4)).to(gpu)
    bvh = cubvh_amd.cuBVH(v, f)
    d0, f0, u0 = bvh.unsigned_distance(p, return_uvw=True)
    s0, g0, w0 = bvh.signed_distance(p, return_uvw=True, mode="watertight")
    order = cubvh_amd.morton_order(p)
    assert torch.equal(torch.sort(order.long()).values, torch.arange(p.shape[0], device=gpu))
    for perm in (order, torch.randperm(p.shape[0], device=gpu).to(torch.int32)):
        d, fi, u = bvh.distance_ordered(p, perm, return_uvw=True)
        assert torch.equal(d, d0) and torch.equal(fi, f0) and torch.equal(u, u0)
        s, g, w = bvh.distance_ordered(p, perm, return_uvw=True, signed=True)
        assert torch.equal(s, s0) and torch.equal(g, g0) and torch.equal(w, w0)
    od, of, _ = oracle.udf_brute(v, f, _np(p)[:2000])
    np.testing.assert_array_equal(_np(d0)[:2000], od)
    np.testing.assert_array_equal(_np(f0)[:2000], of)
    # an out-of-range entry leaves its query unwritten and the others exact
    bad = order.clone()
    victim = int(bad[7])
    bad[7] = p.shape[0] + 5
    lib = cubvh_amd._lib.load()
    n = p.shape[0]
    dist = torch.full((n,), -7.0, device=gpu)
    face = torch.full((n,), -9, dtype=torch.int64, device=gpu)
    cubvh_amd._lib.check(lib.cubvh_distance_ordered(bvh._h, gpu.index, n, 0, p.data_ptr(), bad.data_ptr(),
                                                    dist.data_ptr(), face.data_ptr(), None,
                                                    torch.cuda.current_stream(gpu).cuda_stream))
    torch.cuda.synchronize()
    assert float(dist[victim]) == -7.0 and int(face[victim]) == -9
    keep = torch.ones(n, dtype=torch.bool, device=gpu)
    keep[victim] = False
    assert torch.equal(dist[keep], d0[keep]) and torch.equal(face[keep], f0[keep])


def test_duplicate_faces_tie_to_lowest_index(gpu, oracle, ico4):
    """Every face appears twice (and some three times) with higher indices: every
    distance, ray depth and sign ties exactly, and the answer must be the lowest
    index -- the original face -- in every query form and on both builders."""
    v, f = ico4
    dup = np.concatenate([f, f[::-1], f[:500]]).astype(np.int32)
    p = S.uniform_points(20000, seed=12)
    P = torch.from_numpy(p).to(gpu)
    ref = cubvh_amd.cuBVH(v, f)
    d0, f0, u0 = ref.unsigned_distance(P, return_uvw=True)
    o, dirs = S.sphere_rays(20000, seed=13)
    O, Dd = torch.from_numpy(o).to(gpu), torch.from_numpy(dirs).to(gpu)
    r0 = ref.ray_trace(O, Dd)
    s0 = ref.signed_distance(P, mode="watertight")
    for build in ("host", "gpu"):
        b = cubvh_amd.cuBVH(v, dup, build=build)
        d, fi, u = b.unsigned_distance(P, return_uvw=True)
        assert torch.equal(d, d0) and torch.equal(fi, f0) and torch.equal(u, u0)
        r = b.ray_trace(O, Dd)
        for x, y in zip(r, r0):
            assert torch.equal(x, y)
        # the sign averages the normals of the faces at the closest point, duplicates
        # included (as the reference does): compare with the oracle on this mesh
        s = b.signed_distance(P, mode="watertight")
        assert torch.equal(s[0].abs(), s0[0].abs()) and torch.equal(s[1], s0[1])
        osd, osf, _ = oracle.sdf_brute(v, dup, p[:2000], 0, return_uvw=False)
        np.testing.assert_array_equal(_np(s[0])[:2000], osd)
        np.testing.assert_array_equal(_np(s[1])[:2000], osf)
    od, of, _ = oracle.udf_brute(v, dup, p[:2000])
    np.testing.assert_array_equal(_np(d0)[:2000], od)
    np.testing.assert_array_equal(_np(f0)[:2000], of)
    assert int(f0.max()) < len(f)


def test_camera_rays_packet_path_bit_exact(gpu, oracle):
    """Pixel-ordered camera rays (whole waves nearly parallel, one origin) take the
    packet walk; rays of mixed waves walk per lane.  Both against the oracle."""
    v, f, _ = S.perturbed_sphere(5, seed=4)
    o, d = S.camera_rays(width=128, height=48, n_poses=2, seed=3)
    # second half: the same rays shuffled, so that no wave is coherent
    perm = np.random.default_rng(5).permutation(len(o))
    O = np.concatenate([o, o[perm]])
    D = np.concatenate([d, d[perm]])
    b = cubvh_amd.cuBVH(v, f)
    pos, fi, t = b.ray_trace(torch.from_numpy(O).to(gpu), torch.from_numpy(D).to(gpu))
    op, of, ot = oracle.ray_brute(v, f, O, D)
    np.testing.assert_array_equal(_np(t), ot)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(pos), op)
    assert (of >= 0).mean() > 0.05 and (of < 0).any()  # the sphere fills ~11% of these frames


def test_valu_probe_counts_and_times(gpu):
    """cubvh_valu_probe runs 8 workgroups of 4 waves per CU (8 waves on each SIMD), each
    lane 16 fma chains: its instruction count is CUs x 8 x 4 x iters x 16 and, after a
    first call that pays the one-time launch costs, its time grows with iters; the
    SGPR-operand form is not faster than the VGPR form."""
    import ctypes

    from cubvh_amd import _lib

    lib = _lib.load()
    n_cu = torch.cuda.get_device_properties(gpu).multi_processor_count
    out = {}
    for form in (0, 1):
        warm_wi, warm_ms = ctypes.c_double(0.0), ctypes.c_float(0.0)
        _lib.check(lib.cubvh_valu_probe(gpu.index, form, 2000, ctypes.byref(warm_wi), ctypes.byref(warm_ms)))
        for iters in (2000, 8000):
            wi = ctypes.c_double(0.0)
            ms = ctypes.c_float(0.0)
            _lib.check(lib.cubvh_valu_probe(gpu.index, form, iters, ctypes.byref(wi), ctypes.byref(ms)))
            assert wi.value == n_cu * 8 * 4 * iters * 16
            assert ms.value > 0
            out[form, iters] = ms.value
        assert out[form, 8000] > 2.0 * out[form, 2000]
    assert out[1, 8000] > 0.9 * out[0, 8000]


def test_nan_and_inf_queries(gpu, oracle, ico4):
    """Query points and rays with NaN or infinite coordinates.  The reference's walks
    compare with `<=` and find nothing for them (bvh.cu:300-353, 250-298): a point gets
    the not-found answer (face 0, distance 0; uvw from face 0), a ray a miss; the other
    queries of the same call are unaffected.  Bit-exact against the oracle, on the
    ordered packet path (4,096 queries) and the per-lane path (500)."""
    v, f = ico4
    bvh = cubvh_amd.cuBVH(v, f)
    for n in (4096, 500):
        p = S.uniform_points(n, seed=21)
        p[::7, 0] = np.nan
        p[3::11, 1] = np.inf
        p[5::13, 2] = -np.inf
        P = torch.from_numpy(p).to(gpu)
        d, fi, uvw = bvh.unsigned_distance(P, return_uvw=True)
        od, of, ou = oracle.udf_brute(v, f, p)
        np.testing.assert_array_equal(_np(d), od)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(uvw), ou)
        bad = ~np.isfinite(p).all(axis=1)
        assert (od[bad] == 0).all() and (of[bad] == 0).all() and (od[~bad] > 0).all()
        for mode, m in (("watertight", 0), ("raystab", 1)):
            d, fi, _ = bvh.signed_distance(P, mode=mode)
            od, of, _ = oracle.sdf_brute(v, f, p, m, return_uvw=False)
            np.testing.assert_array_equal(_np(d), od)
            np.testing.assert_array_equal(_np(fi), of)
        ro, rd = S.sphere_rays(n, seed=22)
        ro[::9, 1] = np.nan
        rd[4::10, 0] = np.nan
        ro[2::15, 2] = np.inf
        pos, fi, t = bvh.ray_trace(torch.from_numpy(ro).to(gpu), torch.from_numpy(rd).to(gpu))
        opos, of, ot = oracle.ray_brute(v, f, ro, rd)
        np.testing.assert_array_equal(_np(t), ot)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(pos), opos)
        rbad = ~(np.isfinite(ro).all(axis=1) & np.isfinite(rd).all(axis=1))
        assert (of[rbad] == -1).all() and (of[~rbad] >= 0).any()


def _degenerate_mesh(v, f, seed=4):
    """ico4 with 40 faces collapsed to segments (two corners on one vertex), 20 to points
    and 30 turned into collinear slivers (a third vertex on the segment of the other two):
    zero-length edges and zero normals, whose 0 / 0 terms the reference formula carries
    through its clamp and min (triangle.cuh:46-68, common.h:53-56)."""
    v = v.astype(np.float32)
    f = f.astype(np.int32).copy()
    idx = np.random.default_rng(seed).permutation(len(f))
    f[idx[:40], 1] = f[idx[:40], 0]
    f[idx[40:60], 1] = f[idx[40:60], 0]
    f[idx[40:60], 2] = f[idx[40:60], 0]
    k = idx[60:90]
    mid = (0.25 * v[f[k, 0]] + 0.75 * v[f[k, 1]]).astype(np.float32)
    f[k, 2] = len(v) + np.arange(len(k))
    deg = np.zeros(len(f), bool)
    deg[idx[:90]] = True
    return np.concatenate([v, mid]).astype(np.float32), f, deg


def test_degenerate_faces_bit_exact(gpu, oracle, ico4):
    """Meshes with zero-area faces: distances, faces and uvw (some queries are won by a
    degenerate face), both signed-distance modes and rays equal the oracle bit for bit on
    the host-built tree, on the ordered packet path and the per-lane path; the GPU-built
    trees give the host tree's answers."""
    v, f, deg = _degenerate_mesh(*ico4)
    bvh = cubvh_amd.cuBVH(v, f)
    p = S.uniform_points(20000, seed=5)
    for q in (p, p[:1000]):
        d, fi, uvw = bvh.unsigned_distance(torch.from_numpy(q).to(gpu), return_uvw=True)
        od, of, ou = oracle.udf_brute(v, f, q)
        np.testing.assert_array_equal(_np(d), od)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(uvw), ou)
    assert deg[of].any()
    P = torch.from_numpy(p).to(gpu)
    for mode, m in (("watertight", 0), ("raystab", 1)):
        d, fi, uvw = bvh.signed_distance(P, return_uvw=True, mode=mode)
        od, of, ou = oracle.sdf_brute(v, f, p, m)
        np.testing.assert_array_equal(_np(d), od)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(uvw), ou)
    ro, rd = S.sphere_rays(20000, seed=6)
    pos, fi, t = bvh.ray_trace(torch.from_numpy(ro).to(gpu), torch.from_numpy(rd).to(gpu))
    opos, of, ot = oracle.ray_brute(v, f, ro, rd)
    np.testing.assert_array_equal(_np(t), ot)
    np.testing.assert_array_equal(_np(fi), of)
    np.testing.assert_array_equal(_np(pos), opos)
    ref = [_np(x) for x in bvh.unsigned_distance(P, return_uvw=True)]
    for build in ("gpu", "gpu_sah", "gpu_lbvh"):
        g = cubvh_amd.cuBVH(v, f, build=build)
        for x, y in zip(g.unsigned_distance(P, return_uvw=True), ref):
            np.testing.assert_array_equal(_np(x), y)


def test_ray_edge_cases_bit_exact(gpu, oracle, ico4):
    """Rays with a zero direction, axis-parallel directions (zero components), origins on
    a vertex and on a face's centroid, and directions in a face's plane: hits, faces and
    depths equal the oracle's, on the ordered packet path (4,096 rays) and the per-lane
    path (500)."""
    v, f = ico4
    bvh = cubvh_amd.cuBVH(v, f)
    rng = np.random.default_rng(31)
    ro, rd = S.sphere_rays(4096, seed=32)
    ro, rd = ro.copy(), rd.copy()
    rd[0:64] = 0.0  # zero direction
    axes = np.eye(3, dtype=np.float32)
    rd[64:256] = np.concatenate([axes, -axes])[rng.integers(0, 6, 192)]
    vi = rng.integers(0, len(v), 256)
    ro[256:512] = v[vi]  # origins on vertices
    tri = v[f[rng.integers(0, len(f), 256)]]
    ro[512:768] = tri.mean(axis=1)  # origins on face centroids
    e = tri[:, 1] - tri[:, 0]
    rd[768:1024] = e / np.linalg.norm(e, axis=1, keepdims=True)  # in the face's plane
    ro[768:1024] = tri[:, 0] - 0.5 * rd[768:1024]
    ro, rd = ro.astype(np.float32), rd.astype(np.float32)
    for n in (4096, 500):
        pos, fi, t = bvh.ray_trace(torch.from_numpy(ro[:n]).to(gpu), torch.from_numpy(rd[:n]).to(gpu))
        opos, of, ot = oracle.ray_brute(v, f, ro[:n], rd[:n])
        np.testing.assert_array_equal(_np(t), ot)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(pos), opos)
    assert (of[:64] == -1).all()


def test_points_on_the_surface_bit_exact(gpu, oracle, ico4):
    """Query points on vertices, edge midpoints and face centroids (distance 0 or within
    rounding of it, faces tied at 0, a search bound of 0): distances, faces (the lowest
    tied index), uvw and both signed-distance modes equal the oracle's brute force, on
    the ordered packet path (4,096 points) and the per-lane path (500)."""
    v, f = ico4
    rng = np.random.default_rng(41)
    tri = v[f[rng.integers(0, len(f), 1000)]]
    p = np.concatenate([v[rng.integers(0, len(v), 1000)], 0.5 * (tri[:, 0] + tri[:, 1]), tri.mean(axis=1)])
    p = np.concatenate([p, S.uniform_points(4096 - len(p), seed=42)]).astype(np.float32)
    p = p[rng.permutation(len(p))]
    bvh = cubvh_amd.cuBVH(v, f)
    for q in (p, p[:500]):
        Q = torch.from_numpy(q).to(gpu)
        d, fi, uvw = bvh.unsigned_distance(Q, return_uvw=True)
        od, of, ou = oracle.udf_brute(v, f, q)
        np.testing.assert_array_equal(_np(d), od)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(uvw), ou)
        for mode, m in (("watertight", 0), ("raystab", 1)):
            d, fi, uvw = bvh.signed_distance(Q, return_uvw=True, mode=mode)
            od, of, ou = oracle.sdf_brute(v, f, q, m)
            np.testing.assert_array_equal(_np(d), od)
            np.testing.assert_array_equal(np.signbit(_np(d)), np.signbit(od))
            np.testing.assert_array_equal(_np(fi), of)
            np.testing.assert_array_equal(_np(uvw), ou)
    assert (oracle.udf_brute(v, f, p)[0] == 0).sum() > 500


def _random_soup(seed):
    """A random triangle soup: 200 to 5,000 faces of random sizes (some slivers) at a
    random scale and offset, vertices shared between some faces."""
    rng = np.random.default_rng(seed)
    nf = int(rng.integers(200, 5000))
    scale = float(10.0 ** rng.uniform(-3, 3))
    centre = rng.normal(0.0, 2.0, size=3) * scale
    c = rng.uniform(-1, 1, size=(nf, 1, 3))
    e = rng.normal(0, 1, size=(nf, 3, 3)) * rng.uniform(0.01, 0.3, size=(nf, 1, 1))
    e[: nf // 10, 2] = 0.5 * (e[: nf // 10, 0] + e[: nf // 10, 1])  # slivers
    v = ((c + e).reshape(-1, 3) * scale + centre).astype(np.float32)
    f = np.arange(3 * nf, dtype=np.int32).reshape(-1, 3)
    share = rng.random(nf) < 0.3  # a corner moved onto the previous face's vertex
    f[1:][share[1:], 0] = f[:-1][share[1:], 1]
    return v, f, scale, centre


@pytest.mark.parametrize("seed", [101, 102, 103, 104, 105])
def test_random_soups_bit_exact(gpu, oracle, seed):
    """Random triangle soups over six orders of magnitude of scale: distances, faces and
    uvw, both signed-distance modes and rays equal the oracle's brute force bit for bit,
    on the ordered packet path (3,000 queries) and the per-lane path (300); the queries
    mix uniform points with points just off the surface.  The three GPU-built trees give
    the host tree's answers."""
    v, f, scale, centre = _random_soup(seed)
    rng = np.random.default_rng(seed + 1000)
    t = v[f[rng.integers(0, len(f), 1500)]].astype(np.float64)
    w = rng.dirichlet([1.0, 1.0, 1.0], size=1500)
    near = np.einsum("nk,nkd->nd", w, t) + rng.normal(0, 1e-3 * scale, size=(1500, 3))
    uni = rng.uniform(-1.5, 1.5, size=(1500, 3)) * scale + centre
    p = np.concatenate([near, uni]).astype(np.float32)[rng.permutation(3000)]
    bvh = cubvh_amd.cuBVH(v, f)
    for q in (p, p[:300]):
        Q = torch.from_numpy(q).to(gpu)
        d, fi, uvw = bvh.unsigned_distance(Q, return_uvw=True)
        od, of, ou = oracle.udf_brute(v, f, q)
        np.testing.assert_array_equal(_np(d), od)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(uvw), ou)
        for mode, m in (("watertight", 0), ("raystab", 1)):
            d, fi, _ = bvh.signed_distance(Q, mode=mode)
            od, of, _ = oracle.sdf_brute(v, f, q, m, return_uvw=False)
            np.testing.assert_array_equal(_np(d), od)
            np.testing.assert_array_equal(_np(fi), of)
    ro = (rng.uniform(-2, 2, size=(3000, 3)) * scale + centre).astype(np.float32)
    aim = v[f[rng.integers(0, len(f), 3000), 0]]
    rd = aim - ro
    rd = (rd / np.linalg.norm(rd, axis=1, keepdims=True)).astype(np.float32)
    for n in (3000, 300):
        pos, fi, tt = bvh.ray_trace(torch.from_numpy(ro[:n]).to(gpu), torch.from_numpy(rd[:n]).to(gpu))
        opos, of, ot = oracle.ray_brute(v, f, ro[:n], rd[:n])
        np.testing.assert_array_equal(_np(tt), ot)
        np.testing.assert_array_equal(_np(fi), of)
        np.testing.assert_array_equal(_np(pos), opos)
    # the GPU-built trees (object median, SAH, linear BVH) give the host tree's answers
    Q = torch.from_numpy(p).to(gpu)
    ref = [_np(x) for x in bvh.unsigned_distance(Q, return_uvw=True)]
    R = [torch.from_numpy(x).to(gpu) for x in (ro, rd)]
    rref = [_np(x) for x in bvh.ray_trace(*R)]
    for build in ("gpu", "gpu_sah", "gpu_lbvh"):
        g = cubvh_amd.cuBVH(v, f, build=build)
        for x, y in zip(g.unsigned_distance(Q, return_uvw=True), ref):
            np.testing.assert_array_equal(_np(x), y)
        for x, y in zip(g.ray_trace(*R), rref):
            np.testing.assert_array_equal(_np(x), y)
