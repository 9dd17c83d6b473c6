"""Backward of the unsigned distance (cubvh_udf_backward) against FP64 central finite
differences (h = 1e-4) and against an FP64 torch reference of the same closed form.

Tolerances: analytic vs FD within 2e-3 relative (floor 2e-3 / N) on coordinates whose
query is off the surface (distance > 1e-3) and whose FP64 closest point does not jump
across the finite-difference step (|c(p+h) - c(p-h)| < 4h: the distance is smooth
there); kernel vs FP64 torch reference within 1e-5 * |g| * (1 + 1e-2 / d), the fp32
error of (p - c) / |p - c| growing as the distance d shrinks."""
import numpy as np
import pytest
import torch

import cubvh_amd
from cubvh_amd import synthetic as S

pytestmark = pytest.mark.gpu


def test_point_and_vertex_gradients_vs_finite_differences(gpu, oracle):
    v, f = S.jittered_sphere(3, seed=1)  # 1,280 faces, small enough for FP64 brute force
    pts = S.normal_points(2048, 0.4, seed=3)
    n = len(pts)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu).requires_grad_(True)
    Vt = torch.from_numpy(v).to(gpu).requires_grad_(True)
    d, fid, uvw = cubvh_amd.udf(bvh, P, Vt)
    loss = (d * d).mean()
    loss.backward()
    gp = P.grad.cpu().numpy().astype(np.float64)
    gv = Vt.grad.cpu().numpy().astype(np.float64)

    d64, f64, cp64, gap = oracle.udf_f64(v, f, pts)
    good = d64 > 1e-3
    assert good.sum() > 1900
    h = 1e-4

    # points: dL/dp_i = 2 d_i dd_i/dp_i / n
    rng = np.random.default_rng(0)
    idx = rng.choice(np.nonzero(good)[0], 300, replace=False)
    checked = 0
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        dp, _, cpp, _ = oracle.udf_f64(v, f, pts[idx] + e)
        dm, _, cpm, _ = oracle.udf_f64(v, f, pts[idx] - e)
        smooth = np.linalg.norm(cpp - cpm, axis=1) < 4 * h
        fd = (dp * dp - dm * dm) / (2 * h) / n
        np.testing.assert_allclose(gp[idx, k][smooth], fd[smooth], rtol=2e-3, atol=2e-3 / n)
        checked += smooth.sum()
    assert checked > 0.9 * 3 * len(idx)

    # vertices: FD of the whole loss over the well-conditioned queries
    sub = pts[good]
    ng = len(sub)
    Psub = torch.from_numpy(sub).to(gpu)
    Vt2 = torch.from_numpy(v).to(gpu).requires_grad_(True)
    d2, _, _ = cubvh_amd.udf(bvh, Psub, Vt2)
    ((d2 * d2).sum() / ng).backward()
    gv2 = Vt2.grad.cpu().numpy().astype(np.float64)
    used = np.unique(f[f64[good]])
    coords = rng.choice(len(used) * 3, 64, replace=False)
    for c in coords:
        vi, k = used[c // 3], c % 3
        vp, vm = v.astype(np.float64).copy(), v.astype(np.float64).copy()
        vp[vi, k] += h
        vm[vi, k] -= h
        lp = (oracle.udf_f64(vp.astype(np.float32), f, sub)[0] ** 2).sum() / ng
        lm = (oracle.udf_f64(vm.astype(np.float32), f, sub)[0] ** 2).sum() / ng
        # vertices are fp32: use the actually applied step
        hp = np.float64(np.float32(vp[vi, k])) - np.float64(v[vi, k])
        hm = np.float64(v[vi, k]) - np.float64(np.float32(vm[vi, k]))
        fd = (lp - lm) / (hp + hm)
        assert abs(gv2[vi, k] - fd) <= 3e-3 * max(abs(fd), 1e-3), (vi, k, gv2[vi, k], fd)


def test_gradients_vs_fp64_torch_reference(gpu, oracle):
    """Kernel output equals the closed form n = (p - c)/|p - c|, grad_v = -g w n
    evaluated in FP64 with torch (index_add) from the kernel's own face ids and uvw."""
    v, f = S.jittered_sphere(6, seed=0)
    pts = S.normal_points(1 << 16, 0.4, seed=1)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu).requires_grad_(True)
    Vt = torch.from_numpy(v).to(gpu).requires_grad_(True)
    d, fid, uvw = cubvh_amd.udf(bvh, P, Vt)
    g = torch.linspace(-1, 1, len(pts), device=gpu)
    (d * g).sum().backward()
    Vd = Vt.detach().double()
    F = torch.from_numpy(f).to(gpu).long()
    tri = F[fid]
    w = uvw.double()
    c = (Vd[tri] * w[..., None]).sum(1)
    diff = P.detach().double() - c
    nrm = diff / diff.norm(dim=1, keepdim=True)
    ref_p = g.double()[:, None] * nrm
    ref_v = torch.zeros_like(Vd)
    for k in range(3):
        ref_v.index_add_(0, tri[:, k], -(g.double() * w[:, k])[:, None] * nrm)
    dd = diff.norm(dim=1, keepdim=True)
    tol_p = 1e-5 * g.double().abs()[:, None] * (1 + 1e-2 / dd) + 1e-7
    assert ((P.grad.double() - ref_p).abs() <= tol_p).all()
    tol_v = torch.zeros_like(Vd)
    for k in range(3):
        tol_v.index_add_(0, tri[:, k], (1e-5 * (g.double().abs() * w[:, k].abs())[:, None] * (1 + 1e-2 / dd)).expand(-1, 3))
    assert ((Vt.grad.double() - ref_v).abs() <= tol_v + 1e-6).all()
    # forward values unchanged by the autograd wrapper
    d0, f0, u0 = bvh.unsigned_distance(P.detach(), return_uvw=True)
    assert torch.equal(d0, d.detach()) and torch.equal(f0, fid) and torch.equal(u0, uvw)


def test_backward_bit_exact_vs_oracle_and_repeatable(gpu, oracle, record):
    """The differentiable-UDF workload (jittered level-6 sphere, 2^20 points N(0, 0.4^2),
    loss mean(d^2)): point and vertex gradients equal the oracle's restatement bit for bit
    (same operations, same summation order: per face its queries by ascending index in
    chunks of 32 whose sums are reduced in runs of 32, per vertex its corners 3 f + k
    ascending; this workload has faces with thousands of queries, i.e. two levels of
    runs), and two calls give the same bits."""
    import ctypes

    from cubvh_amd import _lib

    v, f = S.jittered_sphere(6)
    n = 1 << 20
    pts = S.normal_points(n, 0.4, seed=0)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    g = d * (2.0 / n)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    outs = []
    for _ in range(2):
        gp = torch.empty_like(P)
        gv = torch.zeros_like(Vt)
        _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                          Vt.data_ptr(), Vt.shape[0], Ft.data_ptr(), Ft.shape[0], gp.data_ptr(),
                                          gv.data_ptr(), s))
        outs.append((gp, gv))
    # vertex gradients only (no point gradients): the same vertex gradients
    gv_only = torch.zeros_like(Vt)
    _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                      Vt.data_ptr(), Vt.shape[0], Ft.data_ptr(), Ft.shape[0], None,
                                      gv_only.data_ptr(), s))
    torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    assert torch.equal(gv_only, outs[0][1])
    ogp, ogv = oracle.udf_backward(v, f, pts, g.cpu().numpy(), fid.cpu().numpy(), uvw.cpu().numpy())
    np.testing.assert_array_equal(outs[0][0].cpu().numpy(), ogp)
    np.testing.assert_array_equal(outs[0][1].cpu().numpy(), ogv)
    record("backward_level6_2p20", queries=n, vertices=int(len(v)),
           max_abs_vertex_grad=float(np.abs(ogv).max()), point_grad_equal=True, vertex_grad_equal=True)


def test_vertex_corner_table_and_backward_with_it(gpu, oracle):
    """cubvh_vertex_corners equals a stable numpy argsort of the corners by vertex (a
    corner with an out-of-range index goes last and belongs to no vertex), and the
    backward with that table gives the same bits as the one that builds it per call."""
    import ctypes

    from cubvh_amd import _lib

    v, f = S.jittered_sphere(4, seed=2)
    nv = len(v)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    fbad = f.copy()
    fbad[7, 1] = nv + 5  # out of range: the table skips it
    for faces in (f, fbad):
        Ft = torch.from_numpy(faces).to(gpu)
        corners = torch.empty(3 * len(faces), dtype=torch.int32, device=gpu)
        begin = torch.empty(nv, dtype=torch.int32, device=gpu)
        end = torch.empty(nv, dtype=torch.int32, device=gpu)
        _lib.check(lib.cubvh_vertex_corners(gpu.index, Ft.data_ptr(), len(faces), nv, corners.data_ptr(),
                                            begin.data_ptr(), end.data_ptr(), s))
        torch.cuda.synchronize()
        keys = faces.reshape(-1).astype(np.int64)
        keys = np.where((keys >= 0) & (keys < nv), keys, nv)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(corners.cpu().numpy(), order.astype(np.int32))
        sk = keys[order]
        np.testing.assert_array_equal(begin.cpu().numpy(), np.searchsorted(sk, np.arange(nv), "left"))
        np.testing.assert_array_equal(end.cpu().numpy(), np.searchsorted(sk, np.arange(nv), "right"))
    # the backward through the table equals the one building it, bit for bit
    n = 1 << 16
    pts = S.normal_points(n, 0.4, seed=5)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    g = torch.linspace(-1, 1, n, device=gpu)
    corners = torch.empty(3 * len(f), dtype=torch.int32, device=gpu)
    begin = torch.empty(nv, dtype=torch.int32, device=gpu)
    end = torch.empty(nv, dtype=torch.int32, device=gpu)
    _lib.check(lib.cubvh_vertex_corners(gpu.index, Ft.data_ptr(), len(f), nv, corners.data_ptr(), begin.data_ptr(),
                                        end.data_ptr(), s))
    outs = []
    for with_table in (False, True):
        gp = torch.empty_like(P)
        gv = torch.zeros_like(Vt)
        args = [gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(), Vt.data_ptr(), nv,
                Ft.data_ptr(), len(f), gp.data_ptr(), gv.data_ptr()]
        if with_table:
            _lib.check(lib.cubvh_udf_backward_with_corners(*args, corners.data_ptr(), begin.data_ptr(),
                                                           end.data_ptr(), s))
        else:
            _lib.check(lib.cubvh_udf_backward(*args, s))
        outs.append((gp, gv))
    torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    assert outs[0][1].abs().sum() > 0
    ogp, ogv = oracle.udf_backward(v, f, pts, g.cpu().numpy(), fid.cpu().numpy(), uvw.cpu().numpy())
    np.testing.assert_array_equal(outs[1][1].cpu().numpy(), ogv)


def test_backward_one_face_takes_almost_every_query(gpu, oracle):
    """Skewed faces: one large triangle under 2^17 points (4,096 chunks of one face,
    reduced in three levels of runs of 32) and ten small triangles far away; gradients
    bit-exact against the oracle."""
    import ctypes

    from cubvh_amd import _lib

    rng = np.random.default_rng(11)
    big = np.array([[-2.0, -2.0, 0.0], [4.0, -2.0, 0.0], [-2.0, 4.0, 0.0]], np.float32)
    far = (rng.uniform(-0.01, 0.01, size=(10, 3, 3)) + np.array([5.0, 5.0, 5.0])).reshape(-1, 3).astype(np.float32)
    v = np.concatenate([big, far]).astype(np.float32)
    f = np.arange(len(v), dtype=np.int32).reshape(-1, 3)
    n = 1 << 17
    pts = np.concatenate([rng.uniform(-1, 1, size=(n, 2)), rng.uniform(0.05, 0.5, size=(n, 1))], axis=1).astype(np.float32)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    assert (fid == 0).float().mean() > 0.99
    g = torch.linspace(0.5, 1.5, n, device=gpu)
    gp = torch.empty_like(P)
    gv = torch.zeros_like(Vt)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                      Vt.data_ptr(), len(v), Ft.data_ptr(), len(f), gp.data_ptr(), gv.data_ptr(), s))
    torch.cuda.synchronize()
    ogp, ogv = oracle.udf_backward(v, f, pts, g.cpu().numpy(), fid.cpu().numpy(), uvw.cpu().numpy())
    np.testing.assert_array_equal(gp.cpu().numpy(), ogp)
    np.testing.assert_array_equal(gv.cpu().numpy(), ogv)


def test_backward_face_sum_tree_levels(gpu, oracle):
    """Faces holding exactly 1,024, 1,025 and 32,769 queries: 32, 33 and 1,025 chunks of
    32, i.e. sum trees of one, two and three levels of runs of 32 (the call itself runs
    three levels, its largest face could hold all 34,818 queries); gradients bit-exact
    against the oracle."""
    import ctypes

    from cubvh_amd import _lib

    rng = np.random.default_rng(5)
    big = np.array([[-2.0, -2.0, 0.0], [4.0, -2.0, 0.0], [-2.0, 4.0, 0.0]], np.float32)
    shifts = np.array([[0.0, 0.0, 0.0], [20.0, 0.0, 0.0], [40.0, 0.0, 0.0]], np.float32)
    far = (rng.uniform(-0.01, 0.01, size=(10, 3, 3)) + np.array([5.0, 5.0, 5.0])).reshape(-1, 3)
    v = np.concatenate([big + s for s in shifts] + [far]).astype(np.float32)
    f = np.arange(len(v), dtype=np.int32).reshape(-1, 3)
    counts = (1024, 1025, 32769)
    pts = np.concatenate([
        np.concatenate([rng.uniform(-1, 1, size=(c, 2)), rng.uniform(0.05, 0.5, size=(c, 1))], axis=1) + shifts[k]
        for k, c in enumerate(counts)
    ]).astype(np.float32)
    pts = pts[rng.permutation(len(pts))]
    n = len(pts)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    assert [int((fid == k).sum()) for k in range(3)] == list(counts)
    g = torch.linspace(0.25, 2.0, n, device=gpu)
    gp = torch.empty_like(P)
    gv = torch.zeros_like(Vt)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                      Vt.data_ptr(), len(v), Ft.data_ptr(), len(f), gp.data_ptr(), gv.data_ptr(), s))
    torch.cuda.synchronize()
    ogp, ogv = oracle.udf_backward(v, f, pts, g.cpu().numpy(), fid.cpu().numpy(), uvw.cpu().numpy())
    np.testing.assert_array_equal(gp.cpu().numpy(), ogp)
    np.testing.assert_array_equal(gv.cpu().numpy(), ogv)
    assert np.abs(ogv[:9, 2]).min() > 0  # every corner of the three faces takes a share


def test_backward_edge_inputs_bit_exact(gpu, oracle):
    """Backward on the outputs of edge inputs: points on vertices and face centroids
    (distance 0: the zero subgradient), points with NaN or infinite coordinates (face 0,
    NaN uvw: NaN point gradients, nothing added to the vertices) and a mesh with
    zero-area faces; point and vertex gradients equal the oracle's bit for bit."""
    import ctypes

    from cubvh_amd import _lib

    v, f = S.jittered_sphere(3, seed=2)
    v = v.astype(np.float32)
    f = f.astype(np.int32).copy()
    rng = np.random.default_rng(51)
    idx = rng.permutation(len(f))
    f[idx[:20], 1] = f[idx[:20], 0]  # segments
    f[idx[20:30], 1:] = f[idx[20:30], :1]  # points
    tri = v[f[rng.integers(0, len(f), 300)]]
    p = np.concatenate([v[rng.integers(0, len(v), 300)], tri.mean(axis=1), S.normal_points(3000, 0.4, seed=52)])
    p = p.astype(np.float32)
    p[::17, 1] = np.nan
    p[5::23, 0] = np.inf
    p = p[rng.permutation(len(p))]
    n = len(p)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(p).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    g = torch.linspace(0.5, 1.5, n, device=gpu)
    gp = torch.empty_like(P)
    gv = torch.zeros_like(Vt)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                      Vt.data_ptr(), len(v), Ft.data_ptr(), len(f), gp.data_ptr(), gv.data_ptr(), s))
    torch.cuda.synchronize()
    ogp, ogv = oracle.udf_backward(v, f, p, g.cpu().numpy(), fid.cpu().numpy(), uvw.cpu().numpy())
    np.testing.assert_array_equal(gp.cpu().numpy(), ogp)
    np.testing.assert_array_equal(gv.cpu().numpy(), ogv)
    assert np.isfinite(ogv).all() and np.isnan(ogp).any() and (ogp == 0).all(axis=1).any()


def test_backward_with_out_of_range_corner_table_entries(gpu):
    """A caller's vertex -> corners table whose ranges run past the mesh's corners and
    whose entries are all out of range: vertex_sums skips them (no vertex gradient, no
    read outside the table), and the point gradients are those of the plain call."""
    import ctypes

    from cubvh_amd import _lib

    v, f = S.jittered_sphere(2, seed=3)
    pts = S.normal_points(4096, 0.4, seed=4)
    n = len(pts)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(pts).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    g = torch.ones(n, device=gpu)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    corners = torch.full((3 * len(f),), -1, dtype=torch.int32, device=gpu)  # 0xffffffff
    begin = torch.zeros(len(v), dtype=torch.int32, device=gpu)
    end = torch.full((len(v),), 2**31 - 1, dtype=torch.int32, device=gpu)
    gp = torch.empty_like(P)
    gv = torch.zeros_like(Vt)
    _lib.check(lib.cubvh_udf_backward_with_corners(
        gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(), Vt.data_ptr(), len(v),
        Ft.data_ptr(), len(f), gp.data_ptr(), gv.data_ptr(), corners.data_ptr(), begin.data_ptr(), end.data_ptr(), s))
    gp0 = torch.empty_like(P)
    gv0 = torch.zeros_like(Vt)
    _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                      Vt.data_ptr(), len(v), Ft.data_ptr(), len(f), gp0.data_ptr(), gv0.data_ptr(), s))
    torch.cuda.synchronize()
    assert torch.equal(gp, gp0)
    assert (gv == 0).all() and (gv0 != 0).any()


def test_backward_faces_of_one_chunk_bit_exact(gpu, oracle):
    """Every face under at most 32 queries (a few points just off each face): no face has
    a second chunk, so every sum-tree pass is skipped and each face's total is its one
    chunk sum; gradients equal the oracle's bit for bit."""
    import ctypes

    from cubvh_amd import _lib

    v, f = S.jittered_sphere(3, seed=6)
    tri = v[f].astype(np.float64)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rng = np.random.default_rng(61)
    w = rng.dirichlet([4.0, 4.0, 4.0], size=(3, len(f)))
    p = np.concatenate([np.einsum("fk,fkd->fd", w[i], tri) + 0.004 * (i + 1) * nrm for i in range(3)])
    p = p.astype(np.float32)[rng.permutation(3 * len(f))]
    n = len(p)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.from_numpy(p).to(gpu)
    Vt = torch.from_numpy(v).to(gpu)
    Ft = torch.from_numpy(f).to(gpu)
    d, fid, uvw = bvh.unsigned_distance(P, return_uvw=True)
    assert np.bincount(fid.cpu().numpy(), minlength=len(f)).max() <= 32
    g = torch.linspace(0.5, 1.5, n, device=gpu)
    gp = torch.empty_like(P)
    gv = torch.zeros_like(Vt)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    _lib.check(lib.cubvh_udf_backward(gpu.index, n, g.data_ptr(), P.data_ptr(), fid.data_ptr(), uvw.data_ptr(),
                                      Vt.data_ptr(), len(v), Ft.data_ptr(), len(f), gp.data_ptr(), gv.data_ptr(), s))
    torch.cuda.synchronize()
    ogp, ogv = oracle.udf_backward(v, f, p, g.cpu().numpy(), fid.cpu().numpy(), uvw.cpu().numpy())
    np.testing.assert_array_equal(gp.cpu().numpy(), ogp)
    np.testing.assert_array_equal(gv.cpu().numpy(), ogv)
    assert (ogv != 0).any()
