"""Pins of the oracle's geometry (CPU only).

The FP64 brute force is pinned by closed forms on the reference's own fixture (the
dodecahedron) and by an independent numpy closest-point computation; the fp32
restatement of the reference primitives is pinned against the FP64 brute force and
against special cases whose answers are exact in fp32.
"""
import math

import numpy as np
import pytest

# Inradius and midradius of a regular dodecahedron with circumradius 1.
PHI = (1 + 5 ** 0.5) / 2
DODECA_INRADIUS = ((5 + 2 * 5 ** 0.5) / 15) ** 0.5
DODECA_MIDRADIUS = PHI / 3 ** 0.5


def test_dodecahedron_constants_are_consistent(dodecahedron):
    v, f = dodecahedron
    # every face plane is at the inradius, every edge midpoint at the midradius
    a, b, c = (v[f[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    plane = np.abs((n * a).sum(1))
    np.testing.assert_allclose(plane, DODECA_INRADIUS, atol=2e-6)
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    mids = np.linalg.norm(0.5 * (v[edges[:, 0]] + v[edges[:, 1]]).astype(np.float64), axis=1)
    # pentagon edges are at the midradius; the diagonals splitting pentagons are inside
    assert abs(mids.max() - DODECA_MIDRADIUS) < 2e-6
    assert abs(DODECA_INRADIUS - 0.7946544722917661) < 1e-12


def test_f64_closed_forms_on_dodecahedron(oracle, dodecahedron):
    v, f = dodecahedron
    vd = v.astype(np.float64)
    # origin: distance = inradius
    d, _, _, _ = oracle.udf_f64(v, f, np.zeros((1, 3)))
    assert abs(d[0] - DODECA_INRADIUS) < 2e-6
    # along vertex directions at radius t: nearest feature is the vertex, distance t - 1
    for t in (1.25, 2.0, 5.0):
        d, _, cp, _ = oracle.udf_f64(v, f, vd * t)
        np.testing.assert_allclose(d, t - 1.0, atol=2e-6)
        np.testing.assert_allclose(cp, vd, atol=2e-6)
    # along face-centre directions (the pentagon centres) at radius t: t - inradius
    a, b, c = (vd[f[:, k]] for k in range(3))
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    for t in (0.3, 1.1, 3.0):
        d, _, _, _ = oracle.udf_f64(v, f, n * t)
        np.testing.assert_allclose(d, abs(t - DODECA_INRADIUS), atol=3e-6)


def _closest_numpy(V, F, p):
    """Independent closest point: minimise over the plane projection (if it falls in
    the triangle) and the three clamped edge projections, in float64."""
    V = V.astype(np.float64)
    best = np.full(len(p), np.inf)
    bf = np.zeros(len(p), np.int64)
    for fi, (ia, ib, ic) in enumerate(F):
        a, b, c = V[ia], V[ib], V[ic]
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        q = p - ((p - a) @ n)[:, None] * n
        # inside test with signed areas
        s1 = np.cross(b - a, q - a) @ n
        s2 = np.cross(c - b, q - b) @ n
        s3 = np.cross(a - c, q - c) @ n
        inside = (s1 >= 0) & (s2 >= 0) & (s3 >= 0)
        cand = np.where(inside, np.linalg.norm(p - q, axis=1), np.inf)
        for x, y in ((a, b), (b, c), (c, a)):
            t = np.clip(((p - x) @ (y - x)) / ((y - x) @ (y - x)), 0, 1)
            e = x + t[:, None] * (y - x)
            cand = np.minimum(cand, np.linalg.norm(p - e, axis=1))
        upd = cand < best
        best = np.where(upd, cand, best)
        bf = np.where(upd, fi, bf)
    return best, bf


def test_f64_matches_independent_numpy(oracle, dodecahedron):
    v, f = dodecahedron
    rng = np.random.default_rng(7)
    p = (rng.normal(size=(3000, 3)) * 1.2).astype(np.float32).astype(np.float64)  # the oracle takes fp32 points
    d, fi, cp, gap = oracle.udf_f64(v, f, p)
    d2, _ = _closest_numpy(v, f, p)
    np.testing.assert_allclose(d, d2, atol=1e-12, rtol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(p - cp, axis=1), d, atol=1e-12)
    assert (gap >= 0).all()


def test_fp32_triangle_special_cases(oracle):
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    # above the interior: plane branch, exact in fp32
    p = np.array([[0.25, 0.25, 0.5], [0.125, 0.5, -0.25], [0.25, 0.25, 0.0]], np.float32)
    np.testing.assert_array_equal(oracle.tri_dist2(tri, p), np.array([0.25, 0.0625, 0.0], np.float32))
    # beyond a vertex and beyond an edge
    p = np.array([[-1.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.5, -2.0, 0.0], [1.0, 1.0, 0.0]], np.float32)
    np.testing.assert_allclose(oracle.tri_dist2(tri, p), [1.0, 1.0, 4.0, 0.5], rtol=1e-7)
    # closest point and barycentrics
    cp, uvw = oracle.tri_closest(tri, np.array([[0.25, 0.25, 3.0], [2.0, 2.0, 0.0], [-1, -1, 0]], np.float32))
    np.testing.assert_allclose(cp, [[0.25, 0.25, 0.0], [0.5, 0.5, 0.0], [0, 0, 0]], atol=1e-7)
    np.testing.assert_allclose(uvw, [[0.5, 0.25, 0.25], [0.0, 0.5, 0.5], [1, 0, 0]], atol=1e-7)
    # ray: straight down onto the interior hits at t = 2; outside misses; behind misses
    ro = np.array([[0.25, 0.25, 2.0], [2.0, 2.0, 2.0], [0.25, 0.25, -1.0]], np.float32)
    rd = np.array([[0, 0, -1], [0, 0, -1], [0, 0, -1]], np.float32)
    np.testing.assert_array_equal(oracle.tri_ray(tri, ro, rd), np.array([2.0, 1e6, 1e6], np.float32))
    # parallel ray: safe_divide clamps 1/0 to 1e6 and the hit is rejected by u, v
    assert oracle.tri_ray(tri, np.array([[0.2, 0.2, 1.0]]), np.array([[1.0, 0, 0]]))[0] == np.float32(1e6)


def test_safe_divide(oracle):
    assert oracle.safe_divide(1.0, 2.0) == 0.5
    assert oracle.safe_divide(1.0, 1e-7) == pytest.approx(1e6)
    assert oracle.safe_divide(1.0, -1e-7) == pytest.approx(-1e6)
    assert oracle.safe_divide(1.0, 0.0) == pytest.approx(-1e6)  # den <= 0 takes the negative branch
    assert oracle.safe_divide(2.0, -4.0) == -0.5


def test_fp32_vs_f64_small_workload(oracle, ico4):
    """fp32 brute force within 1e-5 (relative, floor 1) of FP64; clear winners agree on
    the face and on the closest point rebuilt from face + uvw (the reference test's
    check, test/unsigned_distance.py:121-139)."""
    from cubvh_amd import synthetic as S

    v, f = ico4
    p = S.uniform_points(65536, seed=0)
    d, fi, uvw = oracle.udf_brute(v, f, p)
    d64, f64, cp64, gap = oracle.udf_f64(v, f, p)
    err = np.abs(d.astype(np.float64) - d64)
    assert (err <= 1e-5 * np.maximum(d64, 1.0)).all(), err.max()
    clear = gap > 1e-5 * np.maximum(d64, 1.0)
    assert clear.sum() > 10000
    assert (fi[clear] == f64[clear]).all()
    vd = v.astype(np.float64)
    cp = vd[f[fi, 0]] * uvw[:, [0]] + vd[f[fi, 1]] * uvw[:, [1]] + vd[f[fi, 2]] * uvw[:, [2]]
    np.testing.assert_allclose(cp[clear], cp64[clear], atol=1e-5)
    # tie class: the chosen face's closest point is at the minimum distance
    np.testing.assert_allclose(np.linalg.norm(p - cp, axis=1), d64, atol=1e-5)


def test_fp32_rays_vs_f64(oracle, ico4):
    from cubvh_amd import synthetic as S

    v, f = ico4
    ro, rd = S.sphere_rays(16384, seed=0)
    pos, fi, t = oracle.ray_brute(v, f, ro, rd)
    t64, f64, amb = oracle.ray_f64(v, f, ro, rd, tol=1e-5)
    ok = ~amb
    assert ok.sum() > 15000
    assert ((fi >= 0) == (f64 >= 0))[ok].all()
    hit = ok & (f64 >= 0)
    assert (fi[hit] == f64[hit]).all()
    np.testing.assert_allclose(t[hit], t64[hit], rtol=1e-5)
    assert (t[fi < 0] == 1000.0).all()
    np.testing.assert_allclose(pos, ro + t[:, None] * rd, rtol=1e-6, atol=1e-6)


def test_dodecahedron_reference_check(oracle, dodecahedron):
    """The reference's own test (test/unsigned_distance.py:93-139) restated: N=1000
    randn points, distance and reconstructed closest point within atol 1e-5."""
    v, f = dodecahedron
    p = np.random.default_rng(0).normal(size=(1000, 3)).astype(np.float32)
    d, fi, uvw = oracle.udf_brute(v, f, p)
    d64, _, cp64, gap = oracle.udf_f64(v, f, p)
    np.testing.assert_allclose(d, d64, atol=1e-5)
    vd = v.astype(np.float64)
    cp = vd[f[fi, 0]] * uvw[:, [0]] + vd[f[fi, 1]] * uvw[:, [1]] + vd[f[fi, 2]] * uvw[:, [2]]
    clear = gap > 1e-5
    np.testing.assert_allclose(cp[clear], cp64[clear], atol=1e-5)


def test_winding_number_inside_outside(oracle, ico4):
    v, f = ico4
    p = np.array([[0, 0, 0], [0.1, 0.2, -0.1], [0.9, 0, 0], [0, 0, 2.0]], np.float32)
    w = oracle.winding_f64(v, f, p)
    np.testing.assert_allclose(w, [1, 1, 0, 0], atol=1e-9)


def test_planted_mistake_is_caught(oracle, ico4):
    """A plausible slip (dropping the plane-branch division by |n|^2) would be caught by
    the FP64 comparison above: emulate it on the oracle's fp32 output and check."""
    from cubvh_amd import synthetic as S

    v, f = ico4
    p = S.uniform_points(2048, seed=3)
    d, fi, _ = oracle.udf_brute(v, f, p)
    a, b, c = (v[f[fi, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, a - c)
    wrong = np.sqrt(d.astype(np.float64) ** 2 * (n * n).sum(1))  # distance times |n|
    d64, _, _, _ = oracle.udf_f64(v, f, p)
    assert not (np.abs(wrong - d64) <= 1e-5 * np.maximum(d64, 1.0)).all()
