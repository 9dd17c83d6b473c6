"""Pins of the oracle's backward (cubvh_oracle.cpp (9)): its point and vertex gradients of
L = sum_i g_i d_i against FP64 central finite differences of the FP64 brute-force
distance (which shares no code with the fp32 restatement), and its bookkeeping against
the hand-derived cases a plausible mistake would break (a dropped query, a swapped
barycentric weight, a sign)."""
import numpy as np
import pytest

from cubvh_amd import synthetic as S


@pytest.fixture(scope="module")
def mesh():
    return S.jittered_sphere(2, sigma=0.01, seed=4)  # 320 faces


def _loss64(oracle, V, F, P, g):
    d, _, _, _ = oracle.udf_f64(V, F, P)
    return float(np.dot(g.astype(np.float64), d))


def test_backward_matches_finite_differences(oracle, mesh):
    V, F = mesh
    rng = np.random.default_rng(0)
    P = rng.normal(0.0, 0.4, size=(256, 3)).astype(np.float32)
    g = rng.uniform(0.5, 1.5, size=256).astype(np.float32)
    _, face, uvw = oracle.udf_brute(V, F, P)
    gp, gv = oracle.udf_backward(V, F, P, g, face, uvw)
    d64, _, _, gap = oracle.udf_f64(V, F, P)
    h = 1e-4
    # points whose nearest face is a clear winner (a kink of d elsewhere)
    clear = np.nonzero(gap > 1e-3)[0][:24]
    assert len(clear) >= 12
    for i in clear:
        for k in range(3):
            Pp, Pm = P.astype(np.float64).copy(), P.astype(np.float64).copy()
            Pp[i, k] += h
            Pm[i, k] -= h
            fd = (_loss64(oracle, V, F, Pp[i:i + 1], g[i:i + 1]) - _loss64(oracle, V, F, Pm[i:i + 1], g[i:i + 1])) / (2 * h)
            assert abs(gp[i, k] - fd) <= 2e-3 * max(1.0, abs(fd)), (i, k, gp[i, k], fd)
    # vertices: central differences of the whole loss, on vertices that some query uses
    used = np.unique(F[face[clear]].reshape(-1))
    for vtx in used[:12]:
        for k in range(3):
            Vp, Vm = V.astype(np.float64).copy(), V.astype(np.float64).copy()
            Vp[vtx, k] += h
            Vm[vtx, k] -= h
            fd = (_loss64(oracle, Vp, F, P[clear], g[clear]) - _loss64(oracle, Vm, F, P[clear], g[clear])) / (2 * h)
            _, gv_c = oracle.udf_backward(V, F, P[clear], g[clear], face[clear], uvw[clear])
            assert abs(gv_c[vtx, k] - fd) <= 2e-3 * max(1.0, abs(fd)), (vtx, k, gv_c[vtx, k], fd)


def test_backward_bookkeeping(oracle):
    """One face z = 0 with vertices a, b, c; a point straight above its centroid at
    height 2 with g = 3: grad_p = (0, 0, 3); every vertex gets -3 w_k (0, 0, 1) with
    w = (1/3, 1/3, 1/3) up to fp32; a second query below the face adds linearly; an
    out-of-range face and a zero g add nothing."""
    V = np.array([[0, 0, 0], [3, 0, 0], [0, 3, 0], [9, 9, 9]], np.float32)  # vertex 3 unused
    F = np.array([[0, 1, 2]], np.int32)
    P = np.array([[1, 1, 2], [1, 1, -1], [5, 5, 5], [1, 1, 2]], np.float32)
    g = np.array([3, 1, 7, 0], np.float32)
    face = np.array([0, 0, 5, 0], np.int64)
    w = np.float32(1.0) / np.float32(3.0)
    uvw = np.array([[w, w, w], [w, w, w], [w, w, w], [w, w, w]], np.float32)
    gp, gv = oracle.udf_backward(V, F, P, g, face, uvw)
    np.testing.assert_allclose(gp[0], [0, 0, 3], atol=1e-6)
    np.testing.assert_allclose(gp[1], [0, 0, -1], atol=1e-6)
    np.testing.assert_array_equal(gp[2], [0, 0, 0])
    np.testing.assert_array_equal(gp[3], [0, 0, 0])
    # vertex k: -w (3 (0,0,1) + 1 (0,0,-1)) = (0, 0, -2/3)
    np.testing.assert_allclose(gv[:3], np.tile([0, 0, -2.0 / 3.0], (3, 1)), atol=1e-6)
    np.testing.assert_array_equal(gv[3], [0, 0, 0])
    # the same call again gives the same bits (a fixed summation order)
    _, gv1 = oracle.udf_backward(V, F, P, g, face, uvw)
    assert np.array_equal(gv1, gv)
