"""Pins of the oracle's reference-BVH restatement, PCG32, stab directions and
floodfill (CPU only)."""
import json

import numpy as np
import pytest

from conftest import GOLDEN


def test_refbvh_shape(oracle, ico4):
    v, f = ico4
    R = oracle.RefBvh(v, f)
    nodes, tris = R.export()
    # 4-ary tree with leaves <= 8 over 5120 faces: (4^k - 1)/3 nodes for a full tree
    assert nodes.shape == (1365, 9)
    ids = tris.view(np.int64)[:, 5]
    assert sorted(ids.tolist()) == list(range(len(f)))
    left, right = nodes[:, 6], nodes[:, 7]
    leaf = left < 0
    sizes = (-right[leaf] - 1) - (-left[leaf] - 1)
    assert sizes.max() <= 8 and sizes.min() >= 1 and sizes.sum() == len(f)
    assert ((right[~leaf] - left[~leaf]) == 4).all()


def test_refbvh_udf_equals_brute_force(oracle, ico4):
    from cubvh_amd import synthetic as S

    v, f = ico4
    p = S.uniform_points(20000, seed=1)
    d, fi, uvw = oracle.udf_brute(v, f, p)
    R = oracle.RefBvh(v, f)
    dc, fc, uc, counts = R.udf(p, tie_mode=1)
    np.testing.assert_array_equal(dc, d)
    np.testing.assert_array_equal(fc, fi)
    np.testing.assert_array_equal(uc, uvw)
    assert counts[0] > 0 and counts[1] > 0
    # the reference tie rule (last visited wins) changes only the face of tied queries
    d0, f0, _, counts0 = R.udf(p, tie_mode=0)
    np.testing.assert_array_equal(d0, d)
    d64, f64, _, gap = oracle.udf_f64(v, f, p)
    clear = gap > 1e-5 * np.maximum(d64, 1.0)
    assert (f0[clear] == fi[clear]).all()


def test_refbvh_canonical_equals_brute_force_level6(oracle):
    """Deeper tree with many exact ties (shared vertices of a jittered sphere)."""
    from cubvh_amd import synthetic as S

    v, f = S.jittered_sphere(6)
    p = S.normal_points(3000, 0.4, seed=2)
    d, fi, uvw = oracle.udf_brute(v, f, p)
    dc, fc, uc, _ = oracle.RefBvh(v, f).udf(p, tie_mode=1)
    np.testing.assert_array_equal(dc, d)
    np.testing.assert_array_equal(fc, fi)
    np.testing.assert_array_equal(uc, uvw)


def test_refbvh_rays_equal_brute_force(oracle, ico4):
    from cubvh_amd import synthetic as S

    v, f = ico4
    ro, rd = S.sphere_rays(8192, seed=2)
    pos, fi, t = oracle.ray_brute(v, f, ro, rd)
    R = oracle.RefBvh(v, f)
    pc, fc, tc, counts = R.ray(ro, rd, tie_mode=1)
    np.testing.assert_array_equal(tc, t)
    np.testing.assert_array_equal(fc, fi)
    np.testing.assert_array_equal(pc, pos)
    p0, f0, t0, _ = R.ray(ro, rd, tie_mode=0)
    np.testing.assert_array_equal(t0, t)


def test_nan_and_inf_queries_find_nothing(oracle, ico4):
    """NaN or infinite coordinates: the restated reference walks (whose `<=` tests are
    false for them, bvh.cu:300-353, 250-298) and the brute force agree on the not-found
    answers: face 0 at distance 0 for a point, a miss at t = 1000 for a ray."""
    from cubvh_amd import synthetic as S

    v, f = ico4
    p = S.uniform_points(500, seed=21)
    p[::7, 0] = np.nan
    p[3::11, 1] = np.inf
    p[5::13, 2] = -np.inf
    bad = ~np.isfinite(p).all(axis=1)
    d, fi, uvw = oracle.udf_brute(v, f, p)
    R = oracle.RefBvh(v, f)
    wd, wf, wu = R.udf(p, tie_mode=1)[:3]
    np.testing.assert_array_equal(wd, d)
    np.testing.assert_array_equal(wf, fi)
    np.testing.assert_array_equal(wu, uvw)
    assert (d[bad] == 0).all() and (fi[bad] == 0).all() and (d[~bad] > 0).all()
    ro, rd = S.sphere_rays(500, seed=22)
    ro[::9, 1] = np.nan
    rd[4::10, 0] = np.nan
    ro[2::15, 2] = np.inf
    rbad = ~(np.isfinite(ro).all(axis=1) & np.isfinite(rd).all(axis=1))
    pos, fr, t = oracle.ray_brute(v, f, ro, rd)
    pc, fc, tc, _ = R.ray(ro, rd, tie_mode=1)
    np.testing.assert_array_equal(tc, t)
    np.testing.assert_array_equal(fc, fr)
    np.testing.assert_array_equal(pc, pos)
    assert (fr[rbad] == -1).all() and (t[rbad] == 1000.0).all() and (fr[~rbad] >= 0).any()


def test_sdf_signs(oracle, ico4):
    """Watertight and raystab signs agree with the FP64 winding number away from the
    surface; the reference-BVH restatement agrees with the brute force."""
    from cubvh_amd import synthetic as S

    v, f = ico4
    p = S.uniform_points(3000, seed=4)
    w = oracle.winding_f64(v, f, p)
    inside = w > 0.5
    dw, fw, _ = oracle.sdf_brute(v, f, p, 0)
    dr, fr, _ = oracle.sdf_brute(v, f, p, 1)
    du, _, _ = oracle.udf_brute(v, f, p)
    np.testing.assert_array_equal(np.abs(dw), du)
    np.testing.assert_array_equal(np.abs(dr), du)
    far = du > 1e-4
    assert ((dw < 0) == inside)[far].all()
    assert ((dr < 0) == inside)[far].all()
    R = oracle.RefBvh(v, f)
    dw2, _, _, cw = R.sdf(p, 0)
    dr2, _, _, cr = R.sdf(p, 1)
    assert ((dw2 < 0) == (dw < 0))[far].all()
    np.testing.assert_array_equal(dr2, dr)
    assert cw[2] > 0 and cr[2] > 0


def test_pcg32_known_answers(oracle):
    kat = json.loads((GOLDEN / "pcg32_kat.json").read_text())
    for adv, vals in kat.items():
        np.testing.assert_array_equal(oracle.pcg32_uints(int(adv), 8), np.array(vals, np.uint32))


def test_pcg32_jump_equals_stepping(oracle):
    s = oracle.pcg32_uints(0, 3000)
    for k in (1, 2, 17, 1000, 2999 - 8):
        np.testing.assert_array_equal(oracle.pcg32_uints(k, 1), s[k : k + 1])
    # per-query offsets: query i jumps 2i and draws two floats from the top 23 bits
    off = oracle.pcg32_offsets(1000)
    u = oracle.pcg32_uints(0, 2000)
    expect = ((u >> 9).astype(np.float64) / 2.0 ** 23).astype(np.float32).reshape(1000, 2)
    np.testing.assert_array_equal(off, expect)


def test_pcg32_equals_reference_generator(oracle):
    """The oracle's generator against the reference's own (include/gpu/pcg32.h compiled
    from its header into oracle/_ref): outputs after jumps, and the raystab offsets of
    every query index of a 2^22-point call (src/bvh.cu:667-668)."""
    if not oracle.ref_pcg32_available():
        pytest.skip("oracle/_ref/libref_pcg32.so not built (reference sources absent)")
    for adv in (0, 1, 2, 17, 1000, 65537, 1 << 32, (1 << 33) + 5, (1 << 40) + 3, (1 << 62) + 11, -1, -1000):
        np.testing.assert_array_equal(oracle.pcg32_uints(adv, 64), oracle.ref_pcg32_uints(adv, 64))
    n = 1 << 22
    np.testing.assert_array_equal(oracle.pcg32_offsets(n), oracle.ref_pcg32_offsets(n))


def test_pcg32_state_is_the_textbook_lcg(oracle):
    """The state recurrence is the 64-bit LCG s' = 6364136223846793005 s + c (mod 2^64)
    with the reference's default state and increment (pcg32.h:33-35, 47): stepped here
    with Python integers, compared with the reference generator's jump-ahead, and the
    first outputs follow from the states by the XSH-RR rule."""
    if not oracle.ref_pcg32_available():
        pytest.skip("oracle/_ref/libref_pcg32.so not built (reference sources absent)")
    a, c, m = 6364136223846793005, 0xDA3E39CB94B95BDB, 1 << 64
    s = 0x853C49E6748FEA9B
    states = [s]
    for _ in range(3000):
        s = (a * s + c) % m
        states.append(s)
    for k in (0, 1, 2, 3, 100, 2047, 3000):
        assert oracle.ref_pcg32_state(k) == states[k]

    def xsh_rr(old):
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        r = old >> 59
        return ((x >> r) | (x << ((32 - r) & 31))) & 0xFFFFFFFF

    np.testing.assert_array_equal(oracle.pcg32_uints(0, 40), np.array([xsh_rr(x) for x in states[:40]], np.uint32))


def test_fibonacci_lattice_properties(oracle):
    """Properties of the stab directions (common.h:63-96 with 32 directions) that a
    wrong epsilon, golden ratio or axis would break: consecutive azimuths step by
    minus the golden angle 2 pi (2 - phi) (mod 2 pi); at zero offset the 32 values of
    cos(theta) are equally spaced by 2 / (31 + 2 * 1.33), start at
    1 - 2 * 1.33 / (31 + 2 * 1.33) and are symmetric about 0."""
    golden = (1 + 5 ** 0.5) / 2
    for ox, oy in ((0.0, 0.0), (0.25, 0.6), (0.71, 0.05)):
        d = oracle.fibonacci_dirs32(ox, oy).astype(np.float64)
        az = np.arctan2(d[:, 1], d[:, 0])
        step = np.mod(np.diff(az), 2 * np.pi)
        np.testing.assert_allclose(step, 2 * np.pi - 2 * np.pi * (2 - golden), atol=5e-5)
        # the first azimuth is 2 pi (frac(oy) - 1/2)
        np.testing.assert_allclose(np.mod(az[0] - 2 * np.pi * (oy - 0.5) + np.pi, 2 * np.pi) - np.pi, 0, atol=5e-6)
    ct = oracle.fibonacci_dirs32(0.0, 0.0).astype(np.float64)[:, 2]
    gap = 2 / (31 + 2 * 1.33)
    np.testing.assert_allclose(np.diff(ct), -gap, atol=2e-6)
    np.testing.assert_allclose(ct[0], 1 - 1.33 * gap, atol=2e-6)
    np.testing.assert_allclose(ct + ct[::-1], 0, atol=2e-6)
    # an offset shifts the cos(theta) lattice cyclically by 2 ox (mod 2), in units of the gap
    ct2 = oracle.fibonacci_dirs32(0.5 * gap, 0.0).astype(np.float64)[:, 2]
    np.testing.assert_allclose(ct2[:-1], ct[1:], atol=2e-6)


def test_fibonacci_directions(oracle):
    for ox, oy in ((0.0, 0.0), (0.3, 0.7), (0.999, 0.123)):
        d = oracle.fibonacci_dirs32(ox, oy).astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
        i = np.arange(32, dtype=np.float64)
        eps = np.float32(1.33)
        px = np.mod((i + eps) / (31 + 2 * eps) + ox, 1.0)
        py = np.mod(i / ((1 + 5 ** 0.5) / 2) + oy, 1.0)
        ct = 1 - 2 * px
        phi = 2 * np.pi * (py - 0.5)
        st = np.sqrt(np.maximum(1 - ct * ct, 0))
        np.testing.assert_allclose(d, np.stack([st * np.cos(phi), st * np.sin(phi), ct], 1), atol=2e-5)


def _scipy_labels(grid):
    """Minimum-linear-index labels of the free components via scipy.ndimage.label."""
    from scipy import ndimage

    g = np.asarray(grid, bool)
    out = np.full(g.shape, -1, np.int64)
    offset = 0
    for b in range(g.shape[0]):
        lab, n = ndimage.label(~g[b])
        flat = lab.ravel()
        idx = np.arange(flat.size) + offset
        mins = np.full(n + 1, np.iinfo(np.int64).max)
        np.minimum.at(mins, flat, idx)
        res = np.where(flat > 0, mins[flat], -1)
        out[b] = res.reshape(g.shape[1:])
        offset += flat.size
    return out.astype(np.int32)


@pytest.mark.parametrize("shape", [(1, 9, 9, 9), (3, 6, 6, 6), (2, 5, 7, 4), (1, 1, 1, 13)])
def test_floodfill_bfs_matches_scipy(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    for p in (0.2, 0.45, 0.7):
        g = rng.random(shape) < p
        np.testing.assert_array_equal(oracle.floodfill_bfs(g), _scipy_labels(g))


def test_floodfill_reference_restatement_cubic(oracle):
    rng = np.random.default_rng(5)
    for shape in ((1, 8, 8, 8), (2, 12, 12, 12)):
        for p in (0.3, 0.5):
            g = rng.random(shape) < p
            lab, it = oracle.floodfill_ref(g)
            assert it > 0
            np.testing.assert_array_equal(lab, oracle.floodfill_bfs(g))


def test_floodfill_two_rooms_and_corridor(oracle):
    g = np.zeros((1, 10, 10, 10), bool)
    g[0, 5, :, :] = True  # wall splitting the volume
    lab = oracle.floodfill_bfs(g)
    assert set(np.unique(lab[0, :5]).tolist()) == {0}
    assert set(np.unique(lab[0, 6:]).tolist()) == {600}
    assert (lab[0, 5] == -1).all()
    g[0, 5, 9, 9] = False  # a hole joins the rooms
    lab = oracle.floodfill_bfs(g)
    assert set(np.unique(lab[g == False]).tolist()) == {0}  # noqa: E712


def test_floodfill_reference_serpentine_converges_within_cap(oracle):
    """A one-voxel serpentine corridor of ~136 voxels: the reference's hook +
    pointer-jump loop converges well inside its iteration cap (78 at 16^3) and agrees
    with the BFS."""
    n = 16
    g = np.ones((1, n, n, n), bool)
    for y in range(0, n, 2):
        g[0, 0, y, :] = False
        if y + 1 < n:
            x = n - 1 if (y // 2) % 2 == 0 else 0
            g[0, 0, y + 1, x] = False
    lab, it = oracle.floodfill_ref(g)
    assert 0 < it < 78
    bfs = oracle.floodfill_bfs(g)
    np.testing.assert_array_equal(lab, bfs)
    assert len(np.unique(bfs[g == False])) == 1  # noqa: E712
