"""Batched forest: results identical to per-mesh cuBVH calls (bit-exact), for every
launch configuration, under HIP-graph replay, and through the backward."""
import numpy as np
import pytest
import torch

import cubvh_amd
from cubvh_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _forest_case(seed=0):
    rng = np.random.default_rng(seed)
    specs = [(3, 0.4, 700), (5, 0.8, 5000), (2, 0.6, 0), (4, 0.5, 65), (6, 0.9, 40000), (3, 0.3, 1), (1, 0.7, 130)]
    bvhs, pts, counts = [], [], []
    for lvl, scale, nq in specs:
        v, f, _ = S.perturbed_sphere(lvl, seed=int(rng.integers(1 << 30)), scale=scale)
        bvhs.append(cubvh_amd.cuBVH(v, f))
        pts.append(rng.uniform(-1, 1, size=(nq, 3)).astype(np.float32))
        counts.append(nq)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return bvhs, np.concatenate(pts), offsets


def _per_mesh(bvhs, pts, offsets, dev):
    d, f, u = [], [], []
    for m, b in enumerate(bvhs):
        p = torch.from_numpy(pts[offsets[m]:offsets[m + 1]]).to(dev)
        if p.shape[0] == 0:
            continue
        dd, ff, uu = b.unsigned_distance(p, return_uvw=True)
        d.append(dd)
        f.append(ff)
        u.append(uu)
    return torch.cat(d), torch.cat(f), torch.cat(u)


@pytest.mark.parametrize("tuning", [(12, 1, 8), (0, 0, 8), (9, 0, 8), (15, 1, 4), (6, 1, 16), (12, 1, 8, 1),
                                    (12, 0, 8, 2), (9, 1, 4, 8), (0, 0, 8, 8), (15,), (15, 0, 8, 1, 16),
                                    (9, 0, 8, 4, 4), (12, 0, 8, 2, 64), (12, 1, 8, 4, 16)])
def test_forest_matches_per_mesh(gpu, tuning):
    bvhs, pts, offsets = _forest_case()
    forest = cubvh_amd.cuBVHForest(bvhs)
    info = forest.info()
    assert info["n_meshes"] == len(bvhs)
    d, f, u = forest.unsigned_distance(torch.from_numpy(pts).to(gpu), torch.from_numpy(offsets), True, tuning)
    rd, rf, ru = _per_mesh(bvhs, pts, offsets, gpu)
    assert torch.equal(d, rd) and torch.equal(f, rf) and torch.equal(u, ru)


def test_forest_vs_oracle_brute_force(gpu, oracle):
    bvhs, pts, offsets = _forest_case(1)
    forest = cubvh_amd.cuBVHForest(bvhs)
    d, f, _ = forest.unsigned_distance(torch.from_numpy(pts).to(gpu), torch.from_numpy(offsets))
    d, f = d.cpu().numpy(), f.cpu().numpy()
    for m, b in enumerate(bvhs):
        sl = slice(offsets[m], offsets[m + 1])
        if offsets[m + 1] == offsets[m]:
            continue
        p = pts[sl][:500]
        od, of, _ = oracle.udf_brute(b._vertices, b._faces, p, return_uvw=False)
        np.testing.assert_array_equal(d[sl][:500], od)
        np.testing.assert_array_equal(f[sl][:500], of)


def test_forest_graph_replays(gpu):
    """Captured once, replayed several times with eager work and new inputs in between;
    every replay equals the eager result."""
    bvhs, pts, offsets = _forest_case(2)
    forest = cubvh_amd.cuBVHForest(bvhs)
    P = torch.from_numpy(pts).to(gpu)
    off = torch.from_numpy(offsets).to(gpu)
    g = cubvh_amd.ForestGraph(forest, P, off, return_uvw=True, tuning=(12, 1, 8))
    for it in range(4):
        newp = P + 0.01 * it
        g.points.copy_(newp)
        d, f, u = g.replay()
        torch.cuda.synchronize()
        ed, ef, eu = forest.unsigned_distance(newp, off, True, (12, 1, 8))
        _ = torch.randn(1000, device=gpu).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(d, ed) and torch.equal(f, ef) and torch.equal(u, eu), it


def test_forest_backward_matches_per_mesh(gpu):
    bvhs, pts, offsets = _forest_case(3)
    forest = cubvh_amd.cuBVHForest(bvhs)
    P = torch.from_numpy(pts).to(gpu).requires_grad_(True)
    vcat, fcat = forest.concatenated_mesh(gpu)
    V = vcat.clone().requires_grad_(True)
    d, _, _ = cubvh_amd.udf_batched(forest, P, torch.from_numpy(offsets), V)
    (d * d).sum().backward()
    # per mesh with the single-mesh autograd path
    gps, gvs = [], []
    for m, b in enumerate(bvhs):
        p = torch.from_numpy(pts[offsets[m]:offsets[m + 1]]).to(gpu).requires_grad_(True)
        v = torch.from_numpy(b._vertices).to(gpu).requires_grad_(True)
        if p.shape[0]:
            dd, _, _ = cubvh_amd.udf(b, p, v)
            (dd * dd).sum().backward()
            gps.append(p.grad)
        gvs.append(v.grad if v.grad is not None else torch.zeros_like(v))
    assert torch.equal(P.grad, torch.cat(gps))
    # vertex sums gathered per mesh in query order: the same bits as the per-mesh calls
    assert torch.equal(V.grad, torch.cat(gvs))


def test_autotuner_picks_a_valid_config(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("CUBVH_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    bvhs, pts, offsets = _forest_case(4)
    forest = cubvh_amd.cuBVHForest(bvhs)
    P = torch.from_numpy(pts).to(gpu)
    best, ms = cubvh_amd.tune_forest(forest, P, torch.from_numpy(offsets), reps=1)
    assert len(ms) >= 4 and all(v > 0 for v in ms.values())
    # the stages ran: every workgroup size was tried with the best first-stage launch, and
    # for a one-wave-per-tile winner every XCD chunk
    assert len(best) == 5 and {int(k.split(",")[3]) for k in ms} == {1, 2, 4, 8}
    if best[1] == 0:
        assert {int(k.split(",")[4]) for k in ms} == set(cubvh_amd.autotune.XCD_CHUNKS)
    best2, _ = cubvh_amd.tune_forest(forest, P, torch.from_numpy(offsets), reps=1)
    assert tuple(best2) == tuple(best)  # served from the cache
    d, f, _ = forest.unsigned_distance(P, torch.from_numpy(offsets), False, best)
    rd, rf, _ = _per_mesh(bvhs, pts, offsets, gpu)
    assert torch.equal(d, rd) and torch.equal(f, rf)


def test_forest_nan_and_inf_queries_match_per_mesh(gpu):
    """Queries with NaN or infinite coordinates in a batch: each gets what the per-mesh
    call gives it (the not-found answer), in the persistent and one-wave-per-tile
    launches, and the other queries are unaffected."""
    bvhs, pts, offsets = _forest_case(2)
    pts[::7, 0] = np.nan
    pts[3::11, 1] = np.inf
    pts[5::13, 2] = -np.inf
    forest = cubvh_amd.cuBVHForest(bvhs)
    rd, rf, ru = (t.cpu().numpy() for t in _per_mesh(bvhs, pts, offsets, gpu))
    bad = ~np.isfinite(pts).all(axis=1)
    assert (rd[bad] == 0).all() and (rd[~bad] > 0).all()
    for tuning in [(12, 1, 8), (15, 0, 8, 1, 64), (0, 0, 8)]:
        d, f, u = forest.unsigned_distance(torch.from_numpy(pts).to(gpu), torch.from_numpy(offsets), True, tuning)
        np.testing.assert_array_equal(d.cpu().numpy(), rd)
        np.testing.assert_array_equal(f.cpu().numpy(), rf)
        np.testing.assert_array_equal(u.cpu().numpy(), ru)


def test_forest_of_random_soups_matches_per_mesh(gpu):
    """A batch of random triangle soups at scales from 1e-3 to 1e3 (one mesh without
    queries): distances, faces and uvw equal the per-mesh calls for three launch shapes,
    and the batched backward's point and vertex gradients equal the per-mesh ones."""
    from test_gpu_queries import _random_soup

    bvhs, pts, counts = [], [], []
    rng = np.random.default_rng(71)
    for k, seed in enumerate((111, 112, 113, 114, 115, 116)):
        v, f, scale, centre = _random_soup(seed)
        bvhs.append(cubvh_amd.cuBVH(v, f))
        nq = 0 if k == 2 else int(rng.integers(1, 6000))
        pts.append((rng.uniform(-1.5, 1.5, size=(nq, 3)) * scale + centre).astype(np.float32))
        counts.append(nq)
    pts = np.concatenate(pts)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    forest = cubvh_amd.cuBVHForest(bvhs)
    rd, rf, ru = _per_mesh(bvhs, pts, offsets, gpu)
    for tuning in [(12, 1, 8), (15, 0, 8, 1, 64), (0, 0, 8, 2)]:
        d, f, u = forest.unsigned_distance(torch.from_numpy(pts).to(gpu), torch.from_numpy(offsets), True, tuning)
        assert torch.equal(d, rd) and torch.equal(f, rf) and torch.equal(u, ru)
    P = torch.from_numpy(pts).to(gpu).requires_grad_(True)
    vcat, _ = forest.concatenated_mesh(gpu)
    V = vcat.clone().requires_grad_(True)
    d, _, _ = cubvh_amd.udf_batched(forest, P, torch.from_numpy(offsets), V)
    (d * d).sum().backward()
    gps, gvs = [], []
    for m, b in enumerate(bvhs):
        p = torch.from_numpy(pts[offsets[m]:offsets[m + 1]]).to(gpu).requires_grad_(True)
        v = torch.from_numpy(b._vertices).to(gpu).requires_grad_(True)
        if p.shape[0]:
            dd, _, _ = cubvh_amd.udf(b, p, v)
            (dd * dd).sum().backward()
            gps.append(p.grad)
        gvs.append(v.grad if v.grad is not None else torch.zeros_like(v))
    assert torch.equal(P.grad, torch.cat(gps))
    assert torch.equal(V.grad, torch.cat(gvs))
