"""GPU floodfill (union-find kernels) against the oracle's BFS and scipy: bit-exact."""
import numpy as np
import pytest
import torch

import cubvh_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(1, 16, 16, 16), (3, 9, 9, 9), (2, 5, 7, 11), (1, 1, 1, 64), (4, 32, 8, 2)])
def test_floodfill_random_grids(gpu, oracle, shape):
    rng = np.random.default_rng(sum(shape))
    for p in (0.1, 0.3, 0.5, 0.8):
        g = rng.random(shape) < p
        lab = cubvh_amd.floodfill(torch.from_numpy(g).to(gpu))
        assert lab.dtype == torch.int32 and tuple(lab.shape) == shape
        np.testing.assert_array_equal(lab.cpu().numpy(), oracle.floodfill_bfs(g))


def test_floodfill_cubic_equals_reference_loop(gpu, oracle):
    rng = np.random.default_rng(1)
    g = rng.random((2, 24, 24, 24)) < 0.45
    lab = cubvh_amd.floodfill(torch.from_numpy(g).to(gpu)).cpu().numpy()
    ref, it = oracle.floodfill_ref(g)
    assert it > 0
    np.testing.assert_array_equal(lab, ref)


def test_floodfill_3d_input_and_edge_cases(gpu, oracle):
    g = torch.zeros(8, 8, 8, dtype=torch.bool)
    lab = cubvh_amd.floodfill(g)  # 3-D CPU input: moved, unsqueezed, squeezed back
    assert lab.shape == (8, 8, 8) and lab.is_cuda
    assert (lab == 0).all()
    lab = cubvh_amd.floodfill(torch.ones(2, 4, 4, 4, dtype=torch.bool, device=gpu))
    assert (lab == -1).all()
    g = torch.zeros(1, 10, 10, 10, dtype=torch.bool)
    g[0, 5] = True
    lab = cubvh_amd.floodfill(g.to(gpu)).cpu()
    assert set(lab[0, :5].unique().tolist()) == {0} and set(lab[0, 6:].unique().tolist()) == {600}


def test_floodfill_occupancy_of_distance_field(gpu, oracle):
    """Occupancy udf < 2/res of a perturbed sphere at 96^3: inside and outside are two
    components; the labels equal the BFS."""
    from cubvh_amd import synthetic as S

    res = 96
    v, f, _ = S.perturbed_sphere(5, seed=0)
    bvh = cubvh_amd.cuBVH(v, f)
    pts = torch.from_numpy(S.grid_points(res)).to(gpu)
    d, _, _ = bvh.unsigned_distance(pts)
    occ = (d < 2.0 / res).reshape(1, res, res, res)
    lab = cubvh_amd.floodfill(occ).cpu().numpy()
    ref = oracle.floodfill_bfs(occ.cpu().numpy())
    np.testing.assert_array_equal(lab, ref)
    assert len(np.unique(lab[lab >= 0])) >= 2
