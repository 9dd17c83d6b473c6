"""sparse_marching_cubes on the GPU against the oracle: vertices and triangles
bit-identical (same order, same fp32 positions) on sampled surfaces, random cells,
averaged corners, empty and uncut inputs; a large sphere checked by the closed-surface
invariants.
"""
import numpy as np
import pytest
import torch

import cubvh_amd
from oracle import oracle as O
from test_oracle_smc import cells_of, sphere_field

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _gpu(coords, corners, iso, consistent=False):
    V, T = cubvh_amd.sparse_marching_cubes(torch.from_numpy(coords).to(DEV), torch.from_numpy(corners).to(DEV), iso,
                                           ensure_consistency=consistent)
    assert V.dtype == torch.float32 and T.dtype == torch.int32 and V.device.type == "cuda"
    return V.cpu().numpy(), T.cpu().numpy()


def _check(coords, corners, iso, consistent=False):
    V, T = _gpu(coords, corners, iso, consistent)
    OV, OT = O.smc(coords, corners, iso, consistent)
    np.testing.assert_array_equal(V, OV)
    np.testing.assert_array_equal(T, OT)
    return V, T


@pytest.mark.parametrize("iso", [0.0, 0.7, -1.5])
def test_sphere_equals_oracle(iso):
    coords, corners = cells_of(sphere_field(48), iso)
    _check(coords, corners, iso)


@pytest.mark.parametrize("consistent", [False, True])
def test_random_cells_equal_oracle(consistent):
    rng = np.random.default_rng(11)
    coords = rng.integers(-40, 40, size=(20000, 3)).astype(np.int32)
    coords[-5:] = coords[:5]  # duplicated cells
    corners = rng.normal(size=(20000, 8)).astype(np.float32)
    _check(coords, corners, 0.05, consistent)


@pytest.mark.parametrize("consistent", [False, True])
def test_wide_lattice_span_equals_oracle(consistent):
    """Spans above 1022 lattice steps need 64-bit edge keys."""
    rng = np.random.default_rng(12)
    base = rng.integers(-3000, 3000, size=(400, 3))
    coords = (base[:, None, :] + rng.integers(0, 3, size=(400, 10, 3))).reshape(-1, 3).astype(np.int32)
    corners = rng.normal(size=(len(coords), 8)).astype(np.float32)
    _check(coords, corners, 0.0, consistent)


@pytest.mark.parametrize("span", [1022, 1023])
def test_key_width_boundary(span):
    """A span of 1022 lattice steps still fits 32-bit edge keys, 1023 needs 64-bit."""
    coords, corners = cells_of(sphere_field(12))
    far = coords + np.array([span - 11, 3, 5], np.int32)
    _check(np.concatenate([coords, far]), np.concatenate([corners, corners]), 0.0)


def test_noisy_shared_corners_averaged_equal_oracle():
    coords, corners = cells_of(sphere_field(40))
    corners = corners + np.random.default_rng(2).normal(scale=0.05, size=corners.shape).astype(np.float32)
    perm = np.random.default_rng(3).permutation(len(coords))  # cell order matters for the sums
    _check(coords[perm], corners[perm], 0.0, True)


def test_dense_block_with_every_cell_equals_oracle():
    field = np.random.default_rng(4).normal(size=(17, 13, 9)).astype(np.float32)
    coords, corners = cells_of(field, all_cells=True)
    coords = coords + np.array([1000, -2000, 5], np.int32)
    _check(coords, corners, 0.0)
    _check(coords, corners, 0.0, True)


def test_empty_and_uncut():
    V, T = _gpu(np.zeros((0, 3), np.int32), np.zeros((0, 8), np.float32), 0.0)
    assert V.shape == (0, 3) and T.shape == (0, 3)
    V, T = _gpu(np.array([[1, 2, 3], [4, 5, 6]], np.int32), np.full((2, 8), 2.0, np.float32), 0.0)
    assert V.shape == (0, 3) and T.shape == (0, 3)


def test_large_sphere_is_closed_and_matches_oracle_counts():
    coords, corners = cells_of(sphere_field(256, radius=100.0))
    V, T = _gpu(coords, corners, 0.0)
    e = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), 1)
    edges, uses = np.unique(e, axis=0, return_counts=True)
    assert (uses == 2).all()
    assert len(V) - len(edges) + len(T) == 2
    OV, OT = O.smc(coords, corners, 0.0)
    np.testing.assert_array_equal(V, OV)
    np.testing.assert_array_equal(T, OT)


def test_two_extractions_in_flight():
    """A second extraction started before the first is written takes its own scratch."""
    import ctypes

    lib = cubvh_amd._lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    cases = [cells_of(sphere_field(36), 0.0), cells_of(sphere_field(30, radius=9.0), 0.5)]
    inputs = [(torch.from_numpy(c).to(DEV), torch.from_numpy(v).to(DEV)) for c, v in cases]
    handles, sizes = [], []
    for (C, V), iso in zip(inputs, (0.0, 0.5)):
        h, nv, nt = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int64()
        cubvh_amd._lib.check(lib.cubvh_sparse_marching_cubes(0, C.data_ptr(), V.data_ptr(), C.shape[0], iso, 0,
                                                             ctypes.byref(h), ctypes.byref(nv), ctypes.byref(nt),
                                                             stream))
        handles.append(h)
        sizes.append((nv.value, nt.value))
    outs = []
    for h, (nv, nt) in zip(reversed(handles), reversed(sizes)):
        verts = torch.empty(nv, 3, dtype=torch.float32, device=DEV)
        tris = torch.empty(nt, 3, dtype=torch.int32, device=DEV)
        cubvh_amd._lib.check(lib.cubvh_smc_write(h, verts.data_ptr(), tris.data_ptr()))
        lib.cubvh_smc_destroy(h)
        outs.append((verts.cpu().numpy(), tris.cpu().numpy()))
    for (c, v), iso, (V, T) in zip(cases, (0.0, 0.5), reversed(outs)):
        OV, OT = O.smc(c, v, iso)
        np.testing.assert_array_equal(V, OV)
        np.testing.assert_array_equal(T, OT)
    # and the shared scratch is usable again afterwards
    _check(*cases[0], 0.0)


def test_argument_checks():
    c = torch.zeros(4, 3, dtype=torch.int32, device=DEV)
    with pytest.raises(RuntimeError, match=r"\[N,8\]"):
        cubvh_amd.sparse_marching_cubes(c, torch.zeros(4, 7, device=DEV), 0.0)
    with pytest.raises(RuntimeError, match="same first-dim"):
        cubvh_amd.sparse_marching_cubes(c, torch.zeros(5, 8, device=DEV), 0.0)
    far = torch.tensor([[0, 0, 0], [1 << 21, 0, 0]], dtype=torch.int32, device=DEV)
    with pytest.raises(cubvh_amd._lib.CubvhError, match="span"):
        cubvh_amd.sparse_marching_cubes(far, torch.ones(2, 8, device=DEV), 0.0)
