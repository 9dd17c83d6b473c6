"""Pins the oracle's sparse marching cubes and the case table (CPU only).

* the case table (cubvh_amd/csrc/mc_table.hpp) entry by entry against the reference's
  own tables, read from its CPU extraction compiled from its header (oracle/_ref);
* the oracle's extraction against that reference CPU extraction: the same triangles
  with the same vertex positions (the reference's CPU path numbers vertices in
  first-use order, its GPU path and the oracle in edge-key order);
* invariants that hold whatever the implementation: a sampled closed surface gives a
  closed 2-manifold of Euler characteristic 2, every vertex lies on a lattice edge at
  the interpolated position, and a one-corner cell gives one triangle.
"""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
CORNER = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
needs_ref = pytest.mark.skipif(not O.ref_smc_available(), reason="reference CPU extraction not built")


def packed_table():
    src = (ROOT / "cubvh_amd" / "csrc" / "mc_table.hpp").read_text()
    vals = [int(x, 16) for x in re.findall(r"0x([0-9a-f]{16})ull", src)]
    assert len(vals) == 256
    return vals


def unpack(v):
    nt = v >> 60
    return [(v >> (4 * j)) & 15 for j in range(3 * nt)]


def cells_of(field, iso=0.0, all_cells=False):
    """Cells of a dense lattice field whose corners straddle iso (or all cells)."""
    r = np.array(field.shape) - 1
    cells = np.stack(np.meshgrid(*[np.arange(k) for k in r], indexing="ij"), -1).reshape(-1, 3)
    cv = np.stack([field[cells[:, 0] + o[0], cells[:, 1] + o[1], cells[:, 2] + o[2]] for o in CORNER], 1)
    keep = np.ones(len(cells), bool) if all_cells else (cv.min(1) < iso) & (cv.max(1) >= iso)
    return cells[keep].astype(np.int32), cv[keep].astype(np.float32)


def sphere_field(res, centre=None, radius=None):
    g = np.stack(np.meshgrid(*[np.arange(res + 1)] * 3, indexing="ij"), -1).astype(np.float64)
    c = np.full(3, res / 2 + 0.3) if centre is None else np.asarray(centre)
    rad = res * 0.35 if radius is None else radius
    return (np.linalg.norm(g - c, axis=-1) - rad).astype(np.float32)


def test_case_table_invariants():
    for c, v in enumerate(packed_table()):
        edges = set(unpack(v))
        crossing = {e for e, (a, b) in enumerate(EDGE) if ((c >> a) ^ (c >> b)) & 1}
        assert edges == crossing, c
    assert packed_table()[0] == 0 and packed_table()[255] == 0


@needs_ref
def test_case_table_equals_reference():
    edge_masks, tri = O.ref_mc_tables()
    for c, v in enumerate(packed_table()):
        row = unpack(v)
        assert tri[c, : len(row)].tolist() == row, c
        assert (tri[c, len(row):] == -1).all(), c
        crossing = sum(1 << e for e, (a, b) in enumerate(EDGE) if ((c >> a) ^ (c >> b)) & 1)
        assert edge_masks[c] == crossing, c


def _same_mesh_as_reference_cpu(coords, corners, iso, consistent):
    V, T = O.smc(coords, corners, iso, consistent)
    RV, RT = O.ref_smc_cpu(coords, corners, iso, consistent)
    assert V.shape == RV.shape and T.shape == RT.shape
    np.testing.assert_array_equal(V[T], RV[RT])
    np.testing.assert_array_equal(np.unique(V, axis=0), np.unique(RV, axis=0))
    return V, T


@needs_ref
@pytest.mark.parametrize("iso", [0.0, 1.3, -2.0])
def test_oracle_equals_reference_cpu_on_a_sphere(iso):
    coords, corners = cells_of(sphere_field(32), iso)
    _same_mesh_as_reference_cpu(coords, corners, iso, False)


@needs_ref
@pytest.mark.parametrize("consistent", [False, True])
def test_oracle_equals_reference_cpu_on_random_cells(consistent):
    """Independent random corner values per cell: every case, inconsistent neighbours,
    negative coordinates, a duplicated cell; averaging makes shared corners agree."""
    rng = np.random.default_rng(5)
    coords = rng.integers(-6, 6, size=(3000, 3)).astype(np.int32)
    coords[-1] = coords[0]
    corners = rng.normal(size=(3000, 8)).astype(np.float32)
    _same_mesh_as_reference_cpu(coords, corners, 0.1, consistent)


@needs_ref
def test_oracle_equals_reference_cpu_with_noisy_shared_corners():
    coords, corners = cells_of(sphere_field(28))
    corners = corners + np.random.default_rng(1).normal(scale=0.05, size=corners.shape).astype(np.float32)
    _same_mesh_as_reference_cpu(coords, corners, 0.0, True)


def test_sphere_is_a_closed_manifold_of_genus_zero():
    coords, corners = cells_of(sphere_field(40))
    V, T = O.smc(coords, corners, 0.0)
    e = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), 1)
    edges, uses = np.unique(e, axis=0, return_counts=True)
    assert (uses == 2).all()
    assert len(V) - len(edges) + len(T) == 2
    # consistent orientation: each undirected edge appears once in each direction
    d = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    assert len(np.unique(d, axis=0)) == len(d)
    # outward winding (corner value < iso is inside): positive signed volume
    a, b, c = V[T[:, 0]].astype(np.float64), V[T[:, 1]].astype(np.float64), V[T[:, 2]].astype(np.float64)
    assert np.einsum("ij,ij->i", a, np.cross(b, c)).sum() > 0


def test_vertices_on_lattice_edges_at_the_interpolated_point():
    coords, corners = cells_of(sphere_field(20))
    V, _ = O.smc(coords, corners, 0.0)
    frac = (V != np.floor(V)).sum(1)
    assert (frac <= 1).all()  # two coordinates are lattice values


def test_one_corner_cell_known_answer():
    coords = np.array([[3, -2, 7]], np.int32)
    corners = np.array([[-1.0, 3.0, 2.0, 1.0, 4.0, 2.0, 2.0, 2.0]], np.float32)  # corner 0 inside
    V, T = O.smc(coords, corners, 0.0)
    assert T.shape == (1, 3) and V.shape == (3, 3)
    # edges 0 (x), 3 (y), 8 (z) from corner 0 at t = 1/4, 1/2, 1/5; keys ascending:
    # (3,-2,7,x) < (3,-2,7,y) < (3,-2,7,z)
    np.testing.assert_array_equal(V, np.array([[3.25, -2, 7], [3, -1.5, 7], [3, -2, 7.2]], np.float32))
    # table case 1 lists edges (0, 8, 3); emitted as (e0, e2, e1)
    assert T.tolist() == [[0, 1, 2]]


def test_empty_and_uncut_inputs():
    V, T = O.smc(np.zeros((0, 3), np.int32), np.zeros((0, 8), np.float32), 0.0)
    assert V.shape == (0, 3) and T.shape == (0, 3)
    V, T = O.smc(np.zeros((2, 3), np.int32), np.full((2, 8), -1.0, np.float32), 0.0)
    assert V.shape == (0, 3) and T.shape == (0, 3)
