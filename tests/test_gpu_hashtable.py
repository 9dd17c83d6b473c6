"""cuHashTable on the GPU against the oracle (the reference's sequential table) and a
brute-force dictionary.

The answer for a key is the lowest row index stored with it, or -1: bit-exact with
the oracle whenever the table has room for every row.  A full table drops rows the
reference's GPU table also drops in an unspecified order; there the answers are
checked to be valid instead.  Scenarios follow the reference's test/cuhashtable.py (the GPU
form of test/hashtable.py).
"""
import numpy as np
import pytest
import torch

import cubvh_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(DEV)


def _search(coords, queries):
    ht = cubvh_amd.cuHashTable(num_dims=coords.shape[1])
    ht.build(_t(coords))
    return ht.search(_t(queries)).cpu().numpy()


def test_reference_known_answers():
    # test/hashtable.py:10-46, 49-76, 125-150
    c3 = np.array([[0, 0, 0], [1, 2, 3], [2, 2, 2], [3, 4, 5], [10, 0, -3]])
    assert _search(c3, [[0, 0, 0], [1, 2, 3], [10, 0, -3], [9, 9, 9]]).tolist() == [0, 1, 4, -1]
    c4 = np.array([[0, 0, 0, 0], [1, 2, 3, 4], [2, 2, 2, 2]])
    assert _search(c4, [[1, 2, 3, 4], [9, 9, 9, 9]]).tolist() == [1, -1]
    ce = np.array([[0, 0, 0], [2147483647, -2147483648, 0], [-2147483648, 2147483647, -2147483648],
                   [123, 456, 789], [123, 456, 789]])
    qe = np.array([[2147483647, -2147483648, 0], [-2147483648, 2147483647, -2147483648], [123, 456, 789], [7, 7, 7]])
    assert _search(ce, qe).tolist() == [1, 2, 3, -1]


def test_reference_large_random_equals_oracle():
    from test_oracle_hashtable import reference_large_random_inputs

    coords, queries, idx = reference_large_random_inputs()
    out = _search(coords, queries)
    np.testing.assert_array_equal(out, O.hash_lookup(coords, queries))
    assert (out[: len(idx)] >= 0).all()


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 7, 8, 15, 16, 31, 63])
def test_key_widths_with_duplicates_equal_oracle(d):
    rng = np.random.default_rng(100 + d)
    n = 5000
    coords = rng.integers(-5, 5, size=(n, d), dtype=np.int32)
    if d > 3:
        coords[:, 3:] = rng.integers(-1, 1, size=(n, d - 3))  # keep duplicates frequent
    queries = np.concatenate([coords[rng.permutation(n)[: n // 2]], rng.integers(-6, 6, size=(n // 2, d))])
    np.testing.assert_array_equal(_search(coords, queries), O.hash_lookup(coords, queries))


def test_empty_and_tiny_tables():
    ht = cubvh_amd.cuHashTable()
    ht.build(torch.zeros(0, 3, dtype=torch.int32, device=DEV))
    assert ht.impl.capacity == 16
    assert ht.search(_t([[0, 0, 0], [1, 2, 3]])).tolist() == [-1, -1]
    assert ht.search(torch.zeros(0, 3, dtype=torch.int32, device=DEV)).numel() == 0
    ht.build(_t([[4, 5, 6]]))
    assert ht.search(_t([[4, 5, 6], [6, 5, 4]])).tolist() == [0, -1]


def test_resize_prepare_insert_and_a_second_batch():
    rng = np.random.default_rng(7)
    a = rng.integers(-30, 30, size=(3000, 3), dtype=np.int32)
    b = rng.integers(-30, 30, size=(2000, 3), dtype=np.int32)
    q = rng.integers(-32, 32, size=(8000, 3), dtype=np.int32)
    impl = cubvh_amd.create_cuHashTable()
    impl.resize(12000)
    impl.prepare()
    impl.insert(_t(a))
    np.testing.assert_array_equal(impl.search(_t(q)).cpu().numpy(), O.hash_lookup(a, q, capacity=12000))
    # a second batch adds its rows under their own indices: the answer is the lowest
    # index stored with the key over both batches
    impl.insert(_t(b))
    first = {}
    for rows in (a, b):
        for i, r in enumerate(rows):
            k = tuple(r.tolist())
            first[k] = min(first.get(k, i), i)
    expect = np.array([first.get(tuple(r.tolist()), -1) for r in q], np.int32)
    np.testing.assert_array_equal(impl.search(_t(q)).cpu().numpy(), expect)
    # prepare() empties the table
    impl.prepare()
    assert (impl.search(_t(q)).cpu().numpy() == -1).all()


@pytest.mark.parametrize("capacity", [16, 1000, 1999])
def test_full_table_answers_stay_valid(capacity):
    rng = np.random.default_rng(capacity)
    coords = rng.integers(-40, 40, size=(2000, 3), dtype=np.int32)
    absent = rng.integers(100, 200, size=(500, 3), dtype=np.int32)
    impl = cubvh_amd.create_cuHashTable()
    impl.resize(capacity)
    impl.insert(_t(coords))
    out = impl.search(_t(np.concatenate([coords, absent]))).cpu().numpy()
    hit = out[:2000]
    found = hit >= 0
    np.testing.assert_array_equal(coords[hit[found]], coords[found])  # a hit holds the key
    stored = np.unique(hit[found])
    assert 0 < len(stored) <= capacity
    assert (out[2000:] == -1).all()  # absent keys end after one full cycle


def test_table_exactly_full():
    """capacity == rows: every row is stored (no empty slot is left), present keys are
    found and absent ones end after one full cycle."""
    rng = np.random.default_rng(21)
    coords = np.unique(rng.integers(-50, 50, size=(3000, 3)), axis=0).astype(np.int32)
    coords = coords[rng.permutation(len(coords))]
    impl = cubvh_amd.create_cuHashTable()
    impl.resize(len(coords))
    impl.insert(_t(coords))
    np.testing.assert_array_equal(impl.search(_t(coords)).cpu().numpy(), np.arange(len(coords), dtype=np.int32))
    absent = rng.integers(60, 90, size=(500, 3), dtype=np.int32)
    assert (impl.search(_t(absent)).cpu().numpy() == -1).all()


def test_sparse_voxel_shell_at_scale():
    """2^22-row sparse voxel set, 2^23 queries: every answer equals the brute force."""
    res = 1024
    rng = np.random.default_rng(0)
    g = rng.normal(size=(1 << 22, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    coords = np.floor((g * 0.45 + 0.5) * res).astype(np.int32)
    n = len(coords)
    queries = np.concatenate([coords[rng.integers(0, n, n)], rng.integers(0, res, size=(n, 3))]).astype(np.int32)
    key = lambda c: (c[:, 0].astype(np.int64) << 22) | (c[:, 1].astype(np.int64) << 11) | c[:, 2]
    kc, kq = key(coords), key(queries)
    uniq, first = np.unique(kc, return_index=True)
    pos = np.clip(np.searchsorted(uniq, kq), 0, len(uniq) - 1)
    expect = np.where(uniq[pos] == kq, first[pos], -1).astype(np.int32)
    np.testing.assert_array_equal(_search(coords, queries), expect)


def test_errors():
    ht = cubvh_amd.cuHashTable()
    with pytest.raises(cubvh_amd._lib.CubvhError, match="not built"):
        ht.impl.search(_t([[1, 2, 3]]))
    ht.build(_t([[1, 2, 3]]))
    with pytest.raises(RuntimeError, match="int32"):
        ht.impl.search(torch.zeros(2, 3, dtype=torch.int64, device=DEV))
    with pytest.raises(AssertionError):
        ht.search(_t([[1, 2, 3, 4]]))
    with pytest.raises(cubvh_amd._lib.CubvhError, match="words"):
        ht.impl.search(_t([[1, 2, 3, 4]]))
    with pytest.raises(cubvh_amd._lib.CubvhError, match="prepare"):
        ht.impl.insert(_t([[1, 2, 3, 4]]))
