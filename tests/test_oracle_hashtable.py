"""Pins the oracle's restatement of the reference hash table (CPU only).

Three independent checks:
  * the known answers of the reference's own test (test/hashtable.py:10-156);
  * a brute-force dictionary: the answer for a key is its lowest row index among the
    rows the sequential table holds (rows 0..capacity-1 when it fills up), else -1;
  * slot for slot against the reference's CPU table compiled from its own header
    (oracle/_ref/libref_hashtable.so, built when the reference sources are present).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def brute(coords, queries, rows=None):
    coords = np.asarray(coords)
    n = len(coords) if rows is None else min(rows, len(coords))
    first = {}
    for i in range(n):
        first.setdefault(tuple(coords[i].tolist()), i)
    return np.array([first.get(tuple(q.tolist()), -1) for q in np.asarray(queries)], np.int32)


# test/hashtable.py:10-46
def test_known_answers_3d():
    coords = np.array([[0, 0, 0], [1, 2, 3], [2, 2, 2], [3, 4, 5], [10, 0, -3]], np.int32)
    queries = np.array([[0, 0, 0], [1, 2, 3], [10, 0, -3], [9, 9, 9]], np.int32)
    assert O.hash_lookup(coords, queries).tolist() == [0, 1, 4, -1]


# test/hashtable.py:49-76
def test_known_answers_4d():
    coords = np.array([[0, 0, 0, 0], [1, 2, 3, 4], [2, 2, 2, 2]], np.int32)
    queries = np.array([[1, 2, 3, 4], [9, 9, 9, 9]], np.int32)
    assert O.hash_lookup(coords, queries).tolist() == [1, -1]


# test/hashtable.py:125-150: int32 extremes and a duplicate key (rows 3 and 4)
def test_known_answers_edge_values():
    coords = np.array([[0, 0, 0], [2147483647, -2147483648, 0], [-2147483648, 2147483647, -2147483648],
                       [123, 456, 789], [123, 456, 789]], np.int32)
    queries = np.array([[2147483647, -2147483648, 0], [-2147483648, 2147483647, -2147483648], [123, 456, 789],
                        [7, 7, 7]], np.int32)
    out = O.hash_lookup(coords, queries).tolist()
    assert out[0] == 1 and out[1] == 2 and out[3] == -1
    assert out[2] in (3, 4)
    assert out[2] == 3  # the sequential table answers with the lower row


def reference_large_random_inputs(num_dims=3, num_keys=200000, num_queries=100000):
    """The inputs of test/hashtable.py:79-122 (torch's CPU generator, seed 42)."""
    torch.manual_seed(42)
    low, high = -5000, 5000
    coords = torch.randint(low, high, (num_keys, num_dims), dtype=torch.int32)
    idx = torch.randperm(num_keys)[: num_queries // 2]
    pos = coords[idx]
    neg = torch.randint(low, high, (num_queries - pos.shape[0], num_dims), dtype=torch.int32)
    return coords.numpy(), torch.cat([pos, neg]).numpy(), idx.numpy()


# test/hashtable.py:79-122, checked exactly (every answer, not only the positives)
def test_reference_large_random_against_brute_force():
    coords, queries, idx = reference_large_random_inputs()
    out = O.hash_lookup(coords, queries)
    np.testing.assert_array_equal(out, brute(coords, queries))
    # the reference test's own criterion: positives map to a row holding the key
    pos = out[: len(idx)]
    assert (pos >= 0).all()
    np.testing.assert_array_equal(coords[pos], queries[: len(idx)])


@pytest.mark.parametrize("d", [1, 2, 3, 4, 7])
def test_brute_force_with_duplicates(d):
    rng = np.random.default_rng(d)
    coords = rng.integers(-6, 6, size=(3000, d), dtype=np.int32)  # many duplicate keys
    queries = np.concatenate([coords[rng.permutation(3000)[:1500]], rng.integers(-8, 8, size=(1500, d))]).astype(
        np.int32)
    np.testing.assert_array_equal(O.hash_lookup(coords, queries), brute(coords, queries))


@pytest.mark.parametrize("capacity", [1, 16, 999, 1000, 1001])
def test_full_table_keeps_the_first_rows(capacity):
    rng = np.random.default_rng(capacity)
    coords = rng.integers(-20, 20, size=(1000, 3), dtype=np.int32)
    queries = np.concatenate([coords, rng.integers(-25, 25, size=(500, 3))]).astype(np.int32)
    table, dropped = O.hash_build(coords, capacity)
    assert dropped == max(0, 1000 - capacity)
    np.testing.assert_array_equal(O.hash_search(table, coords, queries), brute(coords, queries, rows=capacity))


def test_a_planted_mistake_is_caught():
    """The brute force separates "first row" from "last row" on duplicate keys."""
    coords = np.array([[1, 1, 1], [1, 1, 1], [2, 2, 2]], np.int32)
    q = coords[:1]
    assert brute(coords, q).tolist() == [0]
    last = {tuple(c.tolist()): i for i, c in enumerate(coords)}
    assert last[(1, 1, 1)] != brute(coords, q)[0]


@pytest.mark.skipif(not O.ref_hash_available(), reason="reference CPU hash table not built (no reference sources)")
@pytest.mark.parametrize("d,n,capacity", [(3, 5000, None), (1, 3000, None), (4, 2000, 1500), (2, 4000, 4000),
                                          (5, 777, 64), (3, 200000, None)])
def test_slot_for_slot_against_reference_cpu_table(d, n, capacity):
    rng = np.random.default_rng(n + d)
    span = 40 if n < 10000 else 5000
    coords = rng.integers(-span, span, size=(n, d), dtype=np.int32)
    queries = np.concatenate([coords[rng.permutation(n)[: n // 2]], rng.integers(-span, span, size=(n // 2, d))])
    queries = queries.astype(np.int32)
    ref_out, ref_table = O.ref_hash_build_search(coords, queries, capacity)
    table, _ = O.hash_build(coords, capacity)
    np.testing.assert_array_equal(table, ref_table)
    np.testing.assert_array_equal(O.hash_search(table, coords, queries), ref_out)
