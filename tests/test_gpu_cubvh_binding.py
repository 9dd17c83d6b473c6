"""The `_cubvh` extension module on the GPU, driven with the call pattern of the
reference's Python layer (cubvh/api.py:24-185: caller-allocated outputs, [N, 3]
contiguous fp32 queries, uvw optional, mode 0 / 1): every result equals the
cubvh_amd Python API's bit for bit (both run the same HIP kernels), and the oracle's."""
import sys

import numpy as np
import pytest
import torch

import cubvh_amd
from cubvh_amd import synthetic as S
from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def _cubvh():
    d = str(ROOT / "cubvh_amd")
    if d not in sys.path:
        sys.path.insert(0, d)
    import _cubvh

    return _cubvh


def _out(n, dev, uvw=True):
    return (torch.empty(n, dtype=torch.float32, device=dev), torch.empty(n, dtype=torch.int64, device=dev),
            torch.empty(n, 3, dtype=torch.float32, device=dev) if uvw else None)


def test_queries_equal_the_python_api_and_oracle(gpu, _cubvh, oracle):
    v, f, _ = S.perturbed_sphere(5, seed=7)
    # trimesh-style float64 / int64 arrays: the binding casts them (api.py:24)
    impl = _cubvh.create_cuBVH(v.astype(np.float64), f.astype(np.int64))
    mine = cubvh_amd.cuBVH(v, f)
    p = S.uniform_points(50000, seed=3)
    P = torch.from_numpy(p).to(gpu)
    d, fi, uvw = _out(len(p), gpu)
    impl.unsigned_distance(P, d, fi, uvw)  # api.py:88
    for x, y in zip((d, fi, uvw), mine.unsigned_distance(P, return_uvw=True)):
        assert torch.equal(x, y)
    od, of, ou = oracle.udf_brute(v, f, p[:3000])
    np.testing.assert_array_equal(d[:3000].cpu().numpy(), od)
    np.testing.assert_array_equal(fi[:3000].cpu().numpy(), of)
    d2, f2, _ = _out(len(p), gpu, uvw=False)
    impl.unsigned_distance(P, d2, f2, None)  # return_uvw=False passes None
    assert torch.equal(d2, d) and torch.equal(f2, fi)
    for mode, name in ((0, "watertight"), (1, "raystab")):  # api.py:7-10, 119
        s, sf, su = _out(len(p), gpu)
        impl.signed_distance(P, s, sf, su, mode)
        for x, y in zip((s, sf, su), mine.signed_distance(P, return_uvw=True, mode=name)):
            assert torch.equal(x, y)
    ro, rd = S.sphere_rays(40000, seed=5)
    RO, RD = torch.from_numpy(ro).to(gpu), torch.from_numpy(rd).to(gpu)
    pos = torch.empty(len(ro), 3, dtype=torch.float32, device=gpu)
    rf = torch.empty(len(ro), dtype=torch.int64, device=gpu)
    depth = torch.empty(len(ro), dtype=torch.float32, device=gpu)
    impl.ray_trace(RO, RD, pos, rf, depth)  # api.py:59
    for x, y in zip((pos, rf, depth), mine.ray_trace(RO, RD)):
        assert torch.equal(x, y)
    # state_dict -> from_state_dict (api.py:26-36) answers the same
    again = _cubvh.from_state_dict(impl.state_dict())
    d3, f3, u3 = _out(len(p), gpu)
    again.unsigned_distance(P, d3, f3, u3)
    assert torch.equal(d3, d) and torch.equal(f3, fi) and torch.equal(u3, uvw)


def test_queries_run_on_the_current_stream(gpu, _cubvh, ico4):
    v, f = ico4
    impl = _cubvh.create_cuBVH(v, f)
    P = torch.from_numpy(S.uniform_points(20000, seed=4)).to(gpu)
    ref = cubvh_amd.cuBVH(v, f).unsigned_distance(P)
    s = torch.cuda.Stream(device=gpu)
    s.wait_stream(torch.cuda.current_stream(gpu))
    with torch.cuda.stream(s):
        d, fi, _ = _out(P.shape[0], gpu, uvw=False)
        impl.unsigned_distance(P, d, fi, None)
    torch.cuda.current_stream(gpu).wait_stream(s)
    assert torch.equal(d, ref[0]) and torch.equal(fi, ref[1])


def test_query_tensors_are_checked(gpu, _cubvh, ico4):
    v, f = ico4
    impl = _cubvh.create_cuBVH(v, f)
    P = torch.zeros(10, 3, device=gpu)
    d, fi, _ = _out(10, gpu, uvw=False)
    with pytest.raises(RuntimeError, match="int64|Long"):
        impl.unsigned_distance(P, d, fi.int(), None)
    with pytest.raises(RuntimeError, match="GPU"):
        impl.unsigned_distance(P.cpu(), d, fi, None)
    with pytest.raises(RuntimeError, match="elements"):
        impl.unsigned_distance(P, d[:5], fi, None)
    with pytest.raises(RuntimeError, match="\\[N, 3\\]"):
        impl.unsigned_distance(P.view(-1), d, fi, None)


def test_floodfill_hashtable_marching_cubes(gpu, _cubvh, oracle):
    g = np.random.default_rng(2).random((2, 20, 20, 20)) < 0.4
    lab = _cubvh.floodfill(torch.from_numpy(g).to(gpu))  # api.py:136
    assert lab.dtype == torch.int32 and lab.shape == (2, 20, 20, 20)
    np.testing.assert_array_equal(lab.cpu().numpy(), oracle.floodfill_bfs(g))
    rng = np.random.default_rng(1)
    keys = torch.from_numpy(rng.integers(-30, 30, size=(5000, 3), dtype=np.int32)).to(gpu)
    q = torch.from_numpy(rng.integers(-35, 35, size=(8000, 3), dtype=np.int32)).to(gpu)
    ht = _cubvh.create_cuHashTable()  # api.py:153-169
    ht.set_num_dims(3)
    ht.build(keys)
    got = ht.search(q)
    ref = cubvh_amd.cuHashTable()
    ref.build(keys)
    assert torch.equal(got, ref.search(q))
    np.testing.assert_array_equal(got.cpu().numpy(), oracle.hash_lookup(keys.cpu().numpy(), q.cpu().numpy()))
    x = np.stack(np.meshgrid(*[np.arange(13)] * 3, indexing="ij"), -1).astype(np.float32)
    field = np.linalg.norm(x - 6.2, axis=-1) - 4.1
    cells = np.stack(np.meshgrid(*[np.arange(12)] * 3, indexing="ij"), -1).reshape(-1, 3)
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    cv = np.stack([field[cells[:, 0] + a, cells[:, 1] + b, cells[:, 2] + c] for a, b, c in offs], 1)
    C = torch.from_numpy(cells.astype(np.int32)).to(gpu)
    V = torch.from_numpy(cv).to(gpu)
    mv, mt = _cubvh.sparse_marching_cubes(C, V, 0.0, False)  # api.py:185
    ov, ot = oracle.smc(cells.astype(np.int32), cv, 0.0)
    np.testing.assert_array_equal(mv.cpu().numpy(), ov)
    np.testing.assert_array_equal(mt.cpu().numpy(), ot)
