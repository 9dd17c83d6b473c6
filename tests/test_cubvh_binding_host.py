"""The `_cubvh` extension module without a GPU: the reference native module's names
(src/bindings.cpp:17-87), the Eigen caster's force-cast of the mesh arrays
(patch/eigen/matrix.h:490-524), the reference state_dict layout and errors
(src/api_gpu.cu:47-68, 110-121).  Construction and state_dict are host-only."""
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT

# names bound by the reference module (src/bindings.cpp:17-87)
REFERENCE_NAMES = ["cuBVH", "create_cuBVH", "from_state_dict", "floodfill", "sparse_marching_cubes", "cuHashTable",
                   "create_cuHashTable", "fill_holes", "merge_vertices", "sparse_marching_cubes_cpu", "decimate",
                   "parallel_decimate", "HashTable"]


@pytest.fixture(scope="module")
def _cubvh():
    # imported as the reference's Python layer imports it (cubvh/api.py:5)
    d = str(ROOT / "cubvh_amd")
    if d not in sys.path:
        sys.path.insert(0, d)
    import _cubvh

    return _cubvh


def test_module_exports_the_reference_names(_cubvh):
    for name in REFERENCE_NAMES:
        assert hasattr(_cubvh, name), name
    for m in ("ray_trace", "unsigned_distance", "signed_distance", "state_dict"):
        assert hasattr(_cubvh.cuBVH, m)
    for m in ("set_num_dims", "get_num_dims", "resize", "prepare", "insert", "build", "search"):
        assert hasattr(_cubvh.cuHashTable, m)


def test_create_force_casts_and_state_dict_is_the_reference_layout(_cubvh, oracle, ico4):
    """float64 vertices and int64 faces (trimesh's defaults) build the same tree as
    float32 / int32 ones; state_dict() is {"nodes": int32 [n, 9], "triangles": int32
    [M, 12]}, the same as cubvh_amd.cuBVH(...).state_dict(layout="reference"), and the
    oracle's restatement of the reference walk answers from it like brute force."""
    import cubvh_amd
    from cubvh_amd import synthetic as S

    v, f = ico4
    a = _cubvh.create_cuBVH(v.astype(np.float64), f.astype(np.int64)).state_dict()
    b = _cubvh.create_cuBVH(v, f).state_dict()
    c = cubvh_amd.cuBVH(v, f).state_dict(layout="reference")
    assert set(a) == {"nodes", "triangles"}
    assert a["nodes"].dtype == torch.int32 and a["nodes"].shape[1] == 9
    assert a["triangles"].dtype == torch.int32 and tuple(a["triangles"].shape) == (len(f), 12)
    for k in a:
        assert torch.equal(a[k], b[k]) and torch.equal(a[k], c[k])
    # lists and torch CPU tensors go through the same cast
    d = _cubvh.create_cuBVH(torch.from_numpy(v.astype(np.float64)), f.tolist()).state_dict()
    assert torch.equal(d["nodes"], a["nodes"])
    R = oracle.RefBvh.from_arrays(a["nodes"].numpy(), a["triangles"].numpy())
    p = S.uniform_points(3000, seed=21)
    od, of, _ = oracle.udf_brute(v, f, p, return_uvw=False)
    rd, rf, _, _ = R.udf(p, return_uvw=False, tie_mode=1)
    np.testing.assert_array_equal(rd, od)
    np.testing.assert_array_equal(rf, of)


def test_from_state_dict_roundtrip_and_errors(_cubvh, oracle, ico4):
    v, f = ico4
    sd = _cubvh.create_cuBVH(v, f).state_dict()
    again = _cubvh.from_state_dict(sd).state_dict()
    for k in sd:
        assert torch.equal(sd[k], again[k])
    # the reference's own state dict (its build) loads too
    nodes9, tris12 = oracle.RefBvh(v, f).export()
    ref = _cubvh.from_state_dict({"nodes": torch.from_numpy(nodes9), "triangles": torch.from_numpy(tris12)})
    assert ref.state_dict()["triangles"].shape == (len(f), 12)
    with pytest.raises(RuntimeError, match="missing 'triangles'"):
        _cubvh.from_state_dict({"nodes": sd["nodes"]})
    with pytest.raises(RuntimeError, match="missing 'nodes'"):
        _cubvh.from_state_dict({"triangles": sd["triangles"]})
    with pytest.raises(RuntimeError, match="\\[M, 12\\]"):
        _cubvh.from_state_dict({"nodes": sd["nodes"], "triangles": sd["triangles"][:, :6]})


def test_bad_inputs_and_out_of_scope_names(_cubvh, ico4):
    v, f = ico4
    with pytest.raises(TypeError):
        _cubvh.create_cuBVH(v[:, :2], f)
    with pytest.raises(TypeError):
        _cubvh.create_cuBVH(v, f[:, :2])
    with pytest.raises(RuntimeError, match="references vertex"):
        _cubvh.create_cuBVH(v, np.where(f == 7, len(v), f))
    with pytest.raises(RuntimeError, match="references vertex"):
        _cubvh.create_cuBVH(v, np.where(f == 7, -1, f))  # -1 casts to 2^32 - 1, as the caster does
    for name in ("fill_holes", "merge_vertices", "sparse_marching_cubes_cpu", "decimate", "parallel_decimate"):
        with pytest.raises(NotImplementedError):
            getattr(_cubvh, name)(v, f)
    with pytest.raises(NotImplementedError):
        _cubvh.HashTable()
    ht = _cubvh.create_cuHashTable()
    assert ht.get_num_dims() == 3
    ht.set_num_dims(5)
    assert ht.get_num_dims() == 5
    with pytest.raises(RuntimeError, match="bool"):
        _cubvh.floodfill(torch.zeros(1, 2, 2, 2, dtype=torch.int32))
