"""GPU test of cubvh_unsigned_distance_noalloc, the allocation-free form of
cubvh_unsigned_distance (static per-device scratch for the Morton order): through the
C ABI, its outputs equal the allocating call's bit for bit on two calls in a row (the
second with fewer queries, so it runs over what the first left in the scratch), and
on a batch too small to be ordered; sizes above its limit and a tree that is not on
the device are refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

import cubvh_amd
from cubvh_amd import _lib
from cubvh_amd import synthetic as S

pytestmark = pytest.mark.gpu

MAX_QUERIES = 1 << 20


def _outputs(n, dev):
    return (torch.empty(n, dtype=torch.float32, device=dev), torch.empty(n, dtype=torch.int64, device=dev),
            torch.empty(n, 3, dtype=torch.float32, device=dev))


def _call(fn, bvh, P, out):
    d, f, u = out
    s = ctypes.c_void_p(torch.cuda.current_stream(P.device).cuda_stream)
    return fn(bvh._h, P.device.index, P.shape[0], P.data_ptr(), d.data_ptr(), f.data_ptr(), u.data_ptr(), s)


def test_noalloc_equals_allocating_call(gpu):
    lib = _lib.load()
    v, f, _ = S.perturbed_sphere(6, seed=3)
    bvh = cubvh_amd.cuBVH(v, f)
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    _lib.check(lib.cubvh_bvh_upload(bvh._h, gpu.index, s))
    sizes = (MAX_QUERIES, 300_001, 1000)  # full, smaller, below the Morton-order threshold
    pts = [torch.from_numpy(S.uniform_points(n, seed=10 + k)).to(gpu) for k, n in enumerate(sizes)]
    ref = []
    for P in pts:
        out = _outputs(P.shape[0], gpu)
        _lib.check(_call(lib.cubvh_unsigned_distance, bvh, P, out))
        ref.append(out)
    got = []
    for P in pts:  # back to back: each call reuses the scratch the previous one wrote
        out = _outputs(P.shape[0], gpu)
        _lib.check(_call(lib.cubvh_unsigned_distance_noalloc, bvh, P, out))
        got.append(out)
    torch.cuda.synchronize()
    for n, r, g in zip(sizes, ref, got):
        for a, b in zip(r, g):
            assert torch.equal(a, b), n
    # the allocating call agrees with the oracle on a sample (so both are the right answer)
    from oracle import oracle as O

    idx = np.arange(0, sizes[1], 97)
    od, of, ou = O.udf_brute(v, f, pts[1].cpu().numpy()[idx])
    np.testing.assert_array_equal(got[1][0].cpu().numpy()[idx], od)
    np.testing.assert_array_equal(got[1][1].cpu().numpy()[idx], of)


def test_noalloc_refuses_before_device_work(gpu):
    lib = _lib.load()
    v, f = S.icosphere(3, 0.5)
    bvh = cubvh_amd.cuBVH(v, f)
    P = torch.zeros(16, 3, device=gpu)
    out = _outputs(16, gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    # the tree is not on the device yet
    rc = _call(lib.cubvh_unsigned_distance_noalloc, bvh, P, out)
    assert rc != 0 and b"cubvh_bvh_upload" in lib.cubvh_last_error()
    _lib.check(lib.cubvh_bvh_upload(bvh._h, gpu.index, s))
    # more queries than the scratch holds: refused on the host (no buffer is touched)
    d, fi, u = out
    rc = lib.cubvh_unsigned_distance_noalloc(bvh._h, gpu.index, MAX_QUERIES + 1, P.data_ptr(), d.data_ptr(),
                                             fi.data_ptr(), u.data_ptr(), s)
    assert rc != 0 and b"2^20" in lib.cubvh_last_error()
    _lib.check(_call(lib.cubvh_unsigned_distance_noalloc, bvh, P, out))
    ref = _outputs(16, gpu)
    _lib.check(_call(lib.cubvh_unsigned_distance, bvh, P, ref))
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(out, ref))
