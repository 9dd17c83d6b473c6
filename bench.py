"""Benchmark of the BVH query path on one MI355X.

Headline (the metric's configuration): dense unsigned distance over a 256^3 grid on
[-1, 1]^3 against the perturbed level-8 icosphere (1,310,720 faces).  A step is one
query pass over all 16,777,216 grid points, inputs resident in HBM: Morton-order the
points (cubvh_morton_order), then walk them (cubvh_distance_ordered) -- the two halves
of cubvh_unsigned_distance, whose own rate is reported as config.dropin_call_Mq_s.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu-baseline] [--headline-only]
                 [--hashtable-only] [--smc-only] [--valu-probe]

Prints one JSON line.  `value` = M queries/s over the timed steps (wall clock
bracketed by device synchronizations).  `roofline` is the walk kernel's: its HBM bytes
per launch from the committed counter summary (profiles/*_counters.md) over its
average duration, measured with HIP events on the launch stream around the walk
inside the timed steps, against 8 TB/s; and its VALU instructions per second against
the issue ceiling measured in the same run (cubvh_valu_probe), the bound the walks
sit at.  `reference_walk_equiv_GBs` prices the reference's own walk on the same
workload (tests/golden/algo_bytes.json) over the same time; `copy_GBs` is a measured
1 GiB device-to-device copy rate.  `cpu_baseline` times the CPU restatement of the
reference's BVH walk (oracle/) on a seeded sample of the same grid on host cores.
The JSON also carries the metric's two other forms: `fwd_bwd` (jittered level-6
sphere, 2^20 points, loss = mean(udf^2), forward + hand-written backward) and
`batched` (512 perturbed icospheres of 1,280..327,680 faces with 4,096..262,144
uniform points each: one autotuned, HIP-graph-replayed forest launch against 512
per-mesh calls), `hashtable` (sparse-voxel hash table build and search) and
`sparse_marching_cubes` (the cut cells of a 1024^3 lattice), each with the reference's
CPU implementation timed beside it, and `extra`: watertight SDF and floodfill on the
headline grid, camera rays and raystab on the soup, and the GPU tree builds (skipped
with --no-extras, or with --headline-only unless --extras is given).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def _algo_bytes():
    p = ROOT / "tests" / "golden" / "algo_bytes.json"
    if not p.exists():
        return None
    return json.loads(p.read_text())


def profile_figures(name):
    """Figures of a committed counter summary (profiles/<name>_counters.md): every line
    of the form `key: number`.  Empty when the file is absent."""
    out = {}
    p = ROOT / "profiles" / f"{name}_counters.md"
    try:
        text = p.read_text()
    except OSError:
        return out
    for line in text.splitlines():
        key, sep, val = line.partition(":")
        key = key.strip().lstrip("*- ").strip()
        if not sep or not key.replace("_", "").isalnum() or " " in key:
            continue
        try:
            out[key] = float(val.strip().split()[0])
        except (ValueError, IndexError):
            pass
    return out


# the device's VALU issue ceiling in wave64 instructions per second, measured once per
# run by valu_ceiling() (cubvh_valu_probe)
_VALU_PEAK = {}


def valu_ceiling(dev, iters=20000, reps=3):
    """Best of `reps` runs of each form of cubvh_valu_probe (8 waves per SIMD of
    independent fp32 fma chains; form 0 with VGPR and literal operands, form 1 with an
    SGPR operand), in wave64 instructions per second."""
    import ctypes

    import cubvh_amd

    lib = cubvh_amd._lib.load()
    for form, key in ((0, "wave_instr_per_s"), (1, "sgpr_wave_instr_per_s")):
        best = 0.0
        for _ in range(reps):
            wi = ctypes.c_double(0.0)
            ms = ctypes.c_float(0.0)
            cubvh_amd._lib.check(lib.cubvh_valu_probe(dev.index, form, iters, ctypes.byref(wi), ctypes.byref(ms)))
            if ms.value > 0:
                best = max(best, wi.value / (ms.value * 1e-3))
        _VALU_PEAK[key] = best
    return dict(_VALU_PEAK)


def roofline_block(name, kernel, kernel_s, units=None, algo_bytes_per_unit=None, copy_gbs=None):
    """Roofline of one kernel: HBM bytes per launch from the counter summary (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy of known size) over `kernel_s`, the launch's average
    duration measured here; its VALU issue rate (SQ_INSTS_VALU per launch of the same
    summary over `kernel_s`) against the ceiling valu_ceiling() measured in this run;
    active lanes per VALU instruction and the L2 hit rate of the summary; the algorithmic
    bytes (per-unit figure x units) over the same time when given."""
    f = profile_figures(name)
    traffic = f.get("traffic_bytes_per_launch")
    achieved = traffic / kernel_s / 1e9 if traffic and kernel_s > 0 else None
    blk = {"kernel": kernel, "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
           "achieved": achieved, "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
           "traffic": traffic, "kernel_ms": kernel_s * 1e3,
           "valu_lane_utilization": f.get("valu_lane_utilization"),
           "l2_hit_rate": f.get("l2_hit_rate"), "counters": f"profiles/{name}_counters.md" if f else None}
    insts = f.get("sq_insts_valu")
    peak = _VALU_PEAK.get("wave_instr_per_s")
    if insts and kernel_s > 0:
        rate = insts / kernel_s
        sg = _VALU_PEAK.get("sgpr_wave_instr_per_s")
        blk["valu"] = {"insts_per_launch": insts, "achieved_wave_instr_per_s": rate,
                       "peak_wave_instr_per_s": peak, "frac": rate / peak if peak else None,
                       "sgpr_form_wave_instr_per_s": sg, "frac_of_sgpr_form": rate / sg if sg else None,
                       "peak_source": "cubvh_valu_probe: 8 waves per SIMD, 16 independent fp32 fma chains; "
                                      "peak = VGPR and literal operands, sgpr_form = one SGPR operand (half rate)"}
    if copy_gbs and achieved is not None:
        blk["copy_GBs"] = copy_gbs
        blk["frac_of_copy"] = achieved / copy_gbs
    if units is not None and algo_bytes_per_unit is not None:
        algo = algo_bytes_per_unit * units / kernel_s / 1e9
        blk["algorithmic_bytes_per_unit"] = algo_bytes_per_unit
        blk["algorithmic_GBs"] = algo
        blk["algorithmic_frac"] = algo / HBM_PEAK_GBS
    vf = blk.get("valu", {}).get("frac")
    if vf is not None and vf >= 0.6:
        blk["limiter"] = "VALU issue (SQ_INSTS_VALU per second against the measured fp32 fma issue ceiling)"
    elif blk["frac"] is not None and blk["frac"] >= 0.5:
        blk["limiter"] = "HBM"
    elif vf is not None and vf >= 0.45:
        blk["limiter"] = ("VALU issue with latency (between the SGPR-operand and the VGPR-operand fma ceilings, "
                          "which differ by about 2x)")
    elif blk["frac"] is not None:
        blk["limiter"] = "latency (neither VALU issue nor HBM near its peak)"
    return blk


def _time_kernel(fn, steps, warmup):
    """Wall time (synchronized) and HIP-event time per call on the current stream."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        fn()
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall, float(np.mean(ev)) / 1e3


def cpu_baseline(v, f, pts, sample=1 << 21, sample_1t=1 << 17, threads=None):
    """Reference BVH walk restated on the CPU (oracle), on seeded samples of the grid:
    all host threads (up to 16) on `sample` points, one thread on `sample_1t` points."""
    from oracle import oracle as O

    threads = threads or min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    R = O.RefBvh(v, f)
    build_s = time.perf_counter() - t0
    idx = np.random.default_rng(123).choice(len(pts), sample, replace=False)
    sub = np.ascontiguousarray(pts[idx])
    R.udf(sub[:4096], return_uvw=False, nthreads=threads)  # warm caches / thread pool
    t0 = time.perf_counter()
    R.udf(sub, return_uvw=False, nthreads=threads)
    dt = time.perf_counter() - t0
    sub1 = np.ascontiguousarray(sub[:sample_1t])
    t0 = time.perf_counter()
    R.udf(sub1, return_uvw=False, nthreads=1)
    dt1 = time.perf_counter() - t0
    return {
        "value": sample / dt / 1e6,
        "unit": "Mqueries/s",
        "cores": threads,
        "kind": "port",
        "value_1thread": sample_1t / dt1 / 1e6,
        "build_s_1thread": build_s,
        "sample": f"{sample} of the 256^3 grid points (seeded) on {threads} OpenMP threads ({dt:.1f} s), "
                  f"the first {sample_1t} of them on 1 thread ({dt1:.1f} s); reference BVH4 build "
                  f"({build_s:.2f} s, 1 thread) and closest_triangle walk restated in C++ (oracle/cubvh_oracle.cpp)",
    }


def copy_bandwidth(dev, nbytes=1 << 30, reps=5):
    """Device-to-device copy rate (bytes read + written per second): the achievable HBM bandwidth."""
    import torch

    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / sec / 1e9


def fwd_bwd_leg(dev, steps, warmup):
    """Differentiable UDF: jittered level-6 sphere, 2^20 N(0, 0.4^2) points, loss = mean(udf^2)."""
    import torch

    import cubvh_amd
    from cubvh_amd import synthetic as S

    v6, f6 = S.jittered_sphere(6)
    bvh6 = cubvh_amd.cuBVH(v6, f6)
    n = 1 << 20
    P = torch.from_numpy(S.normal_points(n, 0.4, seed=0)).to(dev)
    V6 = torch.from_numpy(v6).to(dev)

    def fwd():
        bvh6.unsigned_distance(P, return_uvw=True)

    def fwdbwd():
        p = P.detach().requires_grad_(True)
        v = V6.detach().requires_grad_(True)
        d, _, _ = cubvh_amd.udf(bvh6, p, v)
        (d * d).mean().backward()

    _, tf = _time_kernel(fwd, steps, warmup)
    _, tfb = _time_kernel(fwdbwd, steps, warmup)
    # the backward kernel alone, through the C ABI on the current stream: dL/dd of
    # mean(d^2) = 2 d / n, gradients of the points and (accumulated) of the vertices
    import ctypes

    lib = cubvh_amd._lib.load()
    d, fi, uvw = bvh6.unsigned_distance(P, return_uvw=True)
    g = d * (2.0 / n)
    F6 = torch.from_numpy(f6).to(dev)
    gp = torch.empty_like(P)
    gv = torch.zeros_like(V6)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)

    def bwd():
        cubvh_amd._lib.check(lib.cubvh_udf_backward(dev.index, n, g.data_ptr(), P.data_ptr(), fi.data_ptr(),
                                                    uvw.data_ptr(), V6.data_ptr(), V6.shape[0], F6.data_ptr(),
                                                    F6.shape[0], gp.data_ptr(), gv.data_ptr(), stream))

    _, tb = _time_kernel(bwd, steps, warmup)
    return {"workload": "jittered icosphere level 6 (81,920 faces), 2^20 points N(0, 0.4^2), loss mean(udf^2)",
            "queries": n, "fwd_ms": tf * 1e3, "fwd_bwd_ms": tfb * 1e3, "ratio": tfb / tf,
            "fwd_Mq_s": n / tf / 1e6, "fwd_bwd_Mq_s": n / tfb / 1e6, "bwd_kernel_ms": tb * 1e3,
            "roofline_bwd": dict(roofline_block("udf_backward_level6",
                                                "query_terms + face sums + vertex corner table + vertex_sums (3 radix sorts)",
                                                tb, units=n, algo_bytes_per_unit=168),
                                 time_basis="whole cubvh_udf_backward call")}


def batched_leg(dev, steps, warmup, n_meshes=512, gpu_build=True):
    """512 meshes with uneven query sets: per-mesh calls vs one forest launch."""
    import torch

    import cubvh_amd
    from cubvh_amd import synthetic as S

    import ctypes
    from concurrent.futures import ThreadPoolExecutor

    lib = cubvh_amd._lib.load()
    specs = S.forest(n_meshes, seed=0)
    t0 = time.perf_counter()
    rng = np.random.default_rng(1)
    pts = [rng.uniform(-1, 1, size=(sp["n_queries"], 3)).astype(np.float32) for sp in specs]
    for lvl in sorted({sp["level"] for sp in specs}):
        S.icosphere_unit(lvl)  # shared topology, built once per level

    def mesh(sp):
        return S.perturbed_sphere(sp["level"], seed=sp["seed"], scale=sp["scale"])[:2]

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        meshes = list(ex.map(mesh, specs))
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        # the host build runs in the library without the GIL
        bvhs = list(ex.map(lambda m: cubvh_amd.cuBVH(*m), meshes))
    build_s = time.perf_counter() - t0
    counts = np.array([sp["n_queries"] for sp in specs])
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    P = torch.from_numpy(np.concatenate(pts)).to(dev)
    n = P.shape[0]
    off = torch.from_numpy(offsets).to(dev)
    slices = [P[offsets[m]:offsets[m + 1]] for m in range(n_meshes)]
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    face = torch.empty(n, dtype=torch.int64, device=dev)

    def plain():
        for b, q in zip(bvhs, slices):
            b.unsigned_distance(q)

    _, t_plain = _time_kernel(plain, steps, warmup)
    forest = cubvh_amd.cuBVHForest(bvhs)
    tuning, table = cubvh_amd.tune_forest(forest, P, off, reps=2, use_cache=False)
    forest.reserve(n, tuning[0], dev)
    _, t_forest = _time_kernel(lambda: forest.unsigned_distance(P, off, False, tuning, (dist, face, None)),
                               steps, warmup)
    graph = cubvh_amd.ForestGraph(forest, P, off, tuning=tuning)
    _, t_graph = _time_kernel(graph.replay, steps, warmup)
    # batched forward + hand-written backward (loss sum(d^2) over every mesh's queries)
    vcat, _ = forest.concatenated_mesh(dev)

    def fwdbwd():
        p = P.detach().requires_grad_(True)
        vv = vcat.detach().requires_grad_(True)
        d, _, _ = cubvh_amd.udf_batched(forest, p, off_host, vv, tuning=tuning)
        (d * d).sum().backward()

    off_host = torch.from_numpy(offsets)
    _, t_fwdbwd = _time_kernel(fwdbwd, steps, warmup)
    # the batched answers equal the per-mesh answers
    ok = True
    for m in range(0, n_meshes, 37):
        d_ref, f_ref, _ = bvhs[m].unsigned_distance(slices[m])
        ok &= bool(torch.equal(graph.dist[offsets[m]:offsets[m + 1]], d_ref))
        ok &= bool(torch.equal(graph.face[offsets[m]:offsets[m + 1]], f_ref))
    best = min(t_forest, t_graph)
    del graph
    res = {"workload": f"{n_meshes} perturbed icospheres, {int(forest.info()['n_faces'])} faces, "
                       f"{n} uniform points in [-1,1]^3 (4,096..262,144 per mesh)",
           "queries": int(n), "faces": int(forest.info()["n_faces"]), "mesh_generation_s": gen_s,
           "host_build_s": build_s,
           "plain_ms": t_plain * 1e3, "forest_ms": t_forest * 1e3, "graph_ms": t_graph * 1e3,
           "plain_Mq_s": n / t_plain / 1e6, "batched_Mq_s": n / best / 1e6, "speedup": t_plain / best,
           "fwd_bwd_ms": t_fwdbwd * 1e3, "fwd_bwd_Mq_s": n / t_fwdbwd / 1e6, "fwd_bwd_ratio": t_fwdbwd / t_forest,
           "tuning": list(tuning), "equal_to_per_mesh": ok,
           "roofline": dict(roofline_block("forest512", "forest_distance_kernel", t_forest),
                            time_basis="whole forest call (per-mesh Morton ordering + walk)")}
    if not gpu_build:
        return res
    # the same meshes built on the GPU (already resident there) and assembled on the device
    mesh_dev = [(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev)) for v, f in meshes]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # four builds in flight (one host thread and stream each)
    gbvhs = cubvh_amd.cuBVH.build_many(mesh_dev, build="gpu", device=dev, streams=4)
    gforest = ctypes.c_void_p()
    arr = (ctypes.c_void_p * n_meshes)(*[b._h.value for b in gbvhs])
    cubvh_amd._lib.check(lib.cubvh_forest_create(arr, n_meshes, ctypes.byref(gforest)))
    gpu_build_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream(dev).cuda_stream
    cubvh_amd._lib.check(lib.cubvh_forest_reserve(gforest, dev.index, n, tuning[0], stream))
    tun = (ctypes.c_int32 * len(tuning))(*tuning)

    def gpu_forest():
        cubvh_amd._lib.check(lib.cubvh_forest_unsigned_distance(gforest, dev.index, n, P.data_ptr(), off.data_ptr(),
                                                                dist.data_ptr(), face.data_ptr(), None, tun,
                                                                len(tuning), stream))

    _, t_gforest = _time_kernel(gpu_forest, steps, warmup)
    lib.cubvh_forest_destroy(gforest)
    del gbvhs
    res["gpu_build_s"] = gpu_build_s
    res["gpu_built_forest_Mq_s"] = n / t_gforest / 1e6
    return res


def hashtable_leg(dev, steps, warmup, with_cpu=True):
    """Sparse-voxel hash table (cuHashTable): the distinct surface voxels of a sphere
    at 1024^3 (from 2^23 seeded surface samples) inserted in random order, then 2^24
    queries, half of them stored keys and half uniform voxels (mostly absent)."""
    import torch

    import cubvh_amd

    gen = torch.Generator(device=dev).manual_seed(0)
    g = torch.randn(1 << 23, 3, device=dev, generator=gen)
    g = g / g.norm(dim=1, keepdim=True)
    v = ((g * 0.45 + 0.5) * 1024).floor().to(torch.int64).clamp_(0, 1023)
    key = torch.unique((v[:, 0] << 20) | (v[:, 1] << 10) | v[:, 2])
    key = key[torch.randperm(key.numel(), device=dev, generator=gen)]
    C = torch.stack([key >> 20, (key >> 10) & 1023, key & 1023], dim=1).to(torch.int32).contiguous()
    n = C.shape[0]
    m = 1 << 24
    hits = C[torch.randint(0, n, (m // 2,), device=dev, generator=gen)]
    miss = torch.randint(0, 1024, (m - m // 2, 3), device=dev, generator=gen, dtype=torch.int32)
    Q = torch.cat([hits, miss])[torch.randperm(m, device=dev, generator=gen)].contiguous()
    ht = cubvh_amd.cuHashTable()
    ht.build(C)  # allocates the table; the timed builds reuse it
    _, tb = _time_kernel(lambda: ht.impl.build(C), steps, warmup)
    _, ts = _time_kernel(lambda: ht.impl.search(Q), steps, warmup)
    found = int((ht.search(Q) >= 0).sum())
    res = {"workload": f"{n} distinct sphere-surface voxels of a 1024^3 grid (random insertion order), {m} queries "
                       f"(half stored keys)", "keys": n, "queries": m, "found": found,
           "build_ms": tb * 1e3, "build_Mkeys_s": n / tb / 1e6, "search_ms": ts * 1e3, "search_Mq_s": m / ts / 1e6,
           "search_GB_s_32B_per_query": 32 * m / ts / 1e9,
           "search_frac_of_hbm_peak_32B_per_query": 32 * m / ts / 1e9 / HBM_PEAK_GBS}
    if with_cpu:
        from oracle import oracle as O

        c_np = C.cpu().numpy()
        q_np = Q[: 1 << 21].cpu().numpy()
        kind = "reference" if O.ref_hash_available() else "port"
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_hash_build_search(c_np, q_np[:0])
        else:
            O.hash_build(c_np)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_hash_build_search(c_np, q_np)
        else:
            O.hash_lookup(c_np, q_np)
        t_both = time.perf_counter() - t0
        t_search = max(t_both - t_build, 1e-9)
        res["cpu_baseline"] = {"kind": kind, "cores": 1, "build_Mkeys_s": n / t_build / 1e6,
                               "search_Mq_s": len(q_np) / t_search / 1e6,
                               "sample": f"all {n} keys, the first {len(q_np)} queries, one thread: the "
                                         f"reference's CPU table (include/cpu/hashtable.h)"
                                         if kind == "reference" else "the oracle's restatement, one thread"}
    return res


def smc_cells(dev, res=1024):
    """Cells of a res^3 lattice cut by a bumpy sphere (the analytic field
    |p - c| - 0.42 res - 6 sin(x/40) sin(y/40) sin(z/40) in lattice units), selected as
    the reference's remesh script selects them (corners of both signs), with their 8
    corner values in the reference's corner order."""
    import torch

    ax = torch.arange(res + 1, device=dev, dtype=torch.float32)
    c = res / 2 + 0.3
    x, y, z = ax.view(-1, 1, 1), ax.view(1, -1, 1), ax.view(1, 1, -1)
    f = torch.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) - 0.42 * res
    f -= 6.0 * torch.sin(x / 40) * torch.sin(y / 40) * torch.sin(z / 40)
    inside = f < 0
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    shifted = [inside[a:a + res, b:b + res, d:d + res] for a, b, d in offs]
    some = shifted[0].clone()
    every = shifted[0].clone()
    for t in shifted[1:]:
        some |= t
        every &= t
    del shifted, inside
    idx = torch.nonzero(some & ~every)
    corners = torch.stack([f[idx[:, 0] + a, idx[:, 1] + b, idx[:, 2] + d] for a, b, d in offs], dim=1)
    del f
    return idx.to(torch.int32).contiguous(), corners.contiguous()


def smc_leg(dev, steps, warmup, with_cpu=True):
    """Sparse marching cubes over the cut cells of a 1024^3 lattice; one step is one
    cubvh_amd.sparse_marching_cubes call (inputs resident on the GPU, outputs allocated
    by the call, as in the reference's API)."""
    import torch

    import cubvh_amd

    C, V = smc_cells(dev)
    n = C.shape[0]
    wall, _ = _time_kernel(lambda: cubvh_amd.sparse_marching_cubes(C, V, 0.0), steps, warmup)
    verts, tris = cubvh_amd.sparse_marching_cubes(C, V, 0.0)
    wall_c, _ = _time_kernel(lambda: cubvh_amd.sparse_marching_cubes(C, V, 0.0, ensure_consistency=True),
                             max(2, steps // 2), 1)
    res = {"workload": f"{n} cut cells of a 1024^3 lattice (bumpy sphere), iso 0", "cells": n,
           "vertices": int(verts.shape[0]), "triangles": int(tris.shape[0]),
           "ms": wall / steps * 1e3, "Mcells_s": n * steps / wall / 1e6,
           "ensure_consistency_ms": wall_c / max(2, steps // 2) * 1e3}
    if with_cpu:
        from oracle import oracle as O

        sample = 1 << 20
        c_np = C[:sample].cpu().numpy()
        v_np = V[:sample].cpu().numpy()
        kind = "reference" if O.ref_smc_available() else "port"
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_smc_cpu(c_np, v_np, 0.0)
        else:
            O.smc(c_np, v_np, 0.0)
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"kind": kind, "cores": 1, "Mcells_s": len(c_np) / t / 1e6,
                               "sample": f"the first {len(c_np)} cells, one thread: "
                                         + ("the reference's CPU extraction (include/cpu/spmc.h)"
                                            if kind == "reference" else "the oracle's restatement")}
    return res


def run_extras(dev, steps, warmup, out, gpu_build=True):
    import torch

    import cubvh_amd
    from cubvh_amd import synthetic as S

    # watertight SDF on the headline workload
    v8, f8, _ = S.perturbed_sphere(8, seed=0)
    bvh8 = cubvh_amd.cuBVH(v8, f8)
    G = torch.from_numpy(S.grid_points(256)).to(dev)
    _, t = _time_kernel(lambda: bvh8.signed_distance(G, mode="watertight"), steps, 1)
    out["sdf_watertight_grid256"] = {"Mq_s": G.shape[0] / t / 1e6, "ms": t * 1e3,
                                     "roofline": dict(roofline_block("sdf_watertight_level8_grid256",
                                                                     "distance_packet_kernel<1>", t),
                                                      time_basis="whole call (Morton ordering + walk)")}

    # the same mesh built on the GPU (linear BVH): build time and UDF rate on that tree
    V8 = torch.from_numpy(v8).to(dev)
    F8 = torch.from_numpy(f8).to(dev)
    lib = cubvh_amd._lib.load()
    import ctypes

    stream = torch.cuda.current_stream(dev).cuda_stream
    builds = (("gpu_build_level8", 0, "gpu"), ("gpu_build_level8_sah", 0x200, "gpu_sah"),
              ("gpu_build_level8_lbvh", 0x100, "gpu_lbvh"))
    for name, flags, build in builds if gpu_build else ():
        times = []
        for _ in range(3):
            h = ctypes.c_void_p()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cubvh_amd._lib.check(lib.cubvh_bvh_create_gpu(dev.index, V8.data_ptr(), V8.shape[0], F8.data_ptr(),
                                                          F8.shape[0], flags, ctypes.byref(h), stream))
            times.append(time.perf_counter() - t0)
            lib.cubvh_bvh_destroy(h)
        gbvh = cubvh_amd.cuBVH(v8, f8, build=build)
        _, t = _time_kernel(lambda: gbvh.unsigned_distance(G), steps, 1)
        out[name] = {"build_ms": min(times) * 1e3, "udf_Mq_s": G.shape[0] / t / 1e6, "tree": gbvh.info()}
        del gbvh
    del V8, F8

    # floodfill of the occupancy udf < 2/256
    d, _, _ = bvh8.unsigned_distance(G)
    occ = (d < 2.0 / 256).reshape(1, 256, 256, 256)
    _, t = _time_kernel(lambda: cubvh_amd.floodfill(occ), steps, 1)
    out["floodfill_256"] = {"ms": t * 1e3, "Mvoxels_s": occ.numel() / t / 1e6,
                            "GB_s_5B_per_voxel": 5 * occ.numel() / t / 1e9,
                            "roofline": dict(roofline_block("floodfill_grid256", "ff_tiles + 3 ff_merge_axis + ff_flatten",
                                                            t, units=occ.numel(), algo_bytes_per_unit=5),
                                             time_basis="whole call (5 kernel launches)")}
    del G, d, occ

    # soup: camera rays (16 x 1920 x 1080) and raystab on 2^22 points
    vs, fs = S.soup(8, seed=0)
    t0 = time.perf_counter()
    bvhs = cubvh_amd.cuBVH(vs, fs)
    out["soup_build_s"] = time.perf_counter() - t0
    ro, rd = S.camera_rays()
    RO = torch.from_numpy(ro).to(dev)
    RD = torch.from_numpy(rd).to(dev)
    _, t = _time_kernel(lambda: bvhs.ray_trace(RO, RD), max(2, steps // 2), 1)
    out["rays_soup_1080p_x16"] = {"Mrays_s": RO.shape[0] / t / 1e6, "ms": t * 1e3,
                                  "roofline": dict(roofline_block("rays_soup_1080p_x16", "ray_packet_kernel", t),
                                                   time_basis="whole call (one kernel launch)")}
    del RO, RD
    Q = torch.from_numpy(S.uniform_points(1 << 22, seed=0)).to(dev)
    _, t = _time_kernel(lambda: bvhs.signed_distance(Q, mode="raystab"), max(2, steps // 2), 1)
    _, tu = _time_kernel(lambda: bvhs.unsigned_distance(Q), max(2, steps // 2), 1)
    # the stab rays' kernel: the raystab call minus the unsigned call it starts with
    out["raystab_soup_4M"] = {"Mq_s": Q.shape[0] / t / 1e6, "ms": t * 1e3, "udf_part_ms": tu * 1e3,
                              "roofline": dict(roofline_block("raystab_soup_4M", "stab_init + stab_first + 3 x (stab_lines + stab_keep) + stab_negate", max(t - tu, 1e-9)),
                                               time_basis="raystab call minus the unsigned call")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--extras", action="store_true", help="run the extra legs also with --headline-only")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the extra legs (watertight SDF, floodfill, rays, raystab, GPU builds)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--leaf", type=int, default=0, help="leaf_size argument of cuBVH (0 = library default)")
    ap.add_argument("--no-gpu-build", action="store_true",
                    help="skip the GPU tree builds of the batched and extra legs (for counter passes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device-to-device tensor copy first (reference for HBM counters)")
    ap.add_argument("--hashtable-only", action="store_true",
                    help="run only the sparse-voxel hash table leg and print its JSON")
    ap.add_argument("--smc-only", action="store_true",
                    help="run only the sparse marching cubes leg and print its JSON")
    ap.add_argument("--valu-probe", action="store_true",
                    help="run only the VALU issue-ceiling probe (cubvh_valu_probe) and print its rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the query path runs on one "
              f"GPU; queries would shard across GPUs with a replicated tree, which is not implemented)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import cubvh_amd
    from cubvh_amd import synthetic as S

    dev = torch.device("cuda:0")
    if args.valu_probe:
        print(json.dumps({"valu_ceiling": valu_ceiling(dev)}), flush=True)
        return
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads and writes exactly 1 GiB
        torch.cuda.synchronize()
        del src, dst

    if args.hashtable_only:
        print(json.dumps(hashtable_leg(dev, args.steps, args.warmup, with_cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.smc_only:
        print(json.dumps(smc_leg(dev, args.steps, args.warmup, with_cpu=not args.no_cpu_baseline)), flush=True)
        return
    v, f, meta = S.perturbed_sphere(8, seed=0)
    t0 = time.perf_counter()
    bvh = cubvh_amd.cuBVH(v, f, leaf_size=args.leaf)
    build_s = time.perf_counter() - t0
    pts = S.grid_points(args.res)
    P = torch.from_numpy(pts).to(dev)
    n = P.shape[0]
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    face = torch.empty(n, dtype=torch.int64, device=dev)
    lib = cubvh_amd._lib.load()
    import ctypes

    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)

    order = torch.empty(n, dtype=torch.int32, device=dev)

    def order_queries():
        cubvh_amd._lib.check(lib.cubvh_morton_order(dev.index, n, P.data_ptr(), order.data_ptr(), stream))

    def walk():
        cubvh_amd._lib.check(lib.cubvh_distance_ordered(bvh._h, dev.index, n, 0, P.data_ptr(), order.data_ptr(),
                                                        dist.data_ptr(), face.data_ptr(), None, stream))

    def dropin():
        cubvh_amd._lib.check(lib.cubvh_unsigned_distance(bvh._h, dev.index, n, P.data_ptr(), dist.data_ptr(),
                                                         face.data_ptr(), None, stream))

    # One step = the whole query pass: order the points, walk them (what
    # cubvh_unsigned_distance does in one call).  HIP events on the launch stream
    # bracket the walk, i.e. the distance_packet_kernel launch alone.
    for _ in range(args.warmup):
        order_queries()
        walk()
    torch.cuda.synchronize()
    k0 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    k1 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        order_queries()
        k0[i].record()
        walk()
        k1[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = float(np.mean([a.elapsed_time(b) for a, b in zip(k0, k1)])) / 1e3
    value = n * args.steps / wall / 1e6
    dropin_wall, _ = _time_kernel(dropin, args.steps, 1)
    valu_ceiling(dev)
    ab = _algo_bytes()
    roof = None
    if args.res == 256:
        copy_gbs = copy_bandwidth(dev)
        roof = roofline_block("udf_level8_grid256", "distance_packet_kernel<0>", ev, copy_gbs=copy_gbs)
        if ab and "udf_level8_grid256" in ab["workloads"]:
            # the reference's own walk (bvh.cu:300-353) on this workload, priced at 24 B per
            # box test and 36 B per triangle test: the work this walk replaces, not HBM traffic
            bpq = ab["workloads"]["udf_level8_grid256"]["bytes_per_query"]
            roof["reference_walk_bytes_per_query"] = bpq
            roof["reference_walk_equiv_GBs"] = bpq * n / ev / 1e9
    result = {
        "metric": "M UDF queries/s @1.3M tris 256^3 (plain vs fwd+bwd vs batched), % HBM roofline",
        "value": value,
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"dense unsigned_distance, {args.res}^3 grid on [-1,1]^3, perturbed icosphere "
                               f"level 8 ({f.shape[0]} faces, seed 0)", "queries": n, "build_s": build_s, "tree": bvh.info(),
                   "dropin_call_Mq_s": n * args.steps / dropin_wall / 1e6},
        "roofline": roof,
    }
    if not args.headline_only:
        result["fwd_bwd"] = fwd_bwd_leg(dev, max(3, args.steps), args.warmup)
        result["batched"] = batched_leg(dev, max(2, args.steps // 3), 1, gpu_build=not args.no_gpu_build)
        result["hashtable"] = hashtable_leg(dev, max(3, args.steps), 1, with_cpu=not args.no_cpu_baseline)
        result["sparse_marching_cubes"] = smc_leg(dev, max(3, args.steps), 1, with_cpu=not args.no_cpu_baseline)
    extra = {}
    if args.extras or not (args.no_extras or args.headline_only):
        run_extras(dev, max(2, args.steps // 2), 1, extra, gpu_build=not args.no_gpu_build)
        result["extra"] = extra
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(v, f, pts)
        # the timed pass's own answers on a seeded sample against the oracle (canonical ties)
        from oracle import oracle as O

        idx = np.sort(np.random.default_rng(7).choice(n, 1 << 16, replace=False))
        od, of, _, _ = O.RefBvh(v, f).udf(pts[idx], return_uvw=False, tie_mode=1,
                                          nthreads=min(os.cpu_count() or 1, 16))
        gd = dist.cpu().numpy()[idx]
        gf = face.cpu().numpy()[idx]
        result["parity"] = {"sample": f"{len(idx)} seeded points of the timed pass's output vs the oracle",
                            "distance_mismatches": int((gd != od).sum()), "face_mismatches": int((gf != of).sum()),
                            "max_abs_distance_diff": float(np.abs(gd.astype(np.float64) - od).max())}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
