"""Autotuner of the batched forest launch.

Tries the launch configurations of :meth:`cuBVHForest.unsigned_distance` on the
caller's own batch -- Morton-prefix bits used to order the queries inside each mesh
(the query-sort granularity), persistent grid or one wave per tile, waves per CU of
the persistent grid, then waves per workgroup (the block size: 64 to 512 lanes) for
the best of those, then for a one-wave-per-tile best the runs of consecutive
workgroups dealt to each XCD (so that neighbouring tiles share an L2) -- times each
with HIP events and caches the winner per (device,
query-count bucket, mesh-count bucket, face-count bucket) in a JSON file
(``$CUBVH_AMD_TUNE_CACHE`` or ``~/.cache/cubvh_amd/autotune.json``).

The traversal stack of the forest's packet walk is not a knob: it holds one entry per
pending subtree in three registers spread over the wave's 64 lanes, so its depth is
the tree's (at most 3 x depth + 1 <= 64 entries), whatever the launch shape.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import torch

CANDIDATES = [
    (bits, persistent, wpc, 4, 0)
    for bits in (0, 6, 9, 12, 15)
    for persistent, wpc in ((0, 8), (1, 4), (1, 8), (1, 16))
]
WAVES_PER_BLOCK = (1, 2, 4, 8)
XCD_CHUNKS = (0, 4, 16, 64, 256, 1024, 4096)
MAX_BINS = 1 << 30  # n_meshes << bits (the library refuses more Morton bins)


def _cache_path() -> Path:
    p = os.environ.get("CUBVH_AMD_TUNE_CACHE")
    if p:
        return Path(p)
    return Path(os.path.expanduser("~")) / ".cache" / "cubvh_amd" / "autotune.json"


def _bucket(x: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, x)))))


def _key(forest, n: int, device) -> str:
    name = torch.cuda.get_device_name(device)
    info = forest.info()
    return f"{name}|n{_bucket(n)}|m{_bucket(info['n_meshes'])}|f{_bucket(info['n_faces'])}"


def _load():
    p = _cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(cache):
    p = _cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home only costs re-tuning


def tune_forest(forest, points, offsets, candidates=None, reps=3, use_cache=True):
    """Returns (best tuning tuple, {tuning: ms}) for this forest and batch."""
    points = points if points.is_cuda else points.cuda()
    dev = points.device
    n = points.view(-1, 3).shape[0]
    key = _key(forest, n, dev)
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        return tuple(cache[key]["best"]), cache[key]["ms"]
    n_meshes = forest.info()["n_meshes"]
    cands = [tuple(c) for c in (candidates or CANDIDATES) if (n_meshes << c[0]) <= MAX_BINS]
    if not cands:
        raise ValueError("no candidate fits the Morton-bin limit (n_meshes << bits <= 2^30)")
    forest.reserve(n, max(c[0] for c in cands), dev)
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    face = torch.empty(n, dtype=torch.int64, device=dev)
    out = (dist, face, None)
    times = {}

    def measure(cfg):
        forest.unsigned_distance(points, offsets, False, cfg, out)  # warm up
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            forest.unsigned_distance(points, offsets, False, cfg, out)
        ev[1].record()
        torch.cuda.synchronize(dev)
        times[cfg] = ev[0].elapsed_time(ev[1]) / reps

    for cfg in cands:
        measure(cfg)
    best = min(times, key=times.get)
    if candidates is None:
        # second stage: the workgroup size of the best launch
        for wpb in WAVES_PER_BLOCK:
            cfg = best[:3] + (wpb, 0)
            if cfg not in times:
                measure(cfg)
        best = min(times, key=times.get)
        if best[1] == 0:  # third stage: the XCD chunk of a one-wave-per-tile grid
            for xc in XCD_CHUNKS:
                cfg = best[:4] + (xc,)
                if cfg not in times:
                    measure(cfg)
            best = min(times, key=times.get)
    ms = {",".join(map(str, k)): v for k, v in times.items()}
    if use_cache:
        cache[key] = {"best": list(best), "ms": ms}
        _store(cache)
    return best, ms
