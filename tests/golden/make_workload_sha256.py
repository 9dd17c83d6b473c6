"""Writes tests/golden/workload_sha256.json: SHA-256 digests of every array the
synthetic workload generators (cubvh_amd/synthetic.py) produce for the five benchmark
configurations, so that a change of numpy (its random streams) or of a generator shows
up as a failing test instead of silently changing every measured workload.

    python tests/golden/make_workload_sha256.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(f"{a.dtype.str}{a.shape}".encode())
        h.update(a.tobytes())
    return h.hexdigest()


def workloads():
    """(name, description, zero-argument function returning the arrays) per workload."""
    from cubvh_amd import synthetic as S

    def forest():
        specs = S.forest(512, seed=0)
        h = hashlib.sha256(json.dumps(specs, sort_keys=True).encode())
        rng = np.random.default_rng(1)
        for sp in specs:
            v, f, _ = S.perturbed_sphere(sp["level"], seed=sp["seed"], scale=sp["scale"])
            p = rng.uniform(-1, 1, size=(sp["n_queries"], 3)).astype(np.float32)
            h.update(digest(v, f, p).encode())
        return h.hexdigest()

    return [
        ("config1_mesh", "icosphere level 4, radius 0.5", lambda: digest(*S.icosphere(4, 0.5))),
        ("config1_points", "65,536 uniform points, seed 0", lambda: digest(S.uniform_points(65536, seed=0))),
        ("config1_rays", "65,536 sphere rays, seed 0", lambda: digest(*S.sphere_rays(65536, seed=0))),
        ("config2_mesh", "perturbed icosphere level 8, seed 0", lambda: digest(*S.perturbed_sphere(8, seed=0)[:2])),
        ("config2_grid", "256^3 grid on [-1, 1]^3", lambda: digest(S.grid_points(256))),
        ("config3_soup", "level-8 soup, seed 0", lambda: digest(*S.soup(8, seed=0))),
        ("config3_camera_rays", "16 poses x 1920 x 1080, seed 0", lambda: digest(*S.camera_rays())),
        ("config3_points", "2^22 uniform points, seed 0", lambda: digest(S.uniform_points(1 << 22, seed=0))),
        ("config4_mesh", "jittered icosphere level 6, seed 0", lambda: digest(*S.jittered_sphere(6))),
        ("config4_points", "2^20 points N(0, 0.4^2), seed 0", lambda: digest(S.normal_points(1 << 20, 0.4, seed=0))),
        ("config5_forest", "512 meshes + their query sets (specs seed 0, points seed 1)", forest),
    ]


def main():
    out = {"numpy": np.__version__, "digests": {}}
    for name, desc, fn in workloads():
        out["digests"][name] = {"what": desc, "sha256": fn()}
        print(name, out["digests"][name]["sha256"], flush=True)
    (ROOT / "tests" / "golden" / "workload_sha256.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
