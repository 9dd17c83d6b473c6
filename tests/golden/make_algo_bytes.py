"""Computes the algorithmic bytes per query of the benchmark workloads with the oracle's
instrumented restatement of the reference BVH walk, and writes algo_bytes.json.

The reference's work is charged at 24 bytes per child box test (6 fp32), 36 bytes per
triangle test (9 fp32), plus the query's own input and output bytes:

  UDF query        12 (point) + 4 (distance) + 8 (int64 face id) [+ 12 uvw]
                   + 24 * N_box + 36 * N_tri
  ray              24 (origin, direction) + 12 (position) + 4 (depth) + 8 (face id)
                   + 24 * N_box + 36 * N_tri
  watertight SDF   UDF bytes + the normal-averaging walk's 24 * N_box + 36 * N_tri
  backward query   168 (see cubvh_amd/csrc/grad.hip: g, p, face id, uvw, face indices,
                   3 vertices, grad_p write, 3 x 12-byte atomic read-modify-write)

N_box and N_tri are the averages over all queries of the workload (4 box tests per
internal node popped, every triangle of every leaf popped), counted by
oracle.RefBvh (tie rule of the reference).  Run from the repository root:
    python tests/golden/make_algo_bytes.py [--quick]
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))


def main():
    from cubvh_amd import synthetic as S
    from oracle import oracle as O

    quick = "--quick" in sys.argv
    out = {"bytes_per_box_test": 24, "bytes_per_tri_test": 36, "workloads": {}}

    # dense UDF on the perturbed level-8 sphere, 256^3 grid
    t0 = time.time()
    v, f, meta = S.perturbed_sphere(8, seed=0)
    R = O.RefBvh(v, f)
    pts = S.grid_points(256)
    if quick:
        pts = pts[np.random.default_rng(0).choice(len(pts), 1 << 18, replace=False)]
    _, _, _, c = R.udf(pts, return_uvw=False, tie_mode=0)
    n = len(pts)
    nb, nt = c[0] / n, c[1] / n
    out["workloads"]["udf_level8_grid256"] = {
        "queries": int(n), "n_box": nb, "n_tri": nt,
        "bytes_per_query": 12 + 4 + 8 + 24 * nb + 36 * nt,
        "mesh": "perturbed icosphere level 8 (1,310,720 faces), seed 0", "queries_def": "256^3 grid on [-1,1]",
    }
    print("udf", out["workloads"]["udf_level8_grid256"], f"{time.time() - t0:.1f}s", flush=True)

    # watertight SDF on the same grid (sampled: the normal walk is short)
    sub = pts[np.random.default_rng(1).choice(len(pts), 1 << 18, replace=False)]
    _, _, _, c = R.sdf(sub, 0, return_uvw=False)
    m = len(sub)
    out["workloads"]["sdf_watertight_level8_grid256"] = {
        "queries": int(m), "n_box": c[0] / m, "n_tri": c[1] / m, "n_box_sign": c[2] / m, "n_tri_sign": c[3] / m,
        "bytes_per_query": 24 + 24 * (c[0] + c[2]) / m + 36 * (c[1] + c[3]) / m,
    }
    print("sdf", out["workloads"]["sdf_watertight_level8_grid256"], flush=True)

    # differentiable UDF: jittered level-6 sphere, 2^20 N(0, 0.4^2) points
    v6, f6 = S.jittered_sphere(6)
    R6 = O.RefBvh(v6, f6)
    p6 = S.normal_points(1 << 20, 0.4, seed=0)
    _, _, _, c = R6.udf(p6, return_uvw=True, tie_mode=0)
    m = len(p6)
    fwd = 12 + 4 + 8 + 12 + 24 * c[0] / m + 36 * c[1] / m
    out["workloads"]["udf_fwd_bwd_level6"] = {
        "queries": int(m), "n_box": c[0] / m, "n_tri": c[1] / m, "bytes_per_query_fwd": fwd,
        "bytes_per_query_bwd": 168, "bytes_per_query": fwd + 168,
    }
    print("fwdbwd", out["workloads"]["udf_fwd_bwd_level6"], flush=True)

    (HERE / "algo_bytes.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote algo_bytes.json")


if __name__ == "__main__":
    main()
