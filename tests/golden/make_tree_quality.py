"""Tree quality as the reference's own walk sees it: N_box and N_tri per query.

Every tree this library builds (host SAH and host median, GPU median, GPU SAH, GPU
linear BVH), at the reference's leaf size (at most 8 faces) and, for the host SAH
tree, at this library's default (16), is exported to the reference's state_dict layout and walked by the
oracle's instrumented restatement of the reference UDF traversal (src/bvh.cu:300-353,
oracle.RefBvh: 4 box tests per internal node popped, every triangle of every leaf
popped), beside the reference's own build (src/bvh.cu:515-620, restated in the
oracle).  Workload: the perturbed level-8 icosphere (1,310,720 faces, seed 0), a
seeded sample of 2^18 points of the 256^3 grid on [-1, 1]^3.  The distances must equal
the reference tree's (a tree changes the work, not the answer).  Bytes per query are
charged as in make_algo_bytes.py: 24 per box test, 36 per triangle test, plus 24 of
query input and output.

Needs a GPU for the GPU builds.  Run from the repository root:
    python tests/golden/make_tree_quality.py [out.json]
(default output: tests/golden/tree_quality.json).
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))


def measure(tree, pts):
    _, _, _, c = tree.udf(pts, return_uvw=False, tie_mode=0)
    n = len(pts)
    nb, nt = c[0] / n, c[1] / n
    return {"n_box": nb, "n_tri": nt, "bytes_per_query": 24 + 24 * nb + 36 * nt}


def main():
    import cubvh_amd
    from cubvh_amd import synthetic as S
    from oracle import oracle as O

    out_path = Path(sys.argv[1]) if len(sys.argv) > 1 else HERE / "tree_quality.json"
    v, f, _ = S.perturbed_sphere(8, seed=0)
    pts = S.grid_points(256)
    pts = pts[np.random.default_rng(0).choice(len(pts), 1 << 18, replace=False)]
    out = {"mesh": "perturbed icosphere level 8 (1,310,720 faces), seed 0",
           "queries": "2^18 points of the 256^3 grid on [-1,1] (seeded sample)", "trees": {}}
    t0 = time.time()
    ref = O.RefBvh(v, f)
    d_ref, _, _, _ = ref.udf(pts, return_uvw=False, tie_mode=0)
    out["trees"]["reference build (4-ary median, leaves <= 8)"] = measure(ref, pts)
    print("reference", out["trees"]["reference build (4-ary median, leaves <= 8)"], f"{time.time() - t0:.1f}s",
          flush=True)
    for name, kw in (("host SAH, leaves <= 8", {"leaf_size": 8}),
                     ("host object median, leaves <= 8", {"leaf_size": 0x108}),
                     ("GPU object median (build='gpu'), leaves <= 8", {"build": "gpu", "leaf_size": 8}),
                     ("GPU SAH (build='gpu_sah'), leaves <= 8", {"build": "gpu_sah", "leaf_size": 8}),
                     ("GPU linear BVH (build='gpu_lbvh'), leaves <= 8", {"build": "gpu_lbvh", "leaf_size": 8}),
                     ("host SAH, leaves <= 16 (default)", {})):
        b = cubvh_amd.cuBVH(v, f, **kw)
        sd = b.state_dict(layout="reference")
        tree = O.RefBvh.from_arrays(sd["nodes"].numpy(), sd["triangles"].numpy())
        d, _, _, _ = tree.udf(pts, return_uvw=False, tie_mode=0)
        np.testing.assert_array_equal(d, d_ref)
        out["trees"][name] = measure(tree, pts)
        print(name, out["trees"][name], flush=True)
    out_path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", out_path)


if __name__ == "__main__":
    main()
