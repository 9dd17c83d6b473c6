"""Writes the committed fixtures under tests/golden/.

dodecahedron.npz  the mesh of the reference's own tests (test/unsigned_distance.py:9-76):
                  20 vertices on the unit sphere, 36 faces (1-based indices in the
                  reference, stored 0-based here), vertices normalized to radius 1 as
                  the reference's create_dodecahedron(radius=1) does.
pcg32_kat.json    first outputs of PCG32 from an independent pure-Python restatement
                  of O'Neill's PCG-XSH-RR 64/32 with the reference's default state and
                  stream (include/gpu/pcg32.h:33-35), for jumps 0, 1, 2, 1000 and 2^32.
config1_subset.npz  4096 points and 4096 rays of the small workload (icosphere level 4,
                  radius 0.5, seeds 0) with the oracle's fp32 brute-force answers and
                  the FP64 distances, so that a run on any machine is compared with the
                  same numbers.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

DODECA_V = [
    -0.57735, -0.57735, 0.57735, 0.934172, 0.356822, 0, 0.934172, -0.356822, 0,
    -0.934172, 0.356822, 0, -0.934172, -0.356822, 0, 0, 0.934172, 0.356822,
    0, 0.934172, -0.356822, 0.356822, 0, -0.934172, -0.356822, 0, -0.934172,
    0, -0.934172, -0.356822, 0, -0.934172, 0.356822, 0.356822, 0, 0.934172,
    -0.356822, 0, 0.934172, 0.57735, 0.57735, -0.57735, 0.57735, 0.57735, 0.57735,
    -0.57735, 0.57735, -0.57735, -0.57735, 0.57735, 0.57735, 0.57735, -0.57735, -0.57735,
    0.57735, -0.57735, 0.57735, -0.57735, -0.57735, -0.57735,
]
DODECA_F = [
    19, 3, 2, 12, 19, 2, 15, 12, 2, 8, 14, 2, 18, 8, 2, 3, 18, 2, 20, 5, 4, 9, 20, 4,
    16, 9, 4, 13, 17, 4, 1, 13, 4, 5, 1, 4, 7, 16, 4, 6, 7, 4, 17, 6, 4, 6, 15, 2,
    7, 6, 2, 14, 7, 2, 10, 18, 3, 11, 10, 3, 19, 11, 3, 11, 1, 5, 10, 11, 5, 20, 10, 5,
    20, 9, 8, 10, 20, 8, 18, 10, 8, 9, 16, 7, 8, 9, 7, 14, 8, 7, 12, 15, 6, 13, 12, 6,
    17, 13, 6, 13, 1, 11, 12, 13, 11, 19, 12, 11,
]

MASK64 = (1 << 64) - 1
MULT = 0x5851F42D4C957F2D


def pcg32_stream(advance, n, state=0x853C49E6748FEA9B, inc=0xDA3E39CB94B95BDB):
    # jump ahead by repeated squaring of the affine step x -> MULT*x + inc (mod 2^64)
    acc_m, acc_p, cur_m, cur_p, d = 1, 0, MULT, inc, advance
    while d:
        if d & 1:
            acc_m = (acc_m * cur_m) & MASK64
            acc_p = (acc_p * cur_m + cur_p) & MASK64
        cur_p = ((cur_m + 1) * cur_p) & MASK64
        cur_m = (cur_m * cur_m) & MASK64
        d >>= 1
    state = (acc_m * state + acc_p) & MASK64
    out = []
    for _ in range(n):
        old = state
        state = (old * MULT + inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        out.append(((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF)
    return out


def main():
    v = np.array(DODECA_V, dtype=np.float64).reshape(-1, 3)
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(DODECA_F, dtype=np.int64).reshape(-1, 3) - 1
    np.savez(HERE / "dodecahedron.npz", vertices=v, faces=f)

    kat = {str(a): pcg32_stream(a, 8) for a in (0, 1, 2, 1000, 1 << 32)}
    (HERE / "pcg32_kat.json").write_text(json.dumps(kat, indent=1) + "\n")

    from cubvh_amd import synthetic as S
    from oracle import oracle as O

    V, F = S.icosphere(4, 0.5)
    pts = S.uniform_points(65536, seed=0)[:4096]
    ro, rd = S.sphere_rays(65536, seed=0)
    ro, rd = ro[:4096], rd[:4096]
    d, fi, uvw = O.udf_brute(V, F, pts)
    d64, f64, cp64, gap = O.udf_f64(V, F, pts)
    pos, rf, rt = O.ray_brute(V, F, ro, rd)
    np.savez_compressed(HERE / "config1_subset.npz", points=pts, dist=d, face=fi, uvw=uvw, dist64=d64, gap64=gap,
                        rays_o=ro, rays_d=rd, ray_pos=pos, ray_face=rf, ray_t=rt)
    print("wrote", sorted(p.name for p in HERE.iterdir()))


if __name__ == "__main__":
    main()
