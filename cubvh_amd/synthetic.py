"""Seeded synthetic meshes and queries for the five benchmark workloads.

Every generator is deterministic (numpy default_rng with a fixed seed) and returns
fp32 vertices / int32 faces / fp32 queries on the host.

  icosphere(L)          icosahedron + L rounds of 4-way midpoint subdivision, projected
                        to the unit sphere: 10*4^L + 2 vertices, 20*4^L faces
  perturbed_sphere      r(theta, phi) = 0.5 + 0.05 * sum_{k=1..8} sin(a_k theta + b_k phi + psi_k),
                        a_k, b_k integers in 1..4, psi_k ~ U[0, 2 pi)  (theta = polar angle)
  soup                  perturbed level-8 sphere with 20% of faces removed + 200,000
                        random triangles (centroid ~ U[-0.6, 0.6]^3, equilateral with edge
                        ~ N(0.01, 0.003) clipped at 1e-4, uniformly random orientation)
  grid_points(n)        n^3 points, meshgrid(linspace(-1, 1, n) x 3, indexing='ij')
  camera_rays           pinhole 1920x1080, vertical fov 50 deg, 16 poses on a radius-2
                        sphere looking at the origin
"""
from __future__ import annotations

import math

import numpy as np

_ICO_CACHE: dict = {}


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere_unit(level: int):
    """Unit-sphere icosphere (float64 vertices, int64 faces), cached per level."""
    if level in _ICO_CACHE:
        return _ICO_CACHE[level]
    v, f = _icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // nv, uniq % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + nv  # midpoints of edges (0,1), (1,2), (2,0)
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1),
        ], axis=0)
        v = np.concatenate([v, mid], axis=0)
    _ICO_CACHE[level] = (v, f)
    return v, f


def icosphere(level: int, radius: float = 0.5):
    v, f = icosphere_unit(level)
    return (v * radius).astype(np.float32), f.astype(np.int32)


def perturbed_sphere(level: int, seed: int = 0, scale: float = 1.0):
    """Sinusoid-perturbed icosphere (radius 0.5 * scale before perturbation)."""
    rng = np.random.default_rng(seed)
    a = rng.integers(1, 5, size=8)
    b = rng.integers(1, 5, size=8)
    psi = rng.uniform(0.0, 2.0 * np.pi, size=8)
    v, f = icosphere_unit(level)
    theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    r = 0.5 + 0.05 * np.sin(a[None, :] * theta[:, None] + b[None, :] * phi[:, None] + psi[None, :]).sum(1)
    verts = v * (r * scale)[:, None]
    meta = {"a": a.tolist(), "b": b.tolist(), "psi": psi.tolist(), "scale": scale}
    return verts.astype(np.float32), f.astype(np.int32), meta


def jittered_sphere(level: int = 6, sigma: float = 0.005, seed: int = 0):
    rng = np.random.default_rng(seed)
    v, f = icosphere(level, 0.5)
    v = v.astype(np.float64) + rng.normal(0.0, sigma, size=v.shape)
    return v.astype(np.float32), f


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def soup(level: int = 8, seed: int = 0, n_random: int = 200_000, keep: float = 0.8):
    """Non-watertight soup: perturbed sphere minus (1-keep) of its faces plus random triangles."""
    v, f, _ = perturbed_sphere(level, seed)
    rng = np.random.default_rng(seed + 1)
    m = f.shape[0]
    kept = np.sort(rng.permutation(m)[: int(round(keep * m))])
    f = f[kept]
    c = rng.uniform(-0.6, 0.6, size=(n_random, 3))
    s = np.clip(rng.normal(0.01, 0.003, size=n_random), 1e-4, None)
    tmpl = np.array([[1.0, 0.0, 0.0], [-0.5, math.sqrt(3) / 2, 0.0], [-0.5, -math.sqrt(3) / 2, 0.0]]) / math.sqrt(3)
    R = _random_rotations(rng, n_random)
    tv = c[:, None, :] + s[:, None, None] * np.einsum("nij,kj->nki", R, tmpl)
    nv = v.shape[0]
    verts = np.concatenate([v.astype(np.float64), tv.reshape(-1, 3)], axis=0).astype(np.float32)
    faces = np.concatenate([f, nv + np.arange(3 * n_random, dtype=np.int64).reshape(-1, 3)], axis=0).astype(np.int32)
    return verts, faces


def uniform_points(n: int, seed: int = 0, lo: float = -1.0, hi: float = 1.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)


def normal_points(n: int, sigma: float = 0.4, seed: int = 0):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, sigma, size=(n, 3)).astype(np.float32)


def grid_points(res: int = 256):
    x = np.linspace(-1.0, 1.0, res, dtype=np.float32)
    g = np.stack(np.meshgrid(x, x, x, indexing="ij"), axis=-1)
    return np.ascontiguousarray(g.reshape(-1, 3))


def sphere_rays(n: int, seed: int = 0):
    """Origins uniform on a radius-2 sphere aimed at uniform points in [-0.5, 0.5]^3."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = 2.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    t = rng.uniform(-0.5, 0.5, size=(n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def camera_poses(n_poses: int = 16, radius: float = 2.0, seed: int = 0):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n_poses, 3))
    return radius * c / np.linalg.norm(c, axis=1, keepdims=True)


def camera_rays(width: int = 1920, height: int = 1080, fov_deg: float = 50.0, n_poses: int = 16, seed: int = 0):
    """Pinhole rays for n_poses cameras looking at the origin; [n_poses*H*W, 3] each."""
    centers = camera_poses(n_poses, 2.0, seed)
    tan = math.tan(math.radians(fov_deg) / 2.0)
    aspect = width / height
    xs = ((np.arange(width) + 0.5) / width * 2.0 - 1.0) * tan * aspect
    ys = (1.0 - (np.arange(height) + 0.5) / height * 2.0) * tan
    px, py = np.meshgrid(xs, ys, indexing="xy")
    o_all, d_all = [], []
    for c in centers:
        fwd = -c / np.linalg.norm(c)
        up = np.array([0.0, 0.0, 1.0]) if abs(fwd[2]) < 0.99 else np.array([0.0, 1.0, 0.0])
        right = np.cross(fwd, up)
        right /= np.linalg.norm(right)
        up = np.cross(right, fwd)
        d = fwd[None, None, :] + px[..., None] * right + py[..., None] * up
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        o_all.append(np.broadcast_to(c.astype(np.float32), (height * width, 3)))
        d_all.append(d.reshape(-1, 3).astype(np.float32))
    return np.concatenate(o_all), np.concatenate(d_all)


def forest(n_meshes: int = 512, seed: int = 0, min_level: int = 3, max_level: int = 7,
           min_q: int = 4096, max_q: int = 262144):
    """Mesh/query specifications of the batched workload: per mesh a level uniform in
    {min_level..max_level} (face counts 20*4^L, log-uniform over 1,280..327,680), a
    scale U[0.3, 0.9], a perturbation seed and a log-uniform query count."""
    rng = np.random.default_rng(seed)
    levels = rng.integers(min_level, max_level + 1, size=n_meshes)
    scales = rng.uniform(0.3, 0.9, size=n_meshes)
    nq = np.exp(rng.uniform(math.log(min_q), math.log(max_q), size=n_meshes)).round().astype(np.int64)
    seeds = rng.integers(0, 2**31 - 1, size=n_meshes)
    return [
        {"level": int(levels[i]), "scale": float(scales[i]), "seed": int(seeds[i]), "n_queries": int(nq[i])}
        for i in range(n_meshes)
    ]
