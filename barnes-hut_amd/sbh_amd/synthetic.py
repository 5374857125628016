"""Seeded synthetic inputs for the benchmark workloads (host-side numpy).

* plummer: Plummer sphere, scale radius 1, truncated at r = 20 by inverse CDF.
* blobs: mixture of isotropic Gaussian blobs with log-normal masses.
* blob_sets: the batched workload, many blob sets of log-uniform sizes.
* geodesic_sphere / perturbed_sphere: closed triangle meshes (frequency-f
  geodesic subdivision of the icosahedron, 20 f^2 faces, optionally with the
  radius perturbed by real spherical harmonics) and their per-face dipoles
  (centroid, unit normal, area) for winding numbers.
"""
from __future__ import annotations

import math

import numpy as np


def plummer(n: int, seed: int, a: float = 1.0, r_max: float = 20.0) -> np.ndarray:
    """N points of a Plummer sphere truncated at r_max (exact inverse CDF:
    M(<r)/M = r^3 / (r^2 + a^2)^1.5)."""
    rng = np.random.default_rng(seed)
    x_max = r_max ** 3 / (r_max ** 2 + a * a) ** 1.5
    x = rng.uniform(0.0, x_max, n)
    r = a / np.sqrt(x ** (-2.0 / 3.0) - 1.0)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return (r[:, None] * v).astype(np.float32)


def blobs(n: int, k: int, seed: int, sigma: float = 0.02):
    """N points in k isotropic Gaussian blobs (centres uniform in [0,1)^3,
    equal weights) and log-normal(0, 0.5) masses normalised to sum 1."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (k, 3))
    which = rng.integers(0, k, n)
    p = centres[which] + rng.normal(0.0, sigma, (n, 3))
    m = rng.lognormal(0.0, 0.5, n)
    m /= m.sum()
    return p.astype(np.float32), m.astype(np.float32)


def blob_sets(num_sets: int, seed: int = 5000, lo: int = 1024, hi: int = 524288):
    """Sets of sizes round(exp(U(ln lo, ln hi))), each a mixture of U{8..64}
    blobs (sigma 0.02, log-normal masses normalised to sum 1 per set), with
    cotangents N(0, 1).  Returns (sizes, points, masses, cotangents), the last
    three as per-set lists."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), num_sets)))
    sizes = sizes.astype(np.int64)
    ps, ms, cs = [], [], []
    for k, n in enumerate(sizes):
        p, m = blobs(int(n), int(rng.integers(8, 65)), seed + 1 + k)
        ps.append(p)
        ms.append(m)
        cs.append(rng.normal(size=(int(n), 3)).astype(np.float32))
    return sizes, ps, ms, cs


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_sphere(freq: int):
    """Unit sphere mesh: every icosahedron face split into freq^2 triangles
    (vertices projected to the sphere, outward orientation).  Returns
    (vertices (V,3) float64, faces (F,3) int64) with F = 20 freq^2; vertices
    on shared icosahedron edges are merged."""
    v0, f0 = _icosahedron()
    verts, index = [], {}

    def vid(p):
        key = tuple(np.round(p, 12))
        k = index.get(key)
        if k is None:
            k = len(verts)
            index[key] = k
            verts.append(p)
        return k

    faces = []
    for a, b, c in f0:
        A, B, C = v0[a], v0[b], v0[c]
        ids = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (C - A) * (j / freq)
                ids[i, j] = vid(p / np.linalg.norm(p))
        for i in range(freq):
            for j in range(freq - i):
                faces.append((ids[i, j], ids[i + 1, j], ids[i, j + 1]))
                if i + j < freq - 1:
                    faces.append((ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]))
    return np.asarray(verts), np.asarray(faces, np.int64)


def real_sph_harm(l: int, m: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Orthonormal real spherical harmonic Y_lm (theta polar, phi azimuth)."""
    from scipy.special import lpmv

    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) /
                     math.factorial(l + am))
    p = lpmv(am, l, np.cos(theta))
    if m == 0:
        return norm * p
    if m > 0:
        return math.sqrt(2.0) * norm * p * np.cos(m * phi)
    return math.sqrt(2.0) * norm * p * np.sin(am * phi)


def perturbed_sphere(freq: int, seed: int, lmax: int = 6, amp: float = 0.2,
                     sigma: float = 0.3):
    """Geodesic sphere with radius 1 + amp * sum_{l=1..lmax, m} a_lm Y_lm,
    a_lm ~ N(0, (sigma / l)^2)."""
    v, f = geodesic_sphere(freq)
    rng = np.random.default_rng(seed)
    theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    r = np.ones(v.shape[0])
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            r += amp * rng.normal(0.0, sigma / l) * real_sph_harm(l, m, theta, phi)
    if np.any(r <= 0):
        raise ValueError("perturbation turned the surface inside out")
    return v * r[:, None], f


def face_dipoles(verts: np.ndarray, faces: np.ndarray):
    """Per-face (centroid, unit normal, area) as float32 arrays."""
    a, b, c = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    cr = np.cross(b - a, c - a)
    dbl = np.linalg.norm(cr, axis=1)
    normal = cr / np.maximum(dbl, 1e-300)[:, None]
    centroid = (a + b + c) / 3.0
    return (centroid.astype(np.float32), normal.astype(np.float32),
            (0.5 * dbl).astype(np.float32))
