"""Seeded host-side inputs of the differentiable workload: a closed surface
made from a perturbed Fibonacci sphere, its triangle soup, and orthographic
rays.  Plain numpy/scipy data preparation; nothing here runs on the GPU.

Workload: 131,074 vertices at Fibonacci-lattice points of the unit sphere,
triangulated by their convex hull (262,144 triangles, outward winding), then
radially perturbed by a sum of 16 random real spherical-harmonic modes
(l <= 8, coefficients uniform in [-0.05, 0.05]); 1024 x 1024 orthographic
rays along +z from z = -2 over [-1.1, 1.1]^2; the target depths come from the
unperturbed sphere traced with the same rays.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN_ANGLE = math.pi * (3.0 - math.sqrt(5.0))


def fibonacci_sphere(n: int) -> np.ndarray:
    """n points of the Fibonacci lattice on the unit sphere (float64[n, 3])."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * GOLDEN_ANGLE
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def hull_triangles(points: np.ndarray) -> np.ndarray:
    """Convex-hull triangulation with outward winding (int32[m, 3])."""
    from scipy.spatial import ConvexHull

    tri = ConvexHull(points).simplices.astype(np.int64)
    a, b, c = points[tri[:, 0]], points[tri[:, 1]], points[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0.0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return tri.astype(np.int32)


def real_sph_harm(l: int, m: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Orthonormal real spherical harmonic Y_lm (theta polar, phi azimuth)."""
    from scipy.special import lpmv

    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = lpmv(am, l, np.cos(theta))
    if m > 0:
        return math.sqrt(2.0) * norm * p * np.cos(am * phi)
    if m < 0:
        return math.sqrt(2.0) * norm * p * np.sin(am * phi)
    return norm * p


def sphere_mesh(seed: int = 4, n_vertices: int = 131074, modes: int = 16, lmax: int = 8,
                amp: float = 0.05) -> dict:
    """Perturbed sphere: unit points, their hull triangles, perturbed vertices."""
    unit = fibonacci_sphere(n_vertices)
    tri = hull_triangles(unit)
    rng = np.random.default_rng(seed)
    theta = np.arccos(np.clip(unit[:, 2], -1.0, 1.0))
    phi = np.arctan2(unit[:, 1], unit[:, 0])
    radius = np.ones(n_vertices)
    spec = []
    for _ in range(modes):
        l = int(rng.integers(1, lmax + 1))
        m = int(rng.integers(-l, l + 1))
        a = float(rng.uniform(-amp, amp))
        spec.append((l, m, a))
        radius += a * real_sph_harm(l, m, theta, phi)
    return {"unit": unit, "index": tri, "vertices": unit * radius[:, None], "modes": spec}


def soup(vertices: np.ndarray, index: np.ndarray) -> np.ndarray:
    """Triangle soup float4[3m] (w = 0) of an indexed mesh."""
    v = np.zeros((index.size, 4), np.float32)
    v[:, :3] = vertices[index.reshape(-1)].astype(np.float32)
    return v


def mesh_float4(vertices: np.ndarray) -> np.ndarray:
    v = np.zeros((vertices.shape[0], 4), np.float32)
    v[:, :3] = vertices.astype(np.float32)
    return v


def ortho_rays(res: int = 1024, extent: float = 1.1, z0: float = -2.0) -> np.ndarray:
    """res x res rays along +z, row-major, float32[res*res, 8]
    {ox, oy, oz, tmin, dx, dy, dz, tmax} with tmin 0 and tmax inf."""
    c = (np.arange(res, dtype=np.float64) + 0.5) / res * (2.0 * extent) - extent
    y, x = np.meshgrid(c, c, indexing="ij")
    r = np.zeros((res * res, 8), np.float32)
    r[:, 0] = x.reshape(-1)
    r[:, 1] = y.reshape(-1)
    r[:, 2] = z0
    r[:, 6] = 1.0
    r[:, 7] = np.inf
    return r
