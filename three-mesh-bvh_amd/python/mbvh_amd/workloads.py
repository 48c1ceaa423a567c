"""Seeded synthetic query sets for the benchmark configurations and the parity tests.

C1  raycastFirst: TorusKnot(10, 3, 400, 100), SAH, targetLeafSize 10; 65,536 rays with origins
    uniform on the sphere r=40 about the origin, each aimed at a uniform point of the mesh AABB.
C2  closestPointToPoint: TorusKnot(10, 3, 800, 200), SAH; 128^3 grid over the AABB expanded 10%.
C3  differentiable raycastFirst: noise-displaced Icosahedron(1, 200); 1M rays from r=3 toward the
    unit ball.
C4  differentiable closest point: noise-displaced Icosahedron(1, 223); 256^3 grid.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import geometry as G


def unit_sphere(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def rays_toward_box(n: int, radius: float, lo: np.ndarray, hi: np.ndarray, center=(0.0, 0.0, 0.0),
                    seed_origin: int = 1, seed_target: int = 2) -> Tuple[np.ndarray, np.ndarray]:
    """Origins uniform on a sphere, each aimed at a uniform point of the box [lo, hi]."""
    o = np.asarray(center) + radius * unit_sphere(np.random.default_rng(seed_origin), n)
    t = lo + (hi - lo) * np.random.default_rng(seed_target).random((n, 3))
    o32 = o.astype(np.float32)
    d = t - o32.astype(np.float64)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    return o32, d.astype(np.float32)


def rays_toward_ball(n: int, radius: float, ball: float, seed_origin: int = 1, seed_target: int = 2):
    rng_o = np.random.default_rng(seed_origin)
    rng_t = np.random.default_rng(seed_target)
    o = radius * unit_sphere(rng_o, n)
    t = unit_sphere(rng_t, n) * (ball * rng_t.random((n, 1)) ** (1.0 / 3.0))
    o32 = o.astype(np.float32)
    d = t - o32.astype(np.float64)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    return o32, d.astype(np.float32)


def grid_points(lo: np.ndarray, hi: np.ndarray, res: int, expand: float = 0.10) -> np.ndarray:
    """res^3 cell centers of the AABB expanded by `expand` about its center; x varies fastest."""
    c = (lo + hi) / 2
    ext = (hi - lo) * (1 + expand)
    mn = c - ext / 2
    step = ext / res
    k, j, i = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    pts = np.stack([mn[0] + (i + 0.5) * step[0], mn[1] + (j + 0.5) * step[1], mn[2] + (k + 0.5) * step[2]],
                   axis=-1)
    return pts.reshape(-1, 3).astype(np.float32)


@dataclass
class Workload:
    name: str
    geometry: G.Geometry
    origins: Optional[np.ndarray] = None
    directions: Optional[np.ndarray] = None
    points: Optional[np.ndarray] = None


def c1_raycast(n_rays: int = 65536) -> Workload:
    g = G.torus_knot(10, 3, 400, 100)
    lo, hi = g.bounds()
    o, d = rays_toward_box(n_rays, 40.0, lo, hi)
    return Workload("C1 raycastFirst TorusKnot(10,3,400,100) 80k tris, 65,536 rays", g, o, d)


def c2_closest(res: int = 128) -> Workload:
    g = G.torus_knot(10, 3, 800, 200)
    lo, hi = g.bounds()
    return Workload(f"C2 closestPointToPoint TorusKnot(10,3,800,200) 320k tris, {res}^3 grid", g,
                    points=grid_points(lo, hi, res))


def c3_raycast(n_rays: int = 1 << 20, detail: int = 200) -> Workload:
    g = G.displaced_icosphere(detail)
    o, d = rays_toward_ball(n_rays, 3.0, 1.0)
    return Workload(f"C3 raycastFirst displaced Icosahedron(1,{detail}), {n_rays} rays", g, o, d)


def c4_closest(res: int = 256, detail: int = 223) -> Workload:
    g = G.displaced_icosphere(detail)
    lo, hi = g.bounds()
    return Workload(f"C4 closest point displaced Icosahedron(1,{detail}), {res}^3 grid", g,
                    points=grid_points(lo, hi, res))
