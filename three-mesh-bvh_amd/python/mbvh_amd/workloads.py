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


@dataclass
class BatchedWorkload:
    """C5: many geometries, one rigid instance each, rays split across instances."""
    geometries: list           # G.Geometry per mesh
    instance_geometry: np.ndarray  # uint32 (n_instances,)
    matrix_world: np.ndarray   # float32 (n_instances, 16), column-major (three's Matrix4.elements)
    matrix_world_inverse: np.ndarray
    origins: np.ndarray
    directions: np.ndarray
    ray_instance: np.ndarray   # uint32 (n_rays,)


def _rotation(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def c5_batched(n_geometries: int = 512, n_rays: int = 1 << 22, tri_min: int = 100, tri_max: int = 200_000,
               zipf: float = 1.1, seed: int = 5) -> BatchedWorkload:
    """C5 batched multi-BVH raycastFirst: torus knots and spheres with seeded parameters and triangle
    counts log-uniform in [tri_min, tri_max], one random rigid instance each (spread over a cube),
    rays split across instances by Zipf(zipf) over a random ranking; each ray starts on a sphere of
    3x the instance's bounding radius and aims at a uniform point of its world AABB."""
    rng = np.random.default_rng(seed)
    geos = []
    for i in range(n_geometries):
        t = float(np.exp(rng.uniform(np.log(tri_min), np.log(tri_max))))
        if i % 2 == 0:
            radial = max(3, int(round(np.sqrt(t / 16))))
            tubular = max(3, int(round(t / (2 * radial))))
            p, q = int(rng.integers(1, 4)), int(rng.integers(2, 6))
            if np.gcd(p, q) != 1:
                p, q = 2, 3
            geos.append(G.torus_knot(rng.uniform(0.5, 2.0), rng.uniform(0.1, 0.5), tubular, radial, p, q))
        else:
            h = max(2, int(round(np.sqrt(t / 4))))
            w = max(3, int(round(t / (2 * h))))
            geos.append(G.sphere(rng.uniform(0.5, 2.0), w, h))
    n_inst = n_geometries
    inst_geom = rng.permutation(n_geometries).astype(np.uint32)
    world = np.zeros((n_inst, 16), np.float64)
    span = 8.0 * np.cbrt(n_inst)
    for k in range(n_inst):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        m = np.eye(4)
        m[:3, :3] = _rotation(q)
        m[:3, 3] = rng.uniform(-span, span, 3)
        world[k] = m.T.reshape(-1)  # column-major
    world32 = world.astype(np.float32)
    inv32 = np.stack([np.linalg.inv(world32[k].astype(np.float64).reshape(4, 4).T).T.reshape(-1)
                      for k in range(n_inst)]).astype(np.float32)
    weights = 1.0 / np.arange(1, n_inst + 1) ** zipf
    weights /= weights.sum()
    counts = rng.multinomial(n_rays, weights)[rng.permutation(n_inst)]
    ray_inst = np.repeat(np.arange(n_inst, dtype=np.uint32), counts)
    rng.shuffle(ray_inst)
    o = np.empty((n_rays, 3))
    tgt = np.empty((n_rays, 3))
    by_inst = np.argsort(ray_inst, kind="stable")
    starts = np.concatenate([[0], np.cumsum(np.bincount(ray_inst, minlength=n_inst))])
    for k in range(n_inst):
        sel = by_inst[starts[k]:starts[k + 1]]
        m = sel.shape[0]
        if m == 0:
            continue
        g = geos[inst_geom[k]]
        lo, hi = g.bounds()
        corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        mw = world32[k].astype(np.float64).reshape(4, 4).T
        wc = corners @ mw[:3, :3].T + mw[:3, 3]
        wlo, whi = wc.min(axis=0), wc.max(axis=0)
        c = (wlo + whi) / 2
        r = np.linalg.norm(whi - wlo) / 2
        o[sel] = c + 3 * r * unit_sphere(rng, m)
        tgt[sel] = wlo + (whi - wlo) * rng.random((m, 3))
    o32 = o.astype(np.float32)
    d = tgt - o32.astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return BatchedWorkload(geos, inst_geom, world32, inv32, o32, d.astype(np.float32), ray_inst)


@dataclass
class SceneWorkload:
    """Instances of meshes in one scene and world rays through it (closest hit over all instances)."""
    matrix_world: np.ndarray       # float32 (n_instances, 16), column-major
    matrix_world_inverse: np.ndarray
    instance_geometry: np.ndarray  # uint32 (n_instances,)
    origins: np.ndarray            # float32 (n, 3)
    directions: np.ndarray         # float32 (n, 3), unit


def scene_rays(geometries: list, n_instances: int, n_rays: int, seed: int = 11, spread: float = 6.0,
               non_uniform: float = 0.0) -> SceneWorkload:
    """Instances of `geometries` (random mesh, rotation, uniform scale in [0.5, 2] — or, for a fraction
    `non_uniform` of them, a different scale per axis — translation in a cube of half-width
    spread * cbrt(n_instances)); each ray starts on a sphere of 1.25x the scene's bounding radius and
    aims at a uniform point of a random instance's world box, so it crosses other instances' boxes."""
    rng = np.random.default_rng(seed)
    n_geo = len(geometries)
    geom = rng.integers(0, n_geo, n_instances).astype(np.uint32)
    span = spread * np.cbrt(n_instances)
    world = np.zeros((n_instances, 16), np.float64)
    for k in range(n_instances):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        s = np.full(3, rng.uniform(0.5, 2.0))
        if rng.random() < non_uniform:
            s = rng.uniform(0.5, 2.0, 3)
        m = np.eye(4)
        m[:3, :3] = _rotation(q) @ np.diag(s)
        m[:3, 3] = rng.uniform(-span, span, 3)
        world[k] = m.T.reshape(-1)
    world32 = world.astype(np.float32)
    inv32 = np.stack([np.linalg.inv(world32[k].astype(np.float64).reshape(4, 4).T).T.reshape(-1)
                      for k in range(n_instances)]).astype(np.float32)
    o, d = rays_through_instances(geometries, world32, geom, n_rays, rng)
    return SceneWorkload(world32, inv32, geom, o, d)


def rays_through_instances(geometries: list, matrix_world: np.ndarray, instance_geometry: np.ndarray, n_rays: int,
                           rng: np.random.Generator):
    """World rays from a sphere of 1.25x the instances' bounding radius towards uniform points of
    random instances' world boxes."""
    n_instances = matrix_world.shape[0]
    boxes = np.empty((n_instances, 2, 3))
    for k in range(n_instances):
        lo, hi = geometries[instance_geometry[k]].bounds()
        corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        mw = matrix_world[k].astype(np.float64).reshape(4, 4).T
        wc = corners @ mw[:3, :3].T + mw[:3, 3]
        boxes[k, 0], boxes[k, 1] = wc.min(axis=0), wc.max(axis=0)
    slo, shi = boxes[:, 0].min(axis=0), boxes[:, 1].max(axis=0)
    c = (slo + shi) / 2
    r = np.linalg.norm(shi - slo) / 2
    o = (c + 1.25 * r * unit_sphere(rng, n_rays)).astype(np.float32)
    pick = rng.integers(0, n_instances, n_rays)
    tgt = boxes[pick, 0] + (boxes[pick, 1] - boxes[pick, 0]) * rng.random((n_rays, 3))
    d = tgt - o.astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)
