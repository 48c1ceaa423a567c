"""FP64 forward maps and analytic gradients of raycastFirst / closestPoint for FIXED hits.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's check leg, never by the product).

The reference has no derivatives, so nothing here restates reference code. The forward maps are
the reference's per-triangle math with the discrete choice (which face) held fixed:
  - ray: three's Ray.intersectTriangle solved as o + t d = a + u (b - a) + v (c - a)
    (src/utils/ThreeRayIntersectUtilities.js:21-47 calls it; distance = t |d|, point = o + t d,
    barycoord = (1 - u - v, u, v) as Triangle.getBarycoord);
  - closest point: Ericson's Voronoi-region closest point on a triangle (three's
    Triangle.closestPointToPoint, used by src/core/cast/closestPointToPoint.js:6-78), distance |q - p|.
The gradients are computed by a route independent of the kernels' closed forms (batched 3x3 linear
solves for the adjoint instead of cofactors) and are pinned by central finite differences of these
FP64 forward maps (tests/test_oracle_grad.py).
"""
from __future__ import annotations

import numpy as np


def face_vertices(positions: np.ndarray, index: np.ndarray, faces: np.ndarray):
    """(a, b, c) float64 arrays (n, 3) of the given faces of an indexed geometry."""
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    tri = np.asarray(index, np.int64).reshape(-1, 3)[np.asarray(faces, np.int64)]
    return pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]


def _ray_system(d, a, b, c):
    # columns: -d, b - a, c - a
    return np.stack([-d, b - a, c - a], axis=2)


def ray_forward(o, d, a, b, c):
    """Distance, point and barycoord of the ray's hit with the plane of (a, b, c), FP64, (n, ...)."""
    o, d, a, b, c = (np.asarray(x, np.float64) for x in (o, d, a, b, c))
    x = np.linalg.solve(_ray_system(d, a, b, c), (o - a)[:, :, None])[:, :, 0]
    t, u, v = x[:, 0], x[:, 1], x[:, 2]
    dist = t * np.linalg.norm(d, axis=1)
    point = o + t[:, None] * d
    bary = np.stack([1.0 - u - v, u, v], axis=1)
    return dist, point, bary


def ray_grads(o, d, a, b, c, g_dist=None, g_point=None, g_bary=None):
    """Gradients of L w.r.t. o, d, a, b, c given dL/d(distance, point, barycoord), FP64."""
    o, d, a, b, c = (np.asarray(x, np.float64) for x in (o, d, a, b, c))
    n = o.shape[0]
    g_dist = np.zeros(n) if g_dist is None else np.asarray(g_dist, np.float64)
    g_point = np.zeros((n, 3)) if g_point is None else np.asarray(g_point, np.float64)
    g_bary = np.zeros((n, 3)) if g_bary is None else np.asarray(g_bary, np.float64)
    M = _ray_system(d, a, b, c)
    x = np.linalg.solve(M, (o - a)[:, :, None])[:, :, 0]
    t, u, v = x[:, 0], x[:, 1], x[:, 2]
    w = 1.0 - u - v
    dlen = np.linalg.norm(d, axis=1)
    # dL/dx for x = (t, u, v) through distance = t|d|, point = o + t d, bary = (1-u-v, u, v)
    gx = np.stack([g_dist * dlen + np.einsum("ij,ij->i", g_point, d),
                   g_bary[:, 1] - g_bary[:, 0], g_bary[:, 2] - g_bary[:, 0]], axis=1)
    lam = np.linalg.solve(np.transpose(M, (0, 2, 1)), gx[:, :, None])[:, :, 0]  # M^T lam = gx
    # x = M^-1 s:  dL/ds = lam, dL/dM = -lam x^T; s = o - a; M = [-d, b - a, c - a]
    dM = -lam[:, :, None] * x[:, None, :]
    g_o = lam + g_point
    g_d = -dM[:, :, 0] + t[:, None] * g_point + (g_dist * t / dlen)[:, None] * d
    g_b = dM[:, :, 1]
    g_c = dM[:, :, 2]
    g_a = -lam - g_b - g_c
    return {"origins": g_o, "directions": g_d, "a": g_a, "b": g_b, "c": g_c, "t": t, "u": u, "v": v, "w": w}


def closest_on_triangle(q, a, b, c):
    """Closest point on triangle (a, b, c) to q and its barycoord, FP64, vectorised; the Voronoi
    regions are tested in Ericson's order (first match wins)."""
    q, a, b, c = (np.asarray(x, np.float64) for x in (q, a, b, c))
    ab, ac, ap, bp, cp = b - a, c - a, q - a, q - b, q - c
    dt = lambda x, y: np.einsum("ij,ij->i", x, y)
    d1, d2, d3, d4, d5, d6 = dt(ab, ap), dt(ac, ap), dt(ab, bp), dt(ac, bp), dt(ab, cp), dt(ac, cp)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    n = q.shape[0]
    bary = np.zeros((n, 3))
    done = np.zeros(n, bool)

    def take(mask, w, u, v):
        m = mask & ~done
        bary[m, 0], bary[m, 1], bary[m, 2] = w[m], u[m], v[m]
        done[m] = True

    one, zero = np.ones(n), np.zeros(n)
    with np.errstate(divide="ignore", invalid="ignore"):
        take((d1 <= 0) & (d2 <= 0), one, zero, zero)
        take((d3 >= 0) & (d4 <= d3), zero, one, zero)
        s = d1 / (d1 - d3)
        take((vc <= 0) & (d1 >= 0) & (d3 <= 0), 1 - s, s, zero)
        take((d6 >= 0) & (d5 <= d6), zero, zero, one)
        s = d2 / (d2 - d6)
        take((vb <= 0) & (d2 >= 0) & (d6 <= 0), 1 - s, zero, s)
        s = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        take((va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0), zero, 1 - s, s)
        den = va + vb + vc
        u, v = vb / den, vc / den
        take(np.ones(n, bool), 1 - u - v, u, v)
    p = bary[:, :1] * a + bary[:, 1:2] * b + bary[:, 2:3] * c
    return p, bary


def closest_distance(q, a, b, c):
    p, _ = closest_on_triangle(q, a, b, c)
    return np.linalg.norm(np.asarray(q, np.float64) - p, axis=1)


def closest_grads(q, a, b, c, g_dist):
    """Gradients of L w.r.t. q, a, b, c given dL/d(distance), FP64 (zero where the distance is 0)."""
    q = np.asarray(q, np.float64)
    p, bary = closest_on_triangle(q, a, b, c)
    D = np.linalg.norm(q - p, axis=1)
    g = np.asarray(g_dist, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        nrm = np.where(D[:, None] > 0, (q - p) / D[:, None], 0.0)
    g_q = g[:, None] * nrm
    return {"points": g_q, "a": -bary[:, :1] * g_q, "b": -bary[:, 1:2] * g_q, "c": -bary[:, 2:3] * g_q,
            "point": p, "barycoord": bary, "distance": D}


def scatter_vertices(index: np.ndarray, faces: np.ndarray, ga, gb, gc, vertex_count: int, hit=None):
    """Sums per-corner gradients into per-vertex gradients (FP64), skipping rows where hit is False.
    Also returns, per vertex, the sum over its corners of |ga + gb + gc| (the size of the query's
    whole gradient, which the corner shares by barycentric weight): the scale against which an
    FP32 result's rounding (absolute in the weights) is judged."""
    tri = np.asarray(index, np.int64).reshape(-1, 3)[np.asarray(faces, np.int64)]
    out = np.zeros((vertex_count, 3))
    scale = np.zeros((vertex_count, 3))
    keep = np.ones(tri.shape[0], bool) if hit is None else np.asarray(hit, bool)
    whole = np.abs(np.asarray(ga, np.float64) + np.asarray(gb, np.float64) + np.asarray(gc, np.float64))[keep]
    for k, g in enumerate((ga, gb, gc)):
        v = tri[keep, k]
        g = np.asarray(g, np.float64)[keep]
        for j in range(3):
            out[:, j] += np.bincount(v, weights=g[:, j], minlength=vertex_count)
            scale[:, j] += np.bincount(v, weights=whole[:, j], minlength=vertex_count)
    return out, scale
