"""Pins the oracle's per-triangle arithmetic (three's Ray.intersectTriangle, Triangle.closestPointToPoint,
Triangle.getBarycoord restated in FP64) with closed forms and with independent formulations
(a 3x3 linear solve; closest points on segments and the plane) — not by re-typing its formulas."""
import numpy as np
import pytest

import oracle_py as O

A = np.array([0.0, 0.0, 0.0])
B = np.array([1.0, 0.0, 0.0])
C = np.array([0.0, 1.0, 0.0])  # (B-A) x (C-A) = +z: front-facing for rays travelling toward -z


def test_ray_triangle_closed_forms():
    r = O.ray_triangle([0.25, 0.25, 1], [0, 0, -1], A, B, C, True)
    assert r is not None and r[0] == 1.0 and np.array_equal(r[1], [0.25, 0.25, 0.0])
    # back face: culled with backface culling, hit without
    assert O.ray_triangle([0.25, 0.25, -1], [0, 0, 1], A, B, C, True) is None
    r = O.ray_triangle([0.25, 0.25, -1], [0, 0, 1], A, B, C, False)
    assert r is not None and r[0] == 1.0
    # edges and vertices are inclusive
    assert O.ray_triangle([0.5, 0.0, 1], [0, 0, -1], A, B, C, True) is not None
    assert O.ray_triangle([0.0, 0.0, 1], [0, 0, -1], A, B, C, True) is not None
    assert O.ray_triangle([0.5, 0.5, 1], [0, 0, -1], A, B, C, True) is not None
    assert O.ray_triangle([0.5, -1e-9, 1], [0, 0, -1], A, B, C, True) is None
    assert O.ray_triangle([0.5, 0.5 + 1e-9, 1], [0, 0, -1], A, B, C, True) is None
    # parallel ray: DdN == 0
    assert O.ray_triangle([0, 0, 1], [1, 0, 0], A, B, C, False) is None
    # behind the origin: t < 0
    assert O.ray_triangle([0.25, 0.25, 1], [0, 0, 1], A, B, C, False) is None
    # origin on the triangle: t == 0 is a hit
    r = O.ray_triangle([0.25, 0.25, 0], [0, 0, -1], A, B, C, True)
    assert r is not None and r[0] == 0.0
    # unnormalised direction: t is in direction units
    r = O.ray_triangle([0.25, 0.25, 4], [0, 0, -2], A, B, C, True)
    assert r[0] == 2.0


def test_ray_triangle_matches_linear_solve():
    rng = np.random.default_rng(0)
    hits = 0
    for _ in range(2000):
        a, b, c = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 3
        uu, vv = rng.uniform(-0.3, 1.3, size=2)
        d = (a + uu * (b - a) + vv * (c - a) - o) * rng.choice([-1.0, 1.0], p=[0.2, 0.8])
        # independent: solve o + t d = a + u (b - a) + v (c - a)
        M = np.stack([b - a, c - a, -d], axis=1)
        u, v, t = np.linalg.solve(M, o - a)
        inside = u >= 0 and v >= 0 and u + v <= 1 and t >= 0
        r = O.ray_triangle(o, d, a, b, c, False)
        if abs(min(u, v, 1 - u - v, t)) < 1e-9:
            continue  # too close to a boundary for the comparison to be meaningful
        assert (r is not None) == inside
        if r is not None:
            hits += 1
            assert abs(r[0] - t) <= 1e-9 * max(1, abs(t))
            assert np.allclose(r[1], a + u * (b - a) + v * (c - a), atol=1e-9)
            bc = O.barycoord(r[1], a, b, c)
            assert np.allclose(bc, [1 - u - v, u, v], atol=1e-8)
        # backface culling: only rays against the normal survive
        n = np.cross(b - a, c - a)
        rc = O.ray_triangle(o, d, a, b, c, True)
        assert (rc is not None) == (inside and np.dot(d, n) < 0)
    assert hits > 300


def _closest_independent(p, a, b, c):
    """Closest point: plane projection if it falls inside, else the best of the three segments."""
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    M = np.stack([b - a, c - a], axis=1)
    uv, *_ = np.linalg.lstsq(M, q - a, rcond=None)
    if uv[0] >= 0 and uv[1] >= 0 and uv.sum() <= 1:
        return q
    best = None
    for s, e in ((a, b), (b, c), (a, c)):
        t = np.clip(np.dot(p - s, e - s) / np.dot(e - s, e - s), 0, 1)
        x = s + t * (e - s)
        if best is None or np.linalg.norm(p - x) < np.linalg.norm(p - best):
            best = x
    return best


def test_closest_point_regions():
    cases = {
        (0.2, 0.2, 5): (0.2, 0.2, 0),      # face
        (-1, -1, 0): (0, 0, 0),            # vertex a
        (2, -1, 0): (1, 0, 0),             # vertex b
        (-1, 2, 0): (0, 1, 0),             # vertex c
        (0.5, -1, 3): (0.5, 0, 0),         # edge ab
        (1, 1, 0): (0.5, 0.5, 0),          # edge bc
        (-1, 0.5, 0): (0, 0.5, 0),         # edge ac
    }
    for p, want in cases.items():
        assert np.allclose(O.closest_point_triangle(p, A, B, C), want, atol=1e-15), p


def test_closest_point_matches_independent():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        got = O.closest_point_triangle(p, a, b, c)
        want = _closest_independent(p, a, b, c)
        assert abs(np.linalg.norm(p - got) - np.linalg.norm(p - want)) < 1e-9


def test_barycoord():
    assert np.allclose(O.barycoord([1 / 3, 1 / 3, 0], A, B, C), [1 / 3, 1 / 3, 1 / 3])
    assert np.allclose(O.barycoord([1, 0, 0], A, B, C), [0, 1, 0])
    assert np.allclose(O.barycoord([0, 1, 0], A, B, C), [0, 0, 1])
    # degenerate triangle: three.js returns null and zeroes the target
    assert O.barycoord([0, 0, 0], A, A, B) is None
