"""The oracle's closestPointToGeometry pieces, pinned without the oracle's own formulas:
- closed forms: parallel offset triangles, a vertex over a face, skew perpendicular edges, crossing
  triangles (distance 0 and the middle of the known crossing segment), coplanar triangles apart,
  touching at a vertex;
- the segment-segment routine against a dense parameter grid;
- random triangle pairs against dense sampling of both triangles: the oracle's distance lies between
  the sampled minimum less the sampling resolution and the sampled minimum, its points lie on their
  triangles and are that distance apart, and swapping the triangles gives the same distance;
- the tree walk against the brute-force pairing on a mesh (same pair, bit-identical distance)."""
import numpy as np
import pytest

import oracle_py as O
from helpers import build_mesh
from mbvh_amd import geometry as G


def _tri(*pts):
    return np.array(pts, np.float64)


@pytest.mark.parametrize("h", [0.5, 1e-3, 7.0])
def test_parallel_offset(h):
    a = _tri([0, 0, 0], [2, 0, 0], [0, 2, 0])
    b = a + np.array([0.2, 0.1, h])
    d, p1, p2 = O.triangle_distance(a, b)
    assert d == pytest.approx(h, rel=1e-14)
    assert p1[2] == 0 and p2[2] == pytest.approx(h)


def test_vertex_over_face():
    a = _tri([-1, -1, 0], [2, -1, 0], [-1, 2, 0])
    b = _tri([0.1, 0.2, 0.75], [0.5, 0.5, 3], [-0.3, 0.9, 2])
    d, p1, p2 = O.triangle_distance(a, b)
    assert d == pytest.approx(0.75, rel=1e-14)
    np.testing.assert_allclose(p1, [0.1, 0.2, 0], atol=1e-15)
    np.testing.assert_allclose(p2, [0.1, 0.2, 0.75], atol=1e-15)


def test_skew_edges():
    # edge x-axis at z=0 vs edge y-axis at z=h, each triangle turned away from the other
    h = 0.4
    a = _tri([-1, 0, 0], [1, 0, 0], [0, 0, -1])
    b = _tri([0, -1, h], [0, 1, h], [0, 0, h + 1])
    d, p1, p2 = O.triangle_distance(a, b)
    assert d == pytest.approx(h, rel=1e-14)
    np.testing.assert_allclose(p1, [0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(p2, [0, 0, h], atol=1e-15)
    s, t = O.segment_segment([-1, 0, 0], [1, 0, 0], [0, -1, h], [0, 1, h])
    np.testing.assert_allclose(s, [0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(t, [0, 0, h], atol=1e-15)


def test_crossing_triangles_give_segment_middle():
    # a in z=0 covering x in [-2, 4]; b vertical in y=0 crossing z=0 along x in [0, 1]
    a = _tri([-2, -2, 0], [4, -2, 0], [-2, 4, 0])
    b = _tri([0, 0, -1], [2, 0, 1], [0, 0, 1])
    d, p1, p2 = O.triangle_distance(a, b)
    assert d == 0
    # b meets z=0 on x in [0, 1] (its edges from (0,0,-1) to (2,0,1) and to (0,0,1)); a covers it all
    np.testing.assert_allclose(p1, [0.5, 0, 0], atol=1e-15)
    np.testing.assert_array_equal(p1, p2)


def test_coplanar_apart_and_touching():
    a = _tri([0, 0, 0], [1, 0, 0], [0, 1, 0])
    b = _tri([3, 0, 0], [4, 0, 0], [3, 1, 0])
    d, _, _ = O.triangle_distance(a, b)
    assert d == pytest.approx(2.0, rel=1e-15)
    c = _tri([1, 0, 0], [2, 0, 5], [2, 1, 5])
    d, p1, p2 = O.triangle_distance(a, c)
    assert d == 0
    np.testing.assert_allclose(p1, [1, 0, 0], atol=1e-15)


def test_segment_segment_grid():
    rng = np.random.default_rng(1)
    t = np.linspace(0, 1, 2001)
    for _ in range(60):
        s1, e1, s2, e2 = rng.normal(size=(4, 3))
        if rng.random() < 0.3:
            e2 = s2 + (e1 - s1) * rng.uniform(0.5, 2)  # parallel segments
        o1, o2 = O.segment_segment(s1, e1, s2, e2)
        got = np.linalg.norm(o1 - o2)
        P = s1 + t[:, None] * (e1 - s1)
        # the exact distance from each grid point of the first segment to the second segment
        d = e2 - s2
        u = np.clip(((P - s2) @ d) / (d @ d), 0, 1)
        ref = np.linalg.norm(P - (s2 + u[:, None] * d), axis=1).min()
        step = np.linalg.norm(e1 - s1) / 2000
        assert ref - step - 1e-12 <= got <= ref + 1e-12


def _on_triangle(p, t, tol):
    a, b, c = t
    n = np.cross(b - a, c - a)
    nn = np.linalg.norm(n)
    if abs((p - a) @ n) / nn > tol:
        return False
    # barycentric coordinates
    v0, v1, v2 = b - a, c - a, p - a
    d00, d01, d11, d20, d21 = v0 @ v0, v0 @ v1, v1 @ v1, v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    return min(v, w, 1 - v - w) > -tol


def _samples(t, n):
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = i + j <= n
    u, v = i[keep] / n, j[keep] / n
    return t[0] + u[:, None] * (t[1] - t[0]) + v[:, None] * (t[2] - t[0])


def test_random_pairs_against_sampling():
    rng = np.random.default_rng(7)
    n = 48
    zero = 0
    for case in range(100):
        a = rng.normal(size=(3, 3))
        b = rng.normal(size=(3, 3)) * rng.uniform(0.3, 1.5) + rng.normal(size=3) * rng.uniform(0, 2.5)
        d, p1, p2 = O.triangle_distance(a, b)
        d_swap, _, _ = O.triangle_distance(b, a)
        assert d == pytest.approx(d_swap, rel=1e-12, abs=1e-14)
        assert _on_triangle(p1, a, 1e-9) and _on_triangle(p2, b, 1e-9), case
        assert np.linalg.norm(p1 - p2) == pytest.approx(d, rel=1e-12, abs=1e-14)
        sa, sb = _samples(a, n), _samples(b, n)
        diff = sa[:, None, :] - sb[None, :, :]
        u = np.sqrt((diff * diff).sum(-1)).min()
        res = (max(np.linalg.norm(a[1] - a[0]), np.linalg.norm(a[2] - a[0])) +
               max(np.linalg.norm(b[1] - b[0]), np.linalg.norm(b[2] - b[0]))) / n
        assert u - res - 1e-12 <= d <= u + 1e-12, case
        zero += d == 0
    assert 5 < zero < 95  # both intersecting and separated pairs were drawn


def test_tree_walk_equals_brute_force():
    g, idx, t = build_mesh(G.torus_knot(1.0, 0.3, 48, 8))
    om = O.OracleMesh(g.positions, idx, t.roots)
    other = G.sphere(0.5, 10, 6)
    for k, off in enumerate(([0.0, 0.0, 2.2], [1.1, 0.2, 0.0], [0.0, 0.0, 0.0])):
        m = np.eye(4)
        m[:3, :3] *= 0.8
        m[:3, 3] = off
        e = m.T.reshape(-1)  # column-major
        walk = om.closest_to_geometry(other.positions, other.index, e)
        brute = om.closest_to_geometry(other.positions, other.index, e, brute=True)
        assert walk["distance"] == brute["distance"], k
        assert walk["face_index"] == brute["face_index"] and walk["other_face_index"] == brute["other_face_index"]
    # a threshold under the distance: nothing
    far = np.eye(4)
    far[:3, 3] = [0, 0, 5]
    r = om.closest_to_geometry(other.positions, other.index, far.T.reshape(-1))
    assert r is not None and r["distance"] > 2
    assert om.closest_to_geometry(other.positions, other.index, far.T.reshape(-1), max_threshold=r["distance"]) is None


def test_scene_brute_force_against_single_mesh_walk():
    """The oracle's scene closest point (brute force over world triangles) against its closest-point
    tree walk: one identity instance gives the same faces and distances; a rotated, uniformly scaled
    and translated copy scales the distances by the scale and maps the points; a hidden instance never
    answers."""
    g, idx, t = build_mesh(G.torus_knot(1.0, 0.3, 40, 8))
    om = O.OracleMesh(g.positions, idx, t.roots)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2, 2, (60, 3)).astype(np.float32)
    walk = om.closest_point(pts, threads=1)
    ident = np.eye(4, dtype=np.float32).reshape(1, 16)
    r = O.scene_closest_point([om], ident, np.zeros(1, np.uint32), None, pts)
    np.testing.assert_allclose(r["distance"], walk["distance"], rtol=1e-12)
    # faces: the walk keeps the reference's first visited among exact ties, the brute force the lowest
    diff = np.nonzero(r["face"] != walk["face"])[0]
    assert diff.size < 20
    for i in diff:
        assert r["face"][i] < walk["face"][i]
        assert om.point_face_distance(int(walk["face"][i]), pts[i]) == pytest.approx(r["distance"][i], rel=1e-12)
    c, s_ = np.cos(0.7), np.sin(0.7)
    m = np.eye(4)
    m[:3, :3] = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]) * 2.0
    m[:3, 3] = [5.0, -3.0, 1.0]
    world = np.stack([m.T.reshape(-1), np.eye(4).reshape(-1)]).astype(np.float32)
    wp = (pts.astype(np.float64) @ m[:3, :3].T + m[:3, 3]).astype(np.float32)
    vis = np.array([1, 0], np.uint8)
    r2 = O.scene_closest_point([om], world, np.zeros(2, np.uint32), vis, wp)
    assert np.all(r2["instance"] == 0)
    lp = (np.linalg.inv(m) @ np.c_[wp.astype(np.float64), np.ones(60)].T).T[:, :3].astype(np.float32)
    walk2 = om.closest_point(lp, threads=1)
    # the world and local query points are each rounded to float: 1e-6 of the coordinates (~10) absolute
    np.testing.assert_allclose(r2["distance"], 2.0 * walk2["distance"], rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(r2["point"], walk2["point"] @ m[:3, :3].T + m[:3, 3], atol=1e-5)
