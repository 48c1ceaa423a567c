"""The oracle's traversals (restatements of raycastFirst / raycast / closestPointToPoint over the
reference node buffers) against brute force, plus the reference test suite's known answers."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh, threshold_cases


def _mesh(geometry, **kw):
    g, idx, tree = build_mesh(geometry, **kw)
    return g, idx, tree, O.OracleMesh(g.positions, idx, tree.roots, indirect=tree.indirect)


@pytest.mark.parametrize("strategy", [M.CENTER, M.SAH])
def test_raycast_first_bvh_equals_brute_force(strategy):
    g, idx, tree, om = _mesh(G.torus_knot(10, 3, 48, 12), strategy=strategy)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(1500, 40.0, lo, hi, seed_origin=3, seed_target=4)
    for side in (O.FRONT, O.BACK, O.DOUBLE):
        a = om.raycast_first(o, d, side=side)
        b = om.raycast_first(o, d, side=side, brute=True)
        # the first hit is BVH-independent; only exact-distance ties may pick another face
        assert np.array_equal(a["face"] == O.NO_HIT, b["face"] == O.NO_HIT)
        assert np.allclose(a["distance"], b["distance"], rtol=0, atol=0, equal_nan=True)
        hit = a["face"] != O.NO_HIT
        assert np.mean(a["face"][hit] == b["face"][hit]) > 0.999
        # traversal counters: the BVH reads far fewer triangles than brute force
        assert a["tris"].mean() < b["tris"].mean() / 10


def test_double_inputs_match_float_inputs_and_brute_force():
    """The double-precision entry points: on float-representable rays and points they give the float
    entry points' results bit for bit; on genuine double inputs (a float ray plus offsets far below its
    float rounding, and rays from a rigid frame change in double) the traversal equals brute force,
    and the double results differ from those of the float-rounded inputs (so the inputs' low bits
    reach the arithmetic)."""
    g, idx, tree, om = _mesh(G.torus_knot(10, 3, 48, 12))
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(1500, 40.0, lo, hi, seed_origin=5, seed_target=6)
    a32 = om.raycast_first(o, d)
    a64 = om.raycast_first(o.astype(np.float64), d.astype(np.float64))
    for k in ("face", "distance", "point", "barycoord"):
        assert np.array_equal(a32[k], a64[k], equal_nan=True), k
    pts = W.grid_points(lo, hi, 10)
    c32 = om.closest_point(pts)
    c64 = om.closest_point(pts.astype(np.float64))
    for k in ("face", "distance", "point"):
        assert np.array_equal(c32[k], c64[k], equal_nan=True), k
    rng = np.random.default_rng(7)
    od = o.astype(np.float64) + rng.standard_normal(o.shape) * 1e-9
    dd = d.astype(np.float64) + rng.standard_normal(d.shape) * 1e-10
    for side in (O.FRONT, O.DOUBLE):
        a = om.raycast_first(od, dd, side=side)
        b = om.raycast_first(od, dd, side=side, brute=True)
        assert np.array_equal(a["face"] == O.NO_HIT, b["face"] == O.NO_HIT)
        assert np.array_equal(a["distance"], b["distance"])
    a = om.raycast_first(od, dd)
    hit = a["face"] != O.NO_HIT
    assert np.all(a["distance"][hit] != a32["distance"][hit])  # the low bits moved every hit
    pd = pts.astype(np.float64) + rng.standard_normal(pts.shape) * 1e-9
    a = om.closest_point(pd)
    b = om.closest_point(pd, brute=True)
    assert np.array_equal(a["distance"], b["distance"])
    for i in np.nonzero(a["face"] != b["face"])[0]:
        assert om.point_face_distance(a["face"][i], pd[i]) == b["distance"][i]


def test_closest_point_bvh_equals_brute_force():
    g, idx, tree, om = _mesh(G.torus_knot(10, 3, 48, 12))
    lo, hi = g.bounds()
    pts = W.grid_points(lo, hi, 10)
    a = om.closest_point(pts)
    b = om.closest_point(pts, brute=True)
    assert np.array_equal(a["distance"], b["distance"])
    # where faces differ the distances are identical (shared vertex / edge ties)
    diff = a["face"] != b["face"]
    for i in np.nonzero(diff)[0]:
        assert om.point_face_distance(a["face"][i], pts[i]) == b["distance"][i]


def test_sphere_distance_kats():
    """ShapeCasts.test.js:715-752: SphereGeometry(1, 200, 200)."""
    g, idx, tree, om = _mesh(G.sphere(1, 200, 200), strategy=M.CENTER)
    r = om.closest_point(np.array([[0, 0, 0], [0, 1, 0]], np.float32))
    assert 0.999 <= r["distance"][0] <= 1
    assert r["distance"][1] == 0
    rng = np.random.default_rng(2)
    v = rng.random((100, 3)) - 0.5
    length = rng.random(100) * 3
    v = (v / np.linalg.norm(v, axis=1, keepdims=True) * length[:, None]).astype(np.float32)
    r = om.closest_point(v)
    assert np.all(np.abs(r["distance"] - np.abs(1 - np.linalg.norm(v.astype(np.float64), axis=1))) <= 1e-3)


def test_torus_hit_count_kat():
    """ShapeCasts.test.js:906-929: TorusGeometry(5, 5, 40, 10), ray (0,0,-10) -> +z, all hits: 10."""
    g, idx, tree, om = _mesh(G.torus(5, 5, 40, 10), strategy=M.CENTER)
    faces, dist = om.raycast_all([0, 0, -10], [0, 0, 1])
    assert len(faces) == 10
    first = om.raycast_first(np.array([[0, 0, -10]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert first["distance"][0] == dist.min()


def _box_face_mask(g, sides_by_material):
    mask = np.zeros(g.tri_count, np.uint8)
    for (start, count, mat) in g.groups:
        mask[start // 3:(start + count) // 3] = {0: 1, 1: 2, 2: 3}[sides_by_material[mat]]
    return mask


def test_box_group_sides_kat():
    """ShapeCasts.test.js:931-979: per-group Front/Back materials on a BoxGeometry."""
    g, idx, tree, _ = _mesh(G.box(), strategy=M.CENTER)
    F, B = O.FRONT, O.BACK
    om = O.OracleMesh(g.positions, idx, tree.roots, face_mask=_box_face_mask(g, [B, F, B, F, B, F]))
    faces, dist = om.raycast_all([0, 0.25, -10], [0, 0, 1], side=O.PER_FACE)
    assert len(faces) == 2
    om = O.OracleMesh(g.positions, idx, tree.roots, face_mask=_box_face_mask(g, [F, F, F, F, B, B]))
    r = om.raycast_first(np.array([[0, 0.25, -10]], np.float32), np.array([[0, 0, 1]], np.float32), side=O.PER_FACE)
    assert r["point"][0, 2] == 0.5


def test_group_gap_kat():
    """ShapeCasts.test.js:981-1005: triangles in the gap between groups are not in any root."""
    g = G.box()
    # groups (0, 6) and (12, 24) in index units -> triangle ranges [0, 2) and [4, 12)
    g2, idx, tree = build_mesh(g, ranges=[(0, 2), (4, 8)])
    om = O.OracleMesh(g2.positions, idx, tree.roots)
    faces, _ = om.raycast_all([-2, 0, 0], [1, 0, 0])
    assert len(faces) == 0


def test_max_depth_matches_brute_force():
    """MeshBVH.options.test.js:122-140: a depth-3 tree still finds the same hits."""
    g, idx, tree, om = _mesh(G.torus_knot(10, 3, 48, 12), max_depth=3)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(500, 40.0, lo, hi, seed_origin=5, seed_target=6)
    a = om.raycast_first(o, d)
    b = om.raycast_first(o, d, brute=True)
    assert np.allclose(a["distance"], b["distance"], rtol=0, atol=0, equal_nan=True)


def test_closest_point_threshold_branches_by_hand():
    """minThreshold's early exit and maxThreshold's pruning (closestPointToPoint.js:19-61 over
    shapecast.js:41-214) on a hand-built two-leaf mesh whose answers are worked out by hand: the
    first face under minThreshold in the visit order (not the closest), the entry test
    `score < best && score < maxThreshold^2`, a face beyond maxThreshold kept because its leaf's box
    passed, and no result when every box is pruned."""
    geo, cases = threshold_cases()
    g, idx, tree = build_mesh(geo, strategy=M.CENTER)
    root = np.asarray(tree.roots[0]).view(np.uint8)
    assert root.size // 32 == 3  # root + two leaves of six
    assert list(root.view(np.uint16)[14::16][1:]) == [6, 6]
    om = O.OracleMesh(g.positions, idx, tree.roots)
    for dt in (np.float32, np.float64):
        p = np.zeros((1, 3), dt)
        for mn, mx, want in cases:
            r = om.closest_point(p, mn, mx)
            if want is None:
                assert r["face"][0] == O.NO_HIT and np.isinf(r["distance"][0]), (mn, mx)
            else:
                assert r["distance"][0] == want, (mn, mx, r["distance"][0], want)
                assert om.point_face_distance(r["face"][0], p[0]) == want
    assert om.closest_point(np.zeros((1, 3), np.float32), brute=True)["distance"][0] == np.float32(0.6)
