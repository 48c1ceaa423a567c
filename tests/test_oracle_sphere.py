"""Pins the oracle's intersectsSphere (orc_intersects_sphere, the reference's sphereIntersectTriangle
by brute force): closed forms on a box (a face exactly at the radius counts, inclusive; a hair less
does not; radius 0 on a vertex counts; a negative radius never does), and agreement with an
independent formulation - the brute-force closest-point distance compared with the radius - on
random spheres around a torus knot."""
import numpy as np

import mbvh_amd as M
import oracle_py as O
from mbvh_amd import geometry as G


def _mesh(g):
    g = g.copy()
    g.ensure_index()
    idx = g.index.copy()
    tree = M.build(g.positions, idx, strategy=M.SAH, target_leaf_size=10)
    return g, idx, tree, O.OracleMesh(g.positions, idx, tree.roots)


def test_sphere_closed_forms_on_a_box():
    g, idx, tree, om = _mesh(G.box(1, 1, 1))  # [-0.5, 0.5]^3
    c = np.array([[0, 0, 2.0], [0, 0, 2.0], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0, 0, 0], [3, 3, 3]], np.float32)
    r = np.array([1.5, np.nextafter(np.float32(1.5), np.float32(0)), 0.0, -1.0, 0.1, 4.2], np.float32)
    got = om.intersects_sphere(c, r)
    # top face at z = 0.5: distance exactly 1.5; a vertex at distance 0; the centre inside the box is
    # 0.5 from every face (0.1 reaches none); (3,3,3) is 2.5*sqrt(3) = 4.33 from the nearest corner
    assert got.tolist() == [True, False, True, False, False, False]


def test_sphere_agrees_with_closest_distance():
    g, idx, tree, om = _mesh(G.torus_knot(10, 3, 64, 12))
    rng = np.random.default_rng(2)
    lo, hi = g.bounds()
    c = rng.uniform(lo - 3, hi + 3, (400, 3)).astype(np.float32)
    ref = om.closest_point(c, brute=True)
    d = np.array([om.point_face_distance(int(f), c[i]) for i, f in enumerate(ref["face"])])
    r = (d * rng.uniform(0.5, 1.5, len(d))).astype(np.float32)
    got = om.intersects_sphere(c, r)
    assert np.array_equal(got, d <= r.astype(np.float64))
    assert 0.2 < got.mean() < 0.8
