"""GPU intersectsSphere through the C ABI against the oracle's brute force with the reference's
sphere-triangle test: random spheres around a torus knot with radii near the true distance, the box
closed forms (inclusive boundary), groups and indirect meshes."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G
import oracle_py as O
from helpers import build_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _spheres(om, lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo - 2, hi + 2, (n, 3)).astype(np.float32)
    ref = om.closest_point(c)
    d = np.array([om.point_face_distance(int(f), c[i]) for i, f in enumerate(ref["face"])])
    r = (d * rng.uniform(0.8, 1.2, n)).astype(np.float32)
    r[::50] = -1.0  # negative radii never intersect
    return c, r


@pytest.mark.parametrize("indirect", [False, True])
def test_spheres_match_oracle(ctx, indirect):
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 200, 24), indirect=indirect)
    mesh = ctx.upload(tree, g.positions, idx)
    om = O.OracleMesh(g.positions, idx, tree.roots, indirect=tree.indirect)
    lo, hi = g.bounds()
    c, r = _spheres(om, lo, hi, 3000, 3 + indirect)
    got = mesh.intersects_sphere(c, r)
    ref = om.intersects_sphere(c, r)
    assert np.array_equal(got, ref)
    assert 0.3 < got.mean() < 0.7
    mesh.free()


def test_sphere_box_closed_forms(ctx):
    g, idx, tree = build_mesh(G.box(1, 1, 1))
    mesh = ctx.upload(tree, g.positions, idx)
    c = np.array([[0, 0, 2.0], [0, 0, 2.0], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0, 0, 0], [3, 3, 3]], np.float32)
    r = np.array([1.5, np.nextafter(np.float32(1.5), np.float32(0)), 0.0, -1.0, 0.1, 4.2], np.float32)
    assert mesh.intersects_sphere(c, r).tolist() == [True, False, True, False, False, False]
    mesh.free()


def test_spheres_groups_and_many(ctx):
    geo = G.torus_knot(10, 3, 120, 16)
    geo.ensure_index()
    ntri = geo.index.size // 3
    g, idx, tree = build_mesh(geo, ranges=[(0, ntri // 2), (ntri // 2, ntri - ntri // 2)])
    mesh = ctx.upload(tree, g.positions, idx)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    lo, hi = g.bounds()
    c, r = _spheres(om, lo, hi, 8000, 9)
    assert np.array_equal(mesh.intersects_sphere(c, r), om.intersects_sphere(c, r))
    mesh.free()
