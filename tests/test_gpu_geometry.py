"""closestPointToGeometry on the GPU (mbvh_closest_point_to_geometry) through the C ABI, against the
oracle (an independent FP64 restatement of distanceToTriangle, walking the tree per other triangle):
- the distance equals the oracle's to 1e-9 relative (both in double; different formulas round
  differently), the pair is the oracle's or an equal-distance tie (the oracle's distance for the GPU's
  pair within 1e-6 relative of the minimum), the points agree to 1e-6 of the scene size and the other
  point / other distance follow from the inverse matrix;
- separated, touching and interpenetrating placements with rotation and scale; uint16, uint32 and
  non-indexed other geometry; indirect and multi-root (group) meshes;
- maxThreshold below the distance (none) and above it; an empty other geometry; an index entry out of
  range (error); the device form equals the host form; repeated calls give the same pair."""
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


def _placement(seed, offset, scale=1.0):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = R * scale
    m[:3, 3] = offset
    return m.T.reshape(-1)  # column-major, as Matrix4.elements


def _tri_of(g, idx, face):
    v = idx[3 * face:3 * face + 3] if idx is not None else np.arange(3 * face, 3 * face + 3)
    return g.positions[v].astype(np.float64)


def _check(got, ref, om_tri, other_tri, e, size):
    if ref is None:
        assert got is None
        return
    assert got is not None
    d = ref["distance"]
    assert abs(got["distance"] - d) <= 1e-9 * max(d, size) , (got["distance"], d)
    if (got["face_index"], got["other_face_index"]) != (ref["face_index"], ref["other_face_index"]):
        # a tie: the GPU's pair is as close as the oracle's best
        dd, _, _ = O.triangle_distance(om_tri(got["face_index"]), other_tri(got["other_face_index"]))
        assert dd <= d + 1e-6 * max(d, size), (dd, d)
        return
    np.testing.assert_allclose(got["point"], ref["point"], atol=1e-6 * size)
    m = np.asarray(e, np.float64).reshape(4, 4).T
    inv = np.linalg.inv(m)
    p2 = inv[:3, :3] @ ref["point_in_mesh_frame"] + inv[:3, 3]
    np.testing.assert_allclose(got["other_point"], p2, atol=1e-6 * size)
    p1 = inv[:3, :3] @ got["point"] + inv[:3, 3]
    assert got["other_distance"] == pytest.approx(np.linalg.norm(p1 - got["other_point"]), rel=1e-9, abs=1e-12)


CASES = [
    # name, mesh, indirect, other, placements (offset, scale)
    ("knot_vs_sphere", lambda: G.torus_knot(1.0, 0.3, 160, 16), False, lambda: G.sphere(0.4, 24, 12),
     [([0, 0, 2.5], 1.0), ([1.3, 0.3, 0.1], 0.7), ([0, 0, 0], 1.2), ([4, 4, 4], 0.5)]),
    ("icosphere_vs_knot", lambda: G.displaced_icosphere(16).ensure_index(), False,
     lambda: G.torus_knot(0.5, 0.15, 100, 12), [([0, 0, 3.0], 1.0), ([0.6, 0, 0], 1.0), ([0, 2.2, 0], 2.0)]),
    ("sphere_indirect_vs_box", lambda: G.sphere(1.0, 40, 20), True, lambda: G.box(0.6, 0.3, 0.9),
     [([0, 0, 1.8], 1.0), ([0.9, 0, 0], 1.0)]),
]


@pytest.mark.parametrize("name,mesh,indirect,other,places", CASES)
@pytest.mark.parametrize("layout", ["u16", "u32", "soup"])
def test_equals_oracle(ctx, name, mesh, indirect, other, places, layout):
    g, idx, tree = build_mesh(mesh(), indirect=indirect)
    m = ctx.upload(tree, g.positions, idx)
    om = O.OracleMesh(g.positions, idx, tree.roots, indirect=tree.indirect)
    og = other().ensure_index()
    opos, oidx = og.positions, og.index.astype(np.uint16 if layout == "u16" else np.uint32)
    if layout == "soup":
        opos, oidx = og.positions[og.index], None
    size = float(np.abs(g.positions).max())
    slots = (lambda s: int(tree.indirect[s])) if tree.indirect is not None else (lambda s: s)
    om_tri = lambda s: _tri_of(g, idx, slots(s))

    for k, (off, scale) in enumerate(places):
        e = _placement(100 + k, off, scale)
        mm = np.asarray(e).reshape(4, 4).T
        other_tri = lambda j: (_tri_of(G.Geometry(opos, oidx), oidx, j) @ mm[:3, :3].T + mm[:3, 3])
        print(f"case {name} {layout} placement {k}", flush=True)
        got = m.closest_point_to_geometry(opos, oidx, e)
        ref = om.closest_to_geometry(opos, oidx, e)
        _check(got, ref, om_tri, other_tri, e, size)
    m.free()


def test_thresholds_empty_and_errors(ctx):
    g, idx, tree = build_mesh(G.torus_knot(1.0, 0.3, 120, 12))
    m = ctx.upload(tree, g.positions, idx)
    og = G.sphere(0.4, 16, 8).ensure_index()
    e = _placement(7, [0, 0, 3.0])
    print("case thresholds", flush=True)
    full = m.closest_point_to_geometry(og.positions, og.index, e)
    assert full is not None and full["distance"] > 0.5
    assert m.closest_point_to_geometry(og.positions, og.index, e, max_threshold=full["distance"]) is None
    again = m.closest_point_to_geometry(og.positions, og.index, e, min_threshold=10.0,
                                        max_threshold=full["distance"] * 1.0001)
    assert again["distance"] == full["distance"] and again["face_index"] == full["face_index"]
    # repeated calls: the same pair, bit for bit
    for _ in range(3):
        r = m.closest_point_to_geometry(og.positions, og.index, e)
        assert (r["distance"], r["face_index"], r["other_face_index"]) == \
            (full["distance"], full["face_index"], full["other_face_index"])
    assert m.closest_point_to_geometry(og.positions[:0], np.zeros(0, np.uint32), e) is None
    bad = og.index.copy()
    bad[5] = og.positions.shape[0]
    with pytest.raises(M.MbvhError):
        m.closest_point_to_geometry(og.positions, bad, e)
    m.free()


def test_groups_and_device_form(ctx):
    torch = pytest.importorskip("torch")
    g = G.torus_knot(1.0, 0.3, 120, 12).ensure_index()
    n = g.tri_count
    g, idx, tree = build_mesh(g, ranges=[(0, n // 3), (n // 3, n - n // 3)])
    assert len(tree.roots) == 2
    m = ctx.upload(tree, g.positions, idx)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    og = G.torus_knot(0.4, 0.1, 60, 8).ensure_index()
    e = _placement(3, [0.9, -0.4, 0.2])
    print("case groups / device form", flush=True)
    got = m.closest_point_to_geometry(og.positions, og.index, e)
    ref = om.closest_to_geometry(og.positions, og.index, e)
    mm = np.asarray(e).reshape(4, 4).T
    _check(got, ref, lambda s: _tri_of(g, idx, s),
           lambda j: _tri_of(og, og.index, j) @ mm[:3, :3].T + mm[:3, 3], e, 1.5)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(og.positions.reshape(-1).copy()).to(dev)
    oi = torch.from_numpy(og.index.astype(np.int32)).to(dev)
    out = torch.zeros(M.GEOMETRY_HIT.itemsize, dtype=torch.uint8, device=dev)
    ee = np.ascontiguousarray(e, np.float64)
    torch.cuda.synchronize()
    M._check(m.ctx._lib.mbvh_closest_point_to_geometry_d(
        m.ctx.handle, m.handle, pos.data_ptr(), og.positions.shape[0], 3, 0, oi.data_ptr(), 4, og.tri_count,
        ee.ctypes.data, 0.0, float("inf"), out.data_ptr(), None))
    torch.cuda.synchronize()
    rec = np.frombuffer(out.cpu().numpy().tobytes(), M.GEOMETRY_HIT)[0]
    assert rec["status"] == 0
    assert rec["distance"] == got["distance"] and rec["face_index"] == got["face_index"]
    assert rec["other_face_index"] == got["other_face_index"]
    m.free()


def _compose(pos, scale):
    """Matrix4.compose(pos, identity quaternion, (scale, scale, scale)) as column-major elements."""
    m = np.eye(4)
    m[:3, :3] *= scale
    m[:3, 3] = pos
    return m.T.reshape(-1)


def test_reference_distance_to_geometry_cases(ctx):
    """The reference's own 'Distance To Geometry' cases (test/ShapeCasts.test.js:755-870): a BVH of
    SphereGeometry(1, 20, 20) against SphereGeometry(1, 5, 5) shrunk to the center (distance ~ the
    radius), placed on the surface (0), and at random distances from it (|1 - r| - 0.1 within the
    faceting error); spheres (1, 2n, n) three units apart (the returned points are ~1 apart)."""
    g, idx, tree = build_mesh(G.sphere(1, 20, 20))
    m = ctx.upload(tree, g.positions, idx)
    other = G.sphere(1, 5, 5).ensure_index()
    print("case reference sphere cases", flush=True)
    r = m.closest_point_to_geometry(other.positions, other.index, _compose([0, 0, 0], 0.001))
    assert 1 - 0.05 <= r["distance"] <= 1
    r = m.closest_point_to_geometry(other.positions, other.index, _compose([0, 1, 0], 0.1))
    assert r["distance"] == 0
    rng = np.random.default_rng(11)
    for _ in range(10):
        p = rng.random(3) - 0.5
        length = rng.random() * 3
        p = p / np.linalg.norm(p) * length
        expected = abs(1 - length)
        expected = 0 if expected < 0.1 else expected - 0.1
        r = m.closest_point_to_geometry(other.positions, other.index, _compose(p, 0.1))
        assert expected - 0.1 <= r["distance"] <= expected + 0.1
    m.free()
    for n in (4, 8, 16):
        g, idx, tree = build_mesh(G.sphere(1, 2 * n, n))
        m = ctx.upload(tree, g.positions, idx)
        o = G.sphere(1, 2 * n, n).ensure_index()
        e = _compose([3, 0, 0], 1.0)
        r = m.closest_point_to_geometry(o.positions, o.index, e)
        p2 = r["other_point"] + np.array([3.0, 0, 0])  # target2.point.applyMatrix4(matrix)
        assert abs(np.linalg.norm(r["point"] - p2) - 1) < 0.05
        assert abs(r["distance"] - np.linalg.norm(r["point"] - p2)) < 1e-9
        m.free()
