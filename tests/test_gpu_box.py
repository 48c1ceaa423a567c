"""intersectsBox on the GPU (mbvh_intersects_box) through the C ABI: per box, the same answer as the
oracle's brute force with the reference's OrientedBox tests (both evaluate the reference's double
expressions in its order, so the answers are equal, boundary cases included): random oriented boxes
of every size around meshes (indexed, indirect), boxes touching a face exactly, degenerate boxes;
the device form equals the host form."""
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


def _boxes(n, lo, hi, seed, size=(0.01, 0.4)):
    rng = np.random.default_rng(seed)
    ext = np.linalg.norm(hi - lo)
    mats = np.empty((n, 16))
    for i in range(n):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        m = np.eye(4)
        m[:3, :3] = R @ np.diag(rng.uniform(0.5, 2.0, 3))
        m[:3, 3] = lo + (hi - lo) * rng.uniform(-0.1, 1.1, 3)
        mats[i] = m.T.reshape(-1)
    half = ext * rng.uniform(size[0], size[1], (n, 3))
    return -half, half, mats


@pytest.mark.parametrize("name,make,indirect", [
    ("torusknot", lambda: G.torus_knot(1, 0.3, 200, 24), False),
    ("sphere_indirect", lambda: G.sphere(1, 48, 24), True),
    ("displaced_icosphere", lambda: G.displaced_icosphere(20).ensure_index(), False),
])
def test_boxes_equal_oracle(ctx, name, make, indirect):
    print(f"case {name}", flush=True)
    g, idx, tree = build_mesh(make(), indirect=indirect)
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    mn, mx, mats = _boxes(3000, lo, hi, 5, size=(0.003, 0.12))
    got = mesh.intersects_box(mn, mx, mats)
    om = O.OracleMesh(g.positions, idx, tree.roots, indirect=tree.indirect)
    ref = om.intersects_box(mn, mx, mats)
    assert np.array_equal(got, ref)
    assert 0.05 < got.mean() < 0.95
    mesh.free()


def test_touching_and_degenerate(ctx):
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [0, 1, 0], [0, 0, -1]], np.float32)
    g = G.Geometry(p, np.arange(6, dtype=np.uint16))
    g, idx, tree = build_mesh(g)
    mesh = ctx.upload(tree, g.positions, idx)
    I4 = np.eye(4).T.reshape(-1)
    cases = [([0.2, 0.2, 0.0], [0.3, 0.3, 1.0]),        # a face on the triangle's plane (z = 0): touching
             ([0.2, 0.2, 1e-7], [0.3, 0.3, 1.0]),       # just above
             ([-1.0, 0.2, 0.2], [0.0, 0.3, 0.3]),       # a face on the x = 0 plane of the second triangle? (z > 0)
             ([0.3, 0.3, -0.5], [0.3, 0.3, 0.5]),       # a zero-width box (a segment) through the face
             ([0.5, 0.5, 0.0], [0.5, 0.5, 0.0]),        # a point on the hypotenuse
             ([0.6, 0.6, 0.0], [0.6, 0.6, 0.0])]        # a point past it: a zero-size box has zero axes,
                                                        # so the reference's axis set (box axes, the
                                                        # triangle's normal and edges, their crosses) finds
                                                        # no separation and reports an intersection
    mn = np.array([c[0] for c in cases])
    mx = np.array([c[1] for c in cases])
    mats = np.tile(I4, (len(cases), 1))
    got = mesh.intersects_box(mn, mx, mats)
    ref = O.OracleMesh(g.positions, idx, tree.roots).intersects_box(mn, mx, mats)
    assert np.array_equal(got, ref)
    assert got.tolist()[:2] == [True, False] and got[3] and got[4] and got[5]
    mesh.free()


def test_device_form(ctx):
    torch = pytest.importorskip("torch")
    g, idx, tree = build_mesh(G.torus_knot(1, 0.3, 100, 12))
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    mn, mx, mats = _boxes(5000, lo, hi, 8)
    host = mesh.intersects_box(mn, mx, mats)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    dmn, dmx, dm = t(mn), t(mx), t(mats)
    out = torch.empty(len(mn), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    assert ctx._lib.mbvh_intersects_box_d(ctx.handle, mesh.handle, dmn.data_ptr(), dmx.data_ptr(), dm.data_ptr(),
                                          len(mn), out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().astype(bool), host)
    mesh.free()


def _euler_xyz(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rx = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rz = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return rx @ ry @ rz  # three's Euler order 'XYZ'


def test_reference_intersects_cases(ctx):
    """The reference's own IntersectsSphere / IntersectsBox cases (test/ShapeCasts.test.js:567-698) on
    SphereGeometry(1, 50, 50): spheres on, inside and outside the surface; boxes rotated by
    Euler(pi/4, pi/4, 0) outside, with a corner inside, enclosing, inside, and cut by the surface."""
    g, idx, tree = build_mesh(G.sphere(1, 50, 50))
    m = ctx.upload(tree, g.positions, idx)
    print("case reference sphere / box cases", flush=True)
    got = m.intersects_sphere(np.array([[0, 1, 0], [0, 0, 0], [0, 2.01, 0]], np.float32),
                              np.array([0.01, 0.9, 0.9], np.float32))
    assert [bool(x) for x in got] == [True, False, False]
    R = _euler_xyz(np.pi / 4, np.pi / 4, 0)
    cases = [((-1, -1, -1), (1, 1, 1), (0, 3, 0), False), ((-1, -1, -1), (1, 1, 1), (0, 2, 0), True),
             ((-10, -10, -10), (10, 10, 10), (0, 0, 0), True), ((-.5, -.5, -.5), (.5, .5, .5), (0, 0, 0), False),
             ((-10, 0, -10), (10, 10, 10), (0, 0, 0), True)]
    mats = []
    for _, _, pos, _ in cases:
        mm = np.eye(4)
        mm[:3, :3] = R
        mm[:3, 3] = pos
        mats.append(mm.T.reshape(-1))
    res = m.intersects_box(np.array([c[0] for c in cases], np.float64), np.array([c[1] for c in cases], np.float64),
                           np.array(mats))
    assert [bool(x) for x in res] == [c[3] for c in cases]
    m.free()
