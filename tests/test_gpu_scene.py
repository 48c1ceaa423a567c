"""GPU scene raycast (mbvh_scene_raycast_first: a top-level tree over the instances' world boxes and
each instance's mesh, one launch) through the C ABI. The scene's answer for a ray is the closest of
the per-instance first hits: it must equal, bit for bit, the batched launch run on every (ray,
instance) pair (whose per-instance hits test_gpu_batch.py pins to the oracle) reduced by world
distance with ties to the lower instance index — the hit three's stable distance sort puts first
among the reference's per-instance BatchedMesh hits. A sample is also checked against the FP64
oracle directly (instance-local rays, world distances)."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh

pytestmark = pytest.mark.gpu

GEOMS = [lambda: G.torus_knot(1.0, 0.3, 120, 16), lambda: G.sphere(1.2, 40, 20), lambda: G.box(1.5, 0.7, 1.1),
         lambda: G.torus(1.0, 0.35, 16, 60), lambda: G.displaced_icosphere(12).ensure_index()]


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def meshes(ctx):
    built = [build_mesh(f()) for f in GEOMS]
    ms = [ctx.upload(t, g.positions, idx) for g, idx, t in built]
    yield built, ms
    for m in ms:
        m.free()


def _pairwise(ctx, ms, w, visible, side=M.FRONT_SIDE, near=0.0, far=float("inf")):
    """Every (instance, ray) pair through the batched launch, reduced to the closest instance."""
    batch = ctx.batch(ms)
    n = w.origins.shape[0]
    inst = [i for i in range(w.instance_geometry.shape[0]) if visible is None or visible[i]]
    face = np.full(n, M.NO_HIT, np.uint32)
    which = np.full(n, M.NO_HIT, np.uint32)
    dist = np.full(n, np.inf, np.float32)
    point = np.full((n, 3), np.nan, np.float32)
    bary = np.full((n, 3), np.nan, np.float32)
    if inst:
        k = len(inst)
        ri = np.repeat(np.asarray(inst, np.uint32), n)
        res = batch.raycast_first(w.matrix_world, w.matrix_world_inverse, w.instance_geometry,
                                  np.tile(w.origins, (k, 1)), np.tile(w.directions, (k, 1)), ri, side, near, far)
        F = res["face"].reshape(k, n)
        D = np.where(F == M.NO_HIT, np.float32(np.inf), res["distance"].reshape(k, n))
        best = np.argmin(D, axis=0)  # the first minimum: the lowest instance index on ties
        cols = np.arange(n)
        hit = F[best, cols] != M.NO_HIT
        face[hit] = F[best, cols][hit]
        which[hit] = np.asarray(inst, np.uint32)[best[hit]]
        dist[hit] = D[best, cols][hit]
        point[hit] = res["point"].reshape(k, n, 3)[best, cols][hit]
        bary[hit] = res["barycoord"].reshape(k, n, 3)[best, cols][hit]
    batch.free()
    return {"face": face, "instance": which, "distance": dist, "point": point, "barycoord": bary}


def _same(got, ref):
    assert np.array_equal(got["face"], ref["face"])
    assert np.array_equal(got["instance"], ref["instance"])
    hit = ref["face"] != M.NO_HIT
    for k in ("distance", "point", "barycoord"):
        assert np.array_equal(got[k][hit].view(np.uint32), ref[k][hit].view(np.uint32)), k
    assert np.all(np.isinf(got["distance"][~hit]))


@pytest.mark.parametrize("side", [M.FRONT_SIDE, M.DOUBLE_SIDE])
def test_scene_equals_closest_of_instances(ctx, meshes, side):
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 60, 4000, seed=3, non_uniform=0.2)
    visible = np.ones(60, np.uint8)
    visible[[4, 17, 33]] = 0
    print(f"case scene 60 instances side {side}", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry, visible)
    got = sc.raycast_first(w.origins, w.directions, side)
    ref = _pairwise(ctx, ms, w, visible, side)
    _same(got, ref)
    hits = ref["face"] != M.NO_HIT
    assert hits.mean() > 0.3 and len(np.unique(ref["instance"][hits])) > 20
    assert not np.any(np.isin(got["instance"][hits], [4, 17, 33]))
    sc.free()


def test_near_far_and_ties(ctx, meshes):
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 24, 3000, seed=5)
    # instance 7 duplicated as 8 and 20: equal distances go to the lowest index
    for k in (8, 20):
        w.matrix_world[k] = w.matrix_world[7]
        w.matrix_world_inverse[k] = w.matrix_world_inverse[7]
        w.instance_geometry[k] = w.instance_geometry[7]
    print("case near/far and ties", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    for near, far in ((0.0, float("inf")), (5.0, 40.0), (30.0, 31.0)):
        got = sc.raycast_first(w.origins, w.directions, M.FRONT_SIDE, near, far)
        _same(got, _pairwise(ctx, ms, w, None, M.FRONT_SIDE, near, far))
    got = sc.raycast_first(w.origins, w.directions)
    assert not np.any(np.isin(got["instance"], [8, 20]))
    sc.free()


def test_oracle_sample(ctx, meshes):
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 30, 400, seed=9)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    got = sc.raycast_first(w.origins, w.directions)
    oms = [O.OracleMesh(g.positions, idx, t.roots) for g, idx, t in built]
    n = w.origins.shape[0]
    best = np.full(n, np.inf)
    per = np.full((30, n), np.inf)
    for i in range(30):
        mw = w.matrix_world[i].astype(np.float64).reshape(4, 4).T
        mi = w.matrix_world_inverse[i].astype(np.float64).reshape(4, 4).T
        lo = w.origins @ mi[:3, :3].T + mi[:3, 3]
        ld = w.directions @ mi[:3, :3].T
        r = oms[w.instance_geometry[i]].raycast_first(lo.astype(np.float32), ld.astype(np.float32))
        hit = r["face"] != M.NO_HIT
        pw = r["point"] @ mw[:3, :3].T + mw[:3, 3]
        dw = np.linalg.norm(pw - w.origins.astype(np.float64), axis=1)
        per[i, hit] = dw[hit]
    best = per.min(axis=0)
    gh = got["face"] != M.NO_HIT
    assert np.array_equal(gh, np.isfinite(best))
    gi = got["instance"][gh]
    cols = np.nonzero(gh)[0]
    # the chosen instance's oracle distance is the best one (ties within 1e-6), and the distance agrees
    assert np.all(per[gi, cols] <= best[gh] * (1 + 1e-6) + 1e-12)
    assert np.all(np.abs(got["distance"][gh] - best[gh]) <= 1e-5 * np.maximum(best[gh], 1.0))
    sc.free()


def test_empty_and_device_form(ctx, meshes):
    torch = pytest.importorskip("torch")
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 12, 2000, seed=13)
    print("case empty / device form", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry, np.zeros(12, np.uint8))
    got = sc.raycast_first(w.origins, w.directions)
    assert np.all(got["face"] == M.NO_HIT) and np.all(got["instance"] == M.NO_HIT)
    sc.free()
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    host = sc.raycast_first(w.origins, w.directions)
    dev = torch.device("cuda:0")
    o = torch.from_numpy(w.origins).to(dev)
    d = torch.from_numpy(w.directions).to(dev)
    n = o.shape[0]
    face = torch.empty(n, dtype=torch.int32, device=dev)
    inst = torch.empty(n, dtype=torch.int32, device=dev)
    dist = torch.empty(n, device=dev)
    torch.cuda.synchronize()
    sc.raycast_first_d(o.data_ptr(), d.data_ptr(), n, face.data_ptr(), inst.data_ptr(), dist.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(face.cpu().numpy().view(np.uint32), host["face"])
    assert np.array_equal(inst.cpu().numpy().view(np.uint32), host["instance"])
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), host["distance"].view(np.uint32))
    with pytest.raises(M.MbvhError):
        ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, np.full(12, 99, np.uint32))
    sc.free()


def _scene_points(w, n, seed):
    """World points around the instances: each near a random instance's origin, up to ~2 units away."""
    rng = np.random.default_rng(seed)
    t = w.matrix_world[:, 12:15].astype(np.float64)
    pick = rng.integers(0, t.shape[0], n)
    return (t[pick] + rng.normal(0, 1.2, (n, 3))).astype(np.float32)


def _world_tri(built, w, inst, face):
    g, idx, _ = built[w.instance_geometry[inst]]
    v = g.positions[idx[3 * face:3 * face + 3]].astype(np.float64)
    m = w.matrix_world[inst].astype(np.float64).reshape(4, 4).T
    return v @ m[:3, :3].T + m[:3, 3]


@pytest.mark.parametrize("non_uniform", [0.0, 0.3])
def test_scene_closest_point_equals_oracle(ctx, meshes, non_uniform):
    """Scene closest point (mbvh_scene_closest_point) against the oracle's brute force over every
    visible instance's world triangles: the instance and face are the oracle's, or an equal-distance
    tie (the oracle's distance to the GPU's world triangle within 1e-6 of the best); distance and point
    to 1e-5 of the scene scale. Rigid and non-uniformly scaled instances, some hidden."""
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 24, 8, seed=17, non_uniform=non_uniform)
    visible = np.ones(24, np.uint8)
    visible[[3, 11]] = 0
    pts = _scene_points(w, 300, 5)
    print(f"case scene closest point non_uniform {non_uniform}", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry, visible)
    got = sc.closest_point(pts)
    oms = [O.OracleMesh(g.positions, idx, t.roots) for g, idx, t in built]
    ref = O.scene_closest_point(oms, w.matrix_world, w.instance_geometry, visible, pts)
    scale = float(np.abs(w.matrix_world[:, 12:15]).max()) + 1.0
    assert np.all(got["face"] != M.NO_HIT) and not np.any(np.isin(got["instance"], [3, 11]))
    np.testing.assert_allclose(got["distance"], ref["distance"], rtol=1e-5, atol=1e-5 * scale)
    same = (got["instance"] == ref["instance"]) & (got["face"] == ref["face"])
    for i in np.nonzero(~same)[0]:
        d, _, _ = None, None, None
        cp = O.closest_point_triangle(pts[i].astype(np.float64), *_world_tri(built, w, got["instance"][i], got["face"][i]))
        d = np.linalg.norm(cp - pts[i])
        assert d <= ref["distance"][i] * (1 + 1e-6) + 1e-9, (i, d, ref["distance"][i])
    np.testing.assert_allclose(got["point"][same], ref["point"][same], atol=1e-5 * scale)
    assert same.mean() > 0.95
    # barycoords reproduce the point on the world triangle
    for i in np.nonzero(same)[0][:50]:
        a, b, c = _world_tri(built, w, got["instance"][i], got["face"][i])
        bc = got["barycoord"][i].astype(np.float64)
        np.testing.assert_allclose(bc[0] * a + bc[1] * b + bc[2] * c, got["point"][i], atol=1e-4 * scale)
    sc.free()


def test_scene_closest_point_ignores_a_stale_inverse_and_rejects_projective(ctx, meshes):
    """The closest point's pruning bounds use the scene's own double-precision inverse of each
    matrixWorld: a caller inverse that is stale by 5% (or garbage) changes nothing, and the results
    equal the oracle's brute force. A matrixWorld whose last row is not 0 0 0 1 is rejected."""
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 12, 8, seed=41, non_uniform=0.3)
    pts = _scene_points(w, 300, 7)
    print("case scene closest point stale inverse", flush=True)
    good = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    ref_gpu = good.closest_point(pts)
    good.free()
    stale = w.matrix_world_inverse.copy()
    stale[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]] *= 1.05
    stale[1] = np.float32(7.0)
    sc = ctx.scene(ms, w.matrix_world, stale, w.instance_geometry)
    got = sc.closest_point(pts)
    sc.free()
    for k in ("face", "instance"):
        assert np.array_equal(got[k], ref_gpu[k]), k
    assert np.array_equal(got["distance"].view(np.uint32), ref_gpu["distance"].view(np.uint32))
    oms = [O.OracleMesh(g.positions, idx, t.roots) for g, idx, t in built]
    ref = O.scene_closest_point(oms, w.matrix_world, w.instance_geometry, None, pts)
    scale = float(np.abs(w.matrix_world[:, 12:15]).max()) + 1.0
    np.testing.assert_allclose(got["distance"], ref["distance"], rtol=1e-5, atol=1e-5 * scale)
    proj = w.matrix_world.copy()
    proj[5, 15] = 2.0
    with pytest.raises(M.MbvhError) as e:
        ctx.scene(ms, proj, w.matrix_world_inverse, w.instance_geometry)
    assert e.value.code == -4 and "affine" in str(e.value)


def test_scene_closest_point_device_form(ctx, meshes):
    torch = pytest.importorskip("torch")
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 16, 8, seed=23)
    pts = _scene_points(w, 2000, 9)
    print("case scene closest point device form", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    host = sc.closest_point(pts)
    dev = torch.device("cuda:0")
    q = torch.from_numpy(pts).to(dev)
    face = torch.empty(pts.shape[0], dtype=torch.int32, device=dev)
    inst = torch.empty(pts.shape[0], dtype=torch.int32, device=dev)
    dist = torch.empty(pts.shape[0], device=dev)
    torch.cuda.synchronize()
    sc.closest_point_d(q.data_ptr(), pts.shape[0], face.data_ptr(), inst.data_ptr(), dist.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(face.cpu().numpy().view(np.uint32), host["face"])
    assert np.array_equal(inst.cpu().numpy().view(np.uint32), host["instance"])
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), host["distance"].view(np.uint32))
    # repeatable
    again = sc.closest_point(pts)
    assert np.array_equal(again["face"], host["face"]) and np.array_equal(again["instance"], host["instance"])
    sc.free()


def test_scene_closest_point_sheared_and_anisotropic(ctx, meshes):
    """Instance matrices at the edges of the frame bound: strongly anisotropic scales (the local-frame
    bound with its rounding margins) and near-parallel columns (no positive singular-value bound, so
    the world-box path), against the oracle's brute force."""
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 8, 8, seed=31)
    rng = np.random.default_rng(3)
    mats = [np.diag([30.0, 1.0, 0.03, 1.0]), np.diag([0.05, 2.0, 40.0, 1.0])]
    sh = np.eye(4)
    sh[:3, 2] = [0.98, 0.05, 0.2]  # third column nearly the first: Gershgorin gives no positive bound
    mats.append(sh)
    sh2 = np.eye(4)
    sh2[:3, :3] = [[1.0, 0.9, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 3.0]]
    mats.append(sh2)
    for k, m in enumerate(mats):
        m = m.copy()
        m[:3, 3] = rng.uniform(-8, 8, 3)
        w.matrix_world[k] = m.T.reshape(-1).astype(np.float32)
        w.matrix_world_inverse[k] = np.linalg.inv(w.matrix_world[k].astype(np.float64).reshape(4, 4).T).T.reshape(-1)
    pts = _scene_points(w, 300, 13)
    print("case scene closest point sheared / anisotropic", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    got = sc.closest_point(pts)
    oms = [O.OracleMesh(g.positions, idx, t.roots) for g, idx, t in built]
    ref = O.scene_closest_point(oms, w.matrix_world, w.instance_geometry, None, pts)
    scale = float(np.abs(w.matrix_world[:, 12:15]).max()) + 40.0
    np.testing.assert_allclose(got["distance"], ref["distance"], rtol=1e-5, atol=1e-5 * scale)
    same = (got["instance"] == ref["instance"]) & (got["face"] == ref["face"])
    for i in np.nonzero(~same)[0]:
        cp = O.closest_point_triangle(pts[i].astype(np.float64), *_world_tri(built, w, got["instance"][i], got["face"][i]))
        assert np.linalg.norm(cp - pts[i]) <= ref["distance"][i] * (1 + 1e-6) + 1e-9
    assert np.isin(got["instance"], np.arange(4)).sum() > 20  # the edge-case instances do win points
    sc.free()


@pytest.mark.parametrize("kind", ["rigid", "non_uniform", "sheared"])
def test_scene_closest_point_lanes_agree(ctx, meshes, kind):
    """One lane and four lanes per point (mbvh_ctx_set_point_lanes 1 / 4) give the same faces,
    instances, distances, points and barycoords bit for bit (6000 points: the spatial order runs)."""
    built, ms = meshes
    geos = [b[0] for b in built]
    w = W.scene_rays(geos, 20, 8, seed=51, non_uniform=0.5 if kind == "non_uniform" else 0.0)
    if kind == "sheared":
        for k in range(0, 20, 3):
            m = w.matrix_world[k].astype(np.float64).reshape(4, 4).T
            m[:3, 2] += 0.6 * m[:3, 0]
            m[:3, 1] *= 7.0
            w.matrix_world[k] = m.T.reshape(-1).astype(np.float32)
            w.matrix_world_inverse[k] = np.linalg.inv(m).T.reshape(-1).astype(np.float32)
    pts = _scene_points(w, 6000, 13)
    print(f"case scene closest point lanes {kind}", flush=True)
    sc = ctx.scene(ms, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    try:
        ctx.set_point_lanes(1)
        one = sc.closest_point(pts)
        ctx.set_point_lanes(4)
        four = sc.closest_point(pts)
    finally:
        ctx.set_point_lanes(0)
        sc.free()
    assert np.all(one["face"] != M.NO_HIT)
    for k in ("face", "instance", "distance", "point", "barycoord"):
        assert np.array_equal(np.asarray(one[k]).view(np.uint32), np.asarray(four[k]).view(np.uint32)), k


def test_scene_closest_point_long_walks(ctx):
    """Points near the centre of a big closed mesh are nearly equidistant from all of its triangles,
    so their walks run past the four-lane kernel's step budget and finish in the wave-per-point pass:
    the results equal the one-lane kernel's bit for bit and the oracle's brute force on a sample."""
    built = [build_mesh(G.sphere(1.5, 400, 200)), build_mesh(G.torus_knot(1.0, 0.3, 120, 16))]
    ms = [ctx.upload(t, g.positions, idx) for g, idx, t in built]
    rng = np.random.default_rng(61)
    mats = []
    for k, (scale, geo) in enumerate([((1.0, 1.0, 1.0), 0), ((1.3, 0.8, 1.1), 0), ((1.0, 1.0, 1.0), 1)]):
        m = np.eye(4)
        m[:3, :3] = np.diag(scale)
        m[:3, 3] = [6.0 * k, 0.5 * k, -1.0 * k]
        mats.append(m)
    world = np.stack([m.T.reshape(-1) for m in mats]).astype(np.float32)
    inverse = np.stack([np.linalg.inv(m).T.reshape(-1) for m in mats]).astype(np.float32)
    geometry = np.array([0, 0, 1], np.uint32)
    d = rng.standard_normal((600, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rho = np.concatenate([rng.uniform(0.0, 0.15, 300) * 1.5, rng.uniform(0.5, 3.0, 300)])
    centre = world[rng.integers(0, 2, 600), 12:15].astype(np.float64)
    pts = (centre + d * rho[:, None]).astype(np.float32)
    print("case scene closest point long walks", flush=True)
    sc = ctx.scene(ms, world, inverse, geometry)
    try:
        ctx.set_point_lanes(1)
        one = sc.closest_point(pts)
        ctx.set_point_lanes(0)
        got = sc.closest_point(pts)
    finally:
        ctx.set_point_lanes(0)
        sc.free()
    for k in ("face", "instance", "distance", "point", "barycoord"):
        assert np.array_equal(np.asarray(one[k]).view(np.uint32), np.asarray(got[k]).view(np.uint32)), k
    oms = [O.OracleMesh(g.positions, idx, t.roots) for g, idx, t in built]
    sample = np.r_[0:20, 300:320]
    ref = O.scene_closest_point(oms, world, geometry, None, pts[sample])
    np.testing.assert_allclose(got["distance"][sample], ref["distance"], rtol=1e-6, atol=1e-7)
    assert (got["instance"][sample] == ref["instance"]).mean() >= 0.95
    for m in ms:
        m.free()
