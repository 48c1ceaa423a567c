"""Double-precision queries through the C ABI (mbvh_raycast_first_f64, mbvh_raycast_f64,
mbvh_closest_point_f64, their device forms and the batched local-ray launch) against the FP64 oracle
run on the same double inputs.

three.js hands the reference JS numbers (doubles): the rays raycastObject3D builds in the mesh frame
and the points given to closestPointToPoint are not float32 values. The decisive cases here straddle a
shared edge by 1e-11 of the mesh size: the double ray (point) is on one side of the edge, its float32
rounding may be on either side. Faces must equal the oracle's on them (the two straddled faces are
not tied for the double input: only rounding-level ties elsewhere are allowed), and the test checks that rounding the inputs to float32
would have changed the oracle's answer on many of them, so a path that rounds its inputs fails."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh, check_closest_parity, check_raycast_parity, point_faces_tied

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def knot(ctx):
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 400, 100))
    mesh = ctx.upload(tree, g.positions, idx)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    yield g, idx, tree, mesh, om
    mesh.free()


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _shared_edges(idx):
    """(edge vertex pairs, face on one side, face on the other) for edges shared by two faces."""
    tri = idx.reshape(-1, 3).astype(np.int64)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    f = np.tile(np.arange(tri.shape[0]), 3)
    key = np.sort(e, axis=1)
    order = np.lexsort((key[:, 1], key[:, 0]))
    ks, fs = key[order], f[order]
    same = np.nonzero(np.all(ks[1:] == ks[:-1], axis=1))[0]
    return ks[same], fs[same], fs[same + 1]


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _straddle(g, idx, n, seed, eps=1e-11, crease=0.0):
    """Points P on shared edges, the surface normal N there (mean of the two faces'), the tangent w
    across the edge, and the mesh scale L. crease: only edges whose faces' normals n1, n2 have
    1 - n1.n2 above it."""
    pos = g.positions.astype(np.float64)
    edges, fa, fb = _shared_edges(idx)
    tri = idx.reshape(-1, 3).astype(np.int64)

    def normal(f):
        v = pos[tri[f]]
        return _unit(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]))

    if crease > 0:
        keep = 1 - np.sum(normal(fa) * normal(fb), axis=1) > crease
        edges, fa, fb = edges[keep], fa[keep], fb[keep]
    rng = np.random.default_rng(seed)
    pick = rng.choice(edges.shape[0], n, replace=False)
    a, b = pos[edges[pick, 0]], pos[edges[pick, 1]]
    P = a + rng.uniform(0.3, 0.7, (n, 1)) * (b - a)
    N = _unit(normal(fa[pick]) + normal(fb[pick]))
    w = _unit(np.cross(N, _unit(b - a)))
    lo, hi = g.bounds()
    L = float(np.linalg.norm(hi.astype(np.float64) - lo))
    return P, N, w, L, eps * L, rng


def _f32_rounded(x):
    return np.asarray(x, np.float64).astype(np.float32)


def test_edge_straddling_rays_equal_the_oracle_exactly(ctx, knot):
    """4,000 rays aimed across shared edges: half with the double direction carrying the offset, half
    with the double origin carrying it."""
    g, idx, tree, mesh, om = knot
    P, N, w, L, e, rng = _straddle(g, idx, 1000, 3)
    O32 = _f32_rounded(P + 0.01 * L * N).astype(np.float64)  # float origins just off the surface
    sign = np.repeat(np.array([[1.0], [-1.0]]), P.shape[0], axis=0)
    P2, O2, w2 = np.tile(P, (2, 1)), np.tile(O32, (2, 1)), np.tile(w, (2, 1))
    # (a) float origin, the double direction carries the offset
    da = _unit(P2 + sign * e * w2 - O2)
    oa = O2
    # (b) float direction, the double origin carries the offset: the ray through P +- e w
    db = _f32_rounded(_unit(P2 - O2)).astype(np.float64)
    ob = P2 + sign * e * w2 - 0.01 * L * db
    origins, dirs = np.concatenate([oa, ob]), np.concatenate([da, db])
    print(f"case edge-straddling rays: {origins.shape[0]} rays, offset {e:.3g}", flush=True)
    for side in (M.DOUBLE_SIDE, M.FRONT_SIDE):
        got = mesh.raycast_first(origins, dirs, side=side)
        ref = om.raycast_first(origins, dirs, side=side)
        hit = ref["face"] != O.NO_HIT
        assert np.array_equal(got["face"], ref["face"]), (side, np.count_nonzero(got["face"] != ref["face"]))
        de = np.abs(got["distance"][hit] - ref["distance"][hit]) / np.maximum(ref["distance"][hit], 1)
        assert de.max() <= 1e-5
        if side == M.DOUBLE_SIDE:
            assert hit.mean() > 0.95
            # the same rays rounded to float32 give other faces on many of them
            r32 = om.raycast_first(_f32_rounded(origins), _f32_rounded(dirs), side=side)
            changed = np.count_nonzero(r32["face"] != ref["face"])
            print("faces that float32 rounding would change:", changed, flush=True)
            assert changed > 0.2 * origins.shape[0]
    # the device form with the rays split by the caller gives the same bits
    torch = pytest.importorskip("torch")
    hi_o, hi_d = _f32_rounded(origins), _f32_rounded(dirs)
    lo_o = (origins - hi_o.astype(np.float64)).astype(np.float32)
    lo_d = (dirs - hi_d.astype(np.float64)).astype(np.float32)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (hi_o, lo_o, hi_d, lo_d)]
    n = origins.shape[0]
    face = torch.empty(n, dtype=torch.int32, device=dev)
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    mesh.raycast_first_ff_d(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(), n, face.data_ptr(),
                            dist.data_ptr(), side=M.DOUBLE_SIDE)
    ctx.synchronize()
    host = mesh.raycast_first(origins, dirs, side=M.DOUBLE_SIDE)
    assert np.array_equal(face.cpu().numpy().view(np.uint32), host["face"])
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), host["distance"].view(np.uint32))


def test_edge_straddling_all_hits_equal_the_oracle(ctx, knot):
    g, idx, tree, mesh, om = knot
    P, N, w, L, e, rng = _straddle(g, idx, 200, 5)
    O32 = _f32_rounded(P + 0.01 * L * N).astype(np.float64)
    sign = np.repeat(np.array([[1.0], [-1.0]]), P.shape[0], axis=0)
    origins = np.tile(O32, (2, 1))
    dirs = _unit(np.tile(P, (2, 1)) + sign * e * np.tile(w, (2, 1)) - origins)
    print(f"case edge-straddling all hits: {origins.shape[0]} rays", flush=True)
    got = mesh.raycast(origins, dirs, side=M.DOUBLE_SIDE)
    off = got["offsets"].astype(np.int64)
    changed = 0
    for i in range(origins.shape[0]):
        rf, rd = om.raycast_all(origins[i], dirs[i], side=O.DOUBLE)
        gf = got["face"][off[i]:off[i + 1]]
        assert np.array_equal(gf, rf), (i, gf, rf)
        assert np.all(np.abs(got["distance"][off[i]:off[i + 1]] - rd) <= 1e-5 * np.maximum(rd, 1))
        r32, _ = om.raycast_all(_f32_rounded(origins[i]), _f32_rounded(dirs[i]), side=O.DOUBLE)
        changed += int(not np.array_equal(r32, rf))
    print("hit lists that float32 rounding would change:", changed, flush=True)
    assert changed > 0.2 * origins.shape[0]


@pytest.mark.parametrize("lanes", [4, 1])
def test_edge_straddling_points_equal_the_oracle_exactly(knot, lanes):
    """Points under shared edges (on the concave side), offset across them by 1e-11 of the mesh size:
    the closest face of the double point is on the offset's side. The edges are creased (faces' normals
    with 1 - n1.n2 > 1e-5, three quarters of the knot's edges): the two faces' squared distances then
    differ by about 4e-11 of their value or more, nine times the rounding-tie band (point_tie_band,
    about 5e-12 here), so the straddled pair is decided strictly. Across the nearly flat diagonals of
    the knot's quads the difference falls to the rounding level, where the FP64 reference's own
    roundings pick the face; the few points whose closest feature is shared by other faces at the same
    distance are rounding ties too."""
    g, idx, tree, _, om = knot
    c = M.Context(0)
    c.set_point_lanes(lanes)
    mesh = c.upload(tree, g.positions, idx)
    try:
        P, N, w, L, e, rng = _straddle(g, idx, 2000, 7, crease=1e-5)
        sign = np.repeat(np.array([[1.0], [-1.0]]), P.shape[0], axis=0)
        # on the concave side of each edge (the two faces' closest points lie inside them)
        pts = np.tile(P - 2e-3 * L * N, (2, 1)) + sign * e * np.tile(w, (2, 1))
        print(f"case edge-straddling points lanes={lanes}: {pts.shape[0]} points", flush=True)
        got = mesh.closest_point(pts)
        ref = om.closest_point(pts)
        diff = np.nonzero(got["face"] != ref["face"])[0]
        bad = [int(i) for i in diff if not point_faces_tied(om, pts[i], got["face"][i], ref["face"][i])]
        print(f"faces differing: {diff.size}, all rounding ties but {len(bad)}", flush=True)
        assert not bad, bad[:5]
        assert diff.size < 0.02 * pts.shape[0]
        assert np.max(np.abs(got["distance"] - ref["distance"]) / np.maximum(ref["distance"], 1)) <= 1e-5
        r32 = om.closest_point(_f32_rounded(pts))
        changed = np.count_nonzero(r32["face"] != ref["face"])
        print("faces that float32 rounding would change:", changed, flush=True)
        assert changed > 0.2 * pts.shape[0]
        # minThreshold / maxThreshold in double: a threshold just above / below each exact distance
        sub = slice(0, 400)
        d0 = ref["distance"][sub]
        for i in range(0, 400, 40):
            for f in (1 - 1e-12, 1 + 1e-12):
                gmax = mesh.closest_point(pts[sub][i:i + 1], 0.0, float(d0[i] * f))
                rmax = om.closest_point(pts[sub][i:i + 1], 0.0, float(d0[i] * f))
                assert gmax["face"][0] == rmax["face"][0], (i, f, gmax["face"][0], rmax["face"][0])
    finally:
        mesh.free()
        c.close()


def test_generic_double_rays_and_points(ctx, knot):
    """C1-style rays and a grid with offsets below float resolution: the parity rule against the
    oracle on the double inputs, and the device form of the closest point equal to the host form."""
    g, idx, tree, mesh, om = knot
    w = W.c1_raycast(20000)
    rng = np.random.default_rng(9)
    o = w.origins.astype(np.float64) + rng.standard_normal(w.origins.shape) * 1e-8
    d = w.directions.astype(np.float64) + rng.standard_normal(w.directions.shape) * 1e-10
    got = mesh.raycast_first(o, d)
    ref = om.raycast_first(o, d)
    print("double rays", check_raycast_parity(om, o, d, got, ref, "f64 rays"), flush=True)
    lo, hi = g.bounds()
    pts = W.grid_points(lo, hi, 24).astype(np.float64) + rng.standard_normal((24 ** 3, 3)) * 1e-8
    gotc = mesh.closest_point(pts)
    refc = om.closest_point(pts)
    print("double points", check_closest_parity(om, pts, gotc, refc, "f64 points"), flush=True)
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    hi_p = _f32_rounded(pts)
    lo_p = (pts - hi_p.astype(np.float64)).astype(np.float32)
    tq, tl = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (hi_p, lo_p))
    n = pts.shape[0]
    face = torch.empty(n, dtype=torch.int32, device=dev)
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    mesh.closest_point_ff_d(tq.data_ptr(), tl.data_ptr(), n, face.data_ptr(), dist.data_ptr())
    ctx.synchronize()
    assert np.array_equal(face.cpu().numpy().view(np.uint32), gotc["face"])
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), gotc["distance"].view(np.uint32))


def test_batched_local_double_rays(ctx, knot):
    """The batched launch on local double rays (BatchedMesh's path): equal to each mesh's own
    double-precision raycastFirst, and per-ray windows decided in double: near just above / below the
    exact hit distance drops / keeps that hit, as the oracle does."""
    g, idx, tree, mesh, om = knot
    g2, idx2, tree2 = build_mesh(G.sphere(7, 64, 32))
    mesh2 = ctx.upload(tree2, g2.positions, idx2)
    om2 = O.OracleMesh(g2.positions, idx2, tree2.roots)
    try:
        P, N, w, L, e, rng = _straddle(g, idx, 500, 11)
        sign = np.repeat(np.array([[1.0], [-1.0]]), P.shape[0], axis=0)
        o1 = np.tile(_f32_rounded(P + 0.01 * L * N).astype(np.float64), (2, 1))
        d1 = _unit(np.tile(P, (2, 1)) + sign * e * np.tile(w, (2, 1)) - o1)
        w2 = W.c1_raycast(1000)
        o2 = w2.origins.astype(np.float64) * 0.5 + 1e-8
        d2 = _unit(-o2 + rng.standard_normal(o2.shape) * 0.1)
        origins, dirs = np.concatenate([o1, o2]), np.concatenate([d1, d2])
        ray_mesh = np.concatenate([np.zeros(o1.shape[0], np.uint32), np.ones(o2.shape[0], np.uint32)])
        ray_mesh[::97] = 7  # no such mesh: a miss
        batch = ctx.batch([mesh, mesh2])
        print(f"case batched local double rays: {origins.shape[0]} rays", flush=True)
        got = batch.raycast_first_local(origins, dirs, ray_mesh, side=M.DOUBLE_SIDE)
        assert np.all(got["face"][ray_mesh == 7] == M.NO_HIT)
        for k, (m, omk) in enumerate(((mesh, om), (mesh2, om2))):
            sel = ray_mesh == k
            one = m.raycast_first(origins[sel], dirs[sel], side=M.DOUBLE_SIDE)
            assert np.array_equal(got["face"][sel], one["face"])
            assert np.array_equal(got["distance"][sel].view(np.uint32), one["distance"].view(np.uint32))
            ref = omk.raycast_first(origins[sel], dirs[sel], side=O.DOUBLE)
            assert np.array_equal(got["face"][sel], ref["face"])
        # windows: near just past / short of the exact distance of the first hit
        sel = np.nonzero((ray_mesh == 0) & (got["face"] != M.NO_HIT))[0][:300]
        ref0 = om.raycast_first(origins[sel], dirs[sel], side=O.DOUBLE)
        for f in (1 + 1e-12, 1 - 1e-12):
            win = np.stack([ref0["distance"] * f, np.full(sel.size, np.inf)], axis=1)
            g2r = batch.raycast_first_local(origins[sel], dirs[sel], ray_mesh[sel], windows=win, side=M.DOUBLE_SIDE)
            for j, i in enumerate(sel[:60]):
                r = om.raycast_first(origins[i:i + 1], dirs[i:i + 1], side=O.DOUBLE, near=float(win[j, 0]))
                assert g2r["face"][j] == r["face"][0], (f, j, g2r["face"][j], r["face"][0])
            if f > 1:
                assert np.all(g2r["face"] != got["face"][sel])  # the first hit is dropped
            else:
                assert np.array_equal(g2r["face"], got["face"][sel])  # and kept
        batch.free()
    finally:
        mesh2.free()
